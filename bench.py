"""Benchmark: particle-steps/s of the AP-Cloud solve-and-move step on one
MI355X.

A step is one pass of the reference's time loop body (controller.cpp:119-208
without output): Morton keys + radix sort + selection + Neigh24, neighbour
search, GFD stencils, density and Poisson GMRES solves, gradient fit,
interpolation + kick + drift.  Inputs are resident in HBM before the timed
region.  Prints one JSON line.

Workloads (synthetic, regenerated from their seeds):
  zeldovich256  256^3 Einstein-de Sitter perturbed lattice, P(k)~k^-2, RMS
                displacement 0.04, 64 Plummer spheres (b=0.005, 20% of the
                particles), seed 2; 10 steps a=0.02..0.1 (the precision-form
                configuration).
  zeldovich128  128^3 lattice, RMS 0.005, seed 1; 20 steps a=0.02..0.1.
Both: leaf capacity 8, 27 GFD neighbours, GMRES restart 200, max 1000
iterations (the reference's own run options, Makefile:30), rtol 1e-8.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ap-cloud-cosmology_amd"))

import numpy as np  # noqa: E402

WORKLOADS = {
    "zeldovich256": dict(n_side=256, rms=0.04, seed=2, plummer=True, steps=10),
    "zeldovich128": dict(n_side=128, rms=0.005, seed=1, plummer=False, steps=20),
    "zeldovich64": dict(n_side=64, rms=0.01, seed=4, plummer=False, steps=10),
}

HBM_PEAK_GBS = 8000.0


def make_ic(w):
    from apcloud import ic
    x, y, z, vx, vy, vz = ic.zeldovich(w["n_side"], seed=w["seed"], rms=w["rms"])
    if w["plummer"]:
        x, y, z, vx, vy, vz = ic.plummer_clusters(x, y, z, vx, vy, vz, w["seed"], n_spheres=64,
                                                  scale=0.005, fraction=0.2)
    return x, y, z, vx, vy, vz


def eds(a):
    """units mode 2 with Omega=1: C_dens, C_pot, C_field, Cx(a), Cv(a)."""
    cx = 1.0 / np.sqrt(1.0 / a)
    return 1.5 / a, -1.0, cx, cx / a / a


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zeldovich256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-side", type=int, default=64, help="lattice side of the CPU sample")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--restart", type=int, default=200)
    ap.add_argument("--max-it", type=int, default=1000)
    ap.add_argument("--precision", type=int, default=0, help="0 FP64 GMRES, 1 FP32 + FP64 refinement")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import apcloud

    w = WORKLOADS[args.workload]
    nsteps_total = w["steps"]
    K = args.steps if args.steps is not None else min(3, nsteps_total)
    W = args.warmup
    x, y, z, vx, vy, vz = make_ic(w)
    N = len(x)
    params = apcloud.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27,
                                 leaf_capacity=8)
    ctx = apcloud.Context(params)
    ctx.set_particles(x, y, z, vx, vy, vz)
    opts = apcloud.ksp(rtol=args.rtol, restart=args.restart, max_it=args.max_it,
                       precision=args.precision)
    a_start, a_end = 0.02, 0.1
    da = (a_end - a_start) / nsteps_total
    t = a_start
    _, _, cx0, _ = eds(t)
    ctx.drift(cx0, 0.5 * da)  # initial half drift (controller.cpp:106-111)
    t += 0.5 * da
    cd = 1.0 / N

    def one_step(t):
        dt = min(da, a_end - t) if t + da > a_end else da
        cpot, cf, _, cv = eds(t)
        _, _, cxh, _ = eds(t + dt / 2)
        mdv, mdx, info, timing = ctx.step(cd, cpot, cf, cv, cxh, dt, opts)
        return t + dt, info, timing

    for _ in range(W):
        t, info, _ = one_step(t)
    ctx.sync()
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    infos = []
    phases = np.zeros(3)
    t0 = time.perf_counter()
    for _ in range(K):
        t, info, timing = one_step(t)
        infos.append((info.its_rho, info.its_phi))
        phases += timing
    ctx.sync()
    wall = time.perf_counter() - t0
    kt = ctx.kernel_times()
    ctx.set_timing(False)
    M = info.nc + info.nv

    # roofline of the dominant kernel: largest total device time among the
    # kernels whose launches carry algorithmic byte counts
    best = None
    for name, (avg_ms, cnt, avg_bytes) in kt.items():
        if cnt == 0 or avg_bytes <= 0 or name == "stream_copy":
            continue
        tot = avg_ms * cnt
        if best is None or tot > best[3]:
            best = (name, avg_bytes, avg_ms, tot, cnt)
    roof = None
    copy_gbs = ctx.stream_copy(2 ** 30, reps=10)
    if best:
        name, b, avg_ms, _, cnt = best
        ach = b / (avg_ms * 1e-3) / 1e9
        roof = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(b), "avg_launch_ms": round(avg_ms, 4), "launches": cnt,
                "measured_copy_GBs": round(copy_gbs, 1)}

    value = N * K / wall
    out = {
        "metric": "particle-steps/s (solve-and-move step, 1 GPU)",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * wall / K, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": N, "rows": M, "lm": params.lm,
                   "leaf_capacity": 8, "gmres": {"rtol": args.rtol, "restart": args.restart,
                                                 "max_it": args.max_it},
                   "gmres_its": infos, "phase_s": [round(v, 3) for v in phases]},
        "roofline": roof,
        "cpu_baseline": None,
    }
    if args.verbose:
        out["kernels"] = {k: {"avg_ms": round(v[0], 4), "launches": v[1],
                              "GBs": round(v[2] / (v[0] * 1e6), 1) if v[0] > 0 else None}
                          for k, v in kt.items()}
    ctx.close()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


def cpu_baseline(args):
    """The oracle (the FP64 CPU restatement, OpenMP where the reference has
    it) timed on this host on a bounded sample: one step of an n^3
    perturbed lattice with the same settings."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po
    from apcloud import ic

    n_side = args.cpu_side
    x, y, z, vx, vy, vz = ic.zeldovich(n_side, seed=1, rms=0.005)
    N = len(x)
    cores = os.cpu_count() or 1
    p = po.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27, units=2, omega=1.0,
                       mass_per_particle=1.0 / N, rtol=args.rtol, restart=args.restart,
                       max_it=args.max_it, nthreads=cores, leaf_capacity=8)
    o = po.Oracle(p)
    o.set_particles(x, y, z, vx, vy, vz)
    t0 = time.perf_counter()
    o.build_tree()
    S = o.select()
    o.set_constants(1.0 / N, 1.5 / 0.022, -1.0)
    o.solve()
    o.move_velocity(1.0, 0.004)
    o.move_position(1.0, 0.004)
    wall = time.perf_counter() - t0
    return {"value": round(N / wall, 1), "unit": "particle-steps/s", "cores": cores,
            "kind": "port", "sample": "1 step, %d^3 perturbed lattice (S=%d nodes)" % (n_side, S)}


if __name__ == "__main__":
    main()
