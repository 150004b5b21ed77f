"""Benchmark: particle-steps/s of the AP-Cloud solve-and-move step on one
MI355X, FP32-inner / FP64-refinement GMRES (the precision form) against the
plain FP64 GPU step, with the force error between the two.

A step is one pass of the reference's time loop body (controller.cpp:119-208
without output): Morton keys + radix sort + selection + Neigh24, neighbour
search, GFD stencils, density and Poisson GMRES solves, gradient fit,
interpolation + kick + drift.  Inputs are resident in HBM before the timed
region.  Prints one JSON line.

Workloads (synthetic, regenerated from their seeds):
  zeldovich256  256^3 Einstein-de Sitter perturbed lattice, P(k)~k^-2, RMS
                displacement 0.04, 64 Plummer spheres (b=0.005, 20% of the
                particles), seed 2; 10 steps a=0.02..0.1 (the precision-form
                configuration; default).
  zeldovich128  128^3 lattice, RMS 0.005, seed 1; 20 steps a=0.02..0.1.
  zeldovich64   64^3 lattice, RMS 0.01, seed 4; 10 steps.
All: leaf capacity 8, 27 GFD neighbours, GMRES restart 200, max 1000
iterations (the reference's own run options, Makefile:30), rtol 1e-8.  The
density solve uses the reference's Jacobi preconditioner (it converges in
~7 steps); the Poisson solve uses --pc (default amg: the device multigrid,
right preconditioning; the reference's users select a multigrid with
-pc_type through KSPSetFromOptions, solver.cpp:202,274).

Schedule: the reference's time loop (controller.cpp:106-208) from
a_start to a_end with step da: a half drift, then steps of da from
a_start + da/2 while t < a_end, the last one shortened to end at a_end (10
steps for zeldovich256: nine of 0.008 and one of 0.004).  Warmup and timed
steps walk this schedule; when it is used up, the state after the initial
half drift is restored from a device-side copy (apc_checkpoint, inside the
timed region) and the schedule starts again, so every timed step is a step
of the configured trajectory with its own dt > 0.  The dt of every timed
step is recorded.

Legs, all from the same state (the initial conditions after the initial half
drift): (1) accuracy: every step of the schedule run in FP64 and in the
measured precision from the same states, relative RMS difference of the
accelerations and largest position difference per step; (2) the reference
configuration (FP64, Jacobi, the same GMRES settings) for one step after one
untimed step of it (first-use allocations): its time, iterations and its
force difference from the converged FP64 step (skipped with
--no-ref-config); then the launch shapes and the multigrid cycle shape are
autotuned (apcloud.tune, on the second schedule step of the measured form;
--no-tune keeps the built-in shapes) and the choices recorded, for both
timed legs; (3) the FP64 step with the same solver settings and shapes
timed over K steps (skipped with --no-fp64); (4) the measured
precision timed over K steps -> "value".  The roofline entry is for the
kernel with the largest total device time (average launch time x launches)
in one extra step of leg (4) run with the library's kernel timers on (the
timed steps run without them); the next kernels by device time are listed
beside it.  (5) the multi-box form (see multibox_leg; --no-multibox skips
it), tuned the same way.
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
METRIC = "particle-steps/s at 256^3: FP32-IR batched vs FP64 GPU at 1e-6 rel force err"


def make_ic(w, n_side=None):
    from apcloud import ic
    n = n_side or w["n_side"]
    x, y, z, vx, vy, vz = ic.zeldovich(n, seed=w["seed"], rms=w["rms"])
    if w["plummer"]:
        x, y, z, vx, vy, vz = ic.plummer_clusters(x, y, z, vx, vy, vz, w["seed"], n_spheres=64,
                                                  scale=0.005, fraction=0.2)
    return x, y, z, vx, vy, vz


def log(msg):
    """Progress of the legs, on stderr (stdout carries only the JSON line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def eds(a):
    """units mode 2 with Omega=1: C_pot, C_field, Cx(a), Cv(a)."""
    cx = 1.0 / np.sqrt(1.0 / a)
    return 1.5 / a, -1.0, cx, cx / a / a


def traffic_from_profiles(kernel, workload, alg_bytes):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/pmc_traffic.json, written from rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy kernel of known size), or None.  Where
    the summary holds the measured traffic over the algorithmic bytes (per
    row, so independent of the step's row count) it is applied to this
    launch's algorithmic bytes `alg_bytes`."""
    path = os.path.join(REPO, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        e = d["kernels"][kernel]
        if d.get("workload") != workload:
            return None
        if "traffic_over_algorithmic" in e:
            return round(float(e["traffic_over_algorithmic"]) * alg_bytes)
        return float(e["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def issue_from_profiles(kernel, workload):
    """Vector / scalar instruction-issue utilisation of `kernel` from the
    committed counter summary (profiles/pmc_issue.json: rocprofv3 SQ_INSTS_*
    and GRBM_GUI_ACTIVE), or None.  For a kernel bound by instruction issue
    rather than HBM this is its roofline position."""
    try:
        with open(os.path.join(REPO, "profiles", "pmc_issue.json")) as f:
            d = json.load(f)
        if d.get("workload") != workload:
            return None
        e = d["kernels"][kernel]
        return {"valu_busy": e["valu_busy"], "salu_busy": e["salu_busy"],
                "valu_insts_per_launch": e["valu_insts"],
                "source": "profiles/pmc_issue.json (rocprofv3 SQ_INSTS_VALU x 4 cycles / "
                          "(1024 SIMDs x GRBM_GUI_ACTIVE / 8))"}
    except (OSError, KeyError, ValueError, TypeError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zeldovich256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the timed FP64 leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the force-error leg")
    ap.add_argument("--no-ref-config", action="store_true",
                    help="skip the reference-configuration (Jacobi) leg")
    ap.add_argument("--pc", default="amg", choices=["amg", "jacobi"],
                    help="Poisson preconditioner (the density solve always uses Jacobi)")
    ap.add_argument("--cpu-side", type=int, default=80, help="lattice side of the CPU sample")
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="schedule steps in the force-error leg (default: all)")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--restart", type=int, default=200)
    ap.add_argument("--max-it", type=int, default=1000)
    ap.add_argument("--precision", type=int, default=1,
                    help="measured form: 1 FP32 GMRES + FP64 refinement (default), 0 FP64")
    ap.add_argument("--no-multibox", action="store_true", help="skip the packed multi-box leg")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in launch shapes (no autotuning before the timed legs)")
    ap.add_argument("--multibox-boxes", type=int, default=48)
    ap.add_argument("--multibox-steps", type=int, default=20)
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
    if K < 1 or W < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)
    log("generating %s" % args.workload)
    x, y, z, vx, vy, vz = make_ic(w)
    N = len(x)
    params = apcloud.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27,
                                 leaf_capacity=8)
    ctx = apcloud.Context(params)
    ctx.set_particles(x, y, z, vx, vy, vz)
    a_start, a_end = 0.02, 0.1
    da = (a_end - a_start) / nsteps_total
    _, _, cx0, _ = eds(a_start)
    ctx.drift(cx0, 0.5 * da)  # initial half drift (controller.cpp:106-111)
    t_state0 = a_start + 0.5 * da
    state0 = ctx.particles()[:6]
    ctx.checkpoint(0)  # device-side copy of the state after the half drift
    cd = 1.0 / N
    sched = schedule(a_start, a_end, da)

    def opts(precision, pc):
        rho = apcloud.ksp(rtol=args.rtol, restart=args.restart, max_it=args.max_it,
                          precision=precision, pc="jacobi")
        phi = apcloud.ksp(rtol=args.rtol, restart=args.restart, max_it=args.max_it,
                          precision=precision, pc=pc)
        return rho, phi

    def one_step(j, o):
        """Step j of the schedule (from its own state)."""
        t, dt = sched[j]
        cpot, cf, _, cv = eds(t)
        _, _, cxh, _ = eds(t + dt / 2)
        _, _, info, timing = ctx.step(cd, cpot, cf, cv, cxh, dt, o[0], o[1])
        return info, timing

    def walk(j, o):
        """Schedule step j mod len(sched); restores the initial state on the
        device when the schedule starts again.  Returns (info, timing, dt)."""
        jj = j % len(sched)
        if jj == 0 and j > 0:
            ctx.checkpoint(0, restore=True)
        info, timing = one_step(jj, o)
        return info, timing, sched[jj][1]

    def force_err(got, ref):
        d = np.stack([got[6] - ref[6], got[7] - ref[7], got[8] - ref[8]])
        a = np.stack(ref[6:9])
        dx = np.stack([got[k] - ref[k] for k in range(3)])
        dx -= np.round(dx)  # periodic unit box
        return float(np.sqrt((d ** 2).sum() / (a ** 2).sum())), float(np.abs(dx).max())

    # (1) accuracy: every schedule step of each form from the same state (the
    # FP64 trajectory's state before the step), errors per step
    acc, ref0 = None, None
    if not args.no_accuracy:
        log("accuracy leg")
        ctx.checkpoint(0, restore=True)
        per_step, conv = [], []
        nacc = len(sched) if args.accuracy_steps is None else min(args.accuracy_steps, len(sched))
        for j in range(nacc):
            ctx.checkpoint(1)  # the state before step j
            info64, _ = one_step(j, opts(0, args.pc))
            ref = ctx.particles()
            if j == 0:
                ref0 = ref
            ctx.checkpoint(2)  # the FP64 state after step j
            ctx.checkpoint(1, restore=True)
            info1, _ = one_step(j, opts(args.precision, args.pc))
            e, p = force_err(ctx.particles(), ref)
            per_step.append({"a": round(sched[j][0], 6), "dt": sched[j][1],
                             "rel_rms_force_err": e, "max_abs_pos_err": p})
            conv.append([info64.reason_phi > 0, info1.reason_phi > 0])
            ctx.checkpoint(2, restore=True)  # continue along the FP64 trajectory
        acc = {"rel_rms_force_err": max(r["rel_rms_force_err"] for r in per_step),
               "max_abs_pos_err": max(r["max_abs_pos_err"] for r in per_step),
               "per_step": per_step,
               "vs": "FP64 GPU step from the same state (the FP64 trajectory), same solver settings",
               "poisson_converged": conv,
               "amg_levels": ctx.amg_levels() if args.pc == "amg" else None}

    # (2) the reference configuration: FP64, Jacobi, same GMRES limits; one
    # untimed step first so that its first-use allocations are not timed
    refcfg = None
    if not args.no_ref_config and args.pc != "jacobi":
        log("reference-configuration leg")
        ctx.checkpoint(0, restore=True)
        one_step(0, opts(0, "jacobi"))
        ctx.checkpoint(0, restore=True)
        ctx.sync()
        t0 = time.perf_counter()
        infor, _ = one_step(0, opts(0, "jacobi"))
        ctx.sync()
        wall_r = time.perf_counter() - t0
        converged = infor.reason_rho > 0 and infor.reason_phi > 0
        refcfg = {"value": round(N / wall_r, 1), "ms_per_step": round(1e3 * wall_r, 2),
                  "gmres_its": (infor.its_rho, infor.its_phi),
                  "poisson_converged": infor.reason_phi > 0,
                  "note": "FP64, Jacobi GMRES(%d), max %d iterations; the first schedule step, "
                          "timed after one untimed step%s"
                          % (args.restart, args.max_it,
                             "" if converged else "; NOT converged: this is the time to reach "
                             "max_it, not a time to solution")}
        if ref0 is not None:
            e, p = force_err(ctx.particles(), ref0)
            refcfg["rel_rms_force_err_vs_converged_fp64"] = e

    def timed(precision, profile=False):
        """K schedule steps after W warmup steps, from the state after the
        half drift, kernel timers off; with profile, one more step with the
        library's kernel timers on (HIP events on its stream, launches issued
        directly) for the roofline."""
        log("timed leg, precision %d: %d warmup + %d steps" % (precision, W, K))
        ctx.checkpoint(0, restore=True)
        o = opts(precision, args.pc)
        for j in range(W):
            walk(j, o)
        ctx.sync()
        infos, dts, phases = [], [], np.zeros(3)
        t0 = time.perf_counter()
        for j in range(W, W + K):
            info, timing, dt = walk(j, o)
            infos.append((info.its_rho, info.its_phi))
            dts.append(dt)
            phases += timing
        ctx.sync()
        wall = time.perf_counter() - t0
        if not all(d > 0 for d in dts):
            raise RuntimeError("bench: a timed step had dt <= 0: %r" % dts)
        kt = None
        if profile:
            ctx.set_timing(True)
            ctx.reset_kernel_times()
            walk(W + K, o)
            ctx.sync()
            kt = ctx.kernel_times()
            ctx.set_timing(False)
        return wall, infos, dts, phases, kt, info.nc + info.nv

    # launch and cycle shapes: autotuned on the second schedule step of the
    # measured form (the first solves the freshly placed clusters cold, in
    # about twice the Krylov steps of the others), before any timed leg
    tuned = None
    if not args.no_tune:
        from apcloud import tune
        log("autotuning launch and cycle shapes")
        ctx.checkpoint(0, restore=True)
        o_t = opts(args.precision, args.pc)
        one_step(0, o_t)
        tuned = tune.autotune(ctx, lambda: one_step(1, o_t))
        ctx.checkpoint(0, restore=True)

    fp64 = None
    if not args.no_fp64 and args.precision != 0:
        wall64, infos64, dts64, _, _, _ = timed(0)
        fp64 = {"value": round(N * K / wall64, 1), "ms_per_step": round(1e3 * wall64 / K, 2),
                "gmres_its": infos64, "poisson_pc": args.pc}
    wall, infos, dts, phases, kt, M = timed(args.precision, profile=True)

    # roofline of the dominant kernel: the largest total device time (average
    # launch time x launches) over every timed kernel; each carries its
    # algorithmic bytes per launch.  The multigrid cycle and setup timers
    # enclose other timed launches and are not kernels.
    enclosing = {"amg_cycle", "amg_cycle32", "amg_setup", "stream_copy"}
    ranked = sorted(((v[0] * v[1], k, v) for k, v in kt.items()
                     if v[1] > 0 and k not in enclosing), reverse=True)
    copy_gbs = ctx.stream_copy(2 ** 30, reps=10)

    def roof_entry(name, avg_ms, cnt, b):
        if b <= 0:
            return {"kernel": name, "bound": "latency", "avg_launch_ms": round(avg_ms, 4),
                    "launches": cnt, "total_ms": round(avg_ms * cnt, 3)}
        ach = b / (avg_ms * 1e-3) / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(name, args.workload, b),
                "issue": issue_from_profiles(name, args.workload),
                "bytes_per_launch": int(b), "avg_launch_ms": round(avg_ms, 4), "launches": cnt,
                "total_ms": round(avg_ms * cnt, 3)}

    roof, others = None, []
    for q, (_, name, (avg_ms, cnt, b)) in enumerate(ranked[:6]):
        e = roof_entry(name, avg_ms, cnt, b)
        if q == 0:
            roof = e
            roof["measured_copy_GBs"] = round(copy_gbs, 1)
        else:
            others.append(e)

    value = N * K / wall
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * wall / K, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == 1 else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": N, "rows": M, "lm": params.lm,
                   "leaf_capacity": 8,
                   "gmres": {"rtol": args.rtol, "restart": args.restart, "max_it": args.max_it,
                             "poisson_pc": args.pc, "density_pc": "jacobi",
                             "precision": ("fp32 multigrid in fp64 flexible gmres"
                                           if args.pc == "amg" else
                                           "fp32 inner gmres, fp64 refinement")
                             if args.precision == 1 else "fp64"},
                   "schedule": "a=%g..%g, %d steps (last dt shortened to a_end), restarted "
                               "from the initial state when used up" % (a_start, a_end, len(sched)),
                   "timed_dt": dts,
                   "gmres_its": infos, "phase_s": [round(v, 3) for v in phases]},
        "fp64_gpu": fp64,
        "speedup_vs_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "reference_config_fp64_gpu": refcfg,
        "speedup_vs_reference_config": round(value / refcfg["value"], 3) if refcfg else None,
        "accuracy": acc,
        "roofline": roof,
        "roofline_next": others,
        "autotune": tuned,
        "cpu_baseline": None,
    }
    if args.verbose:
        out["kernels"] = {k: {"avg_ms": round(v[0], 4), "launches": v[1],
                              "GBs": round(v[2] / (v[0] * 1e6), 1) if v[0] > 0 else None}
                          for k, v in kt.items()}
    ctx.close()

    if not args.no_multibox:
        out["multibox"] = multibox_leg(args)
    if not args.no_cpu:
        log("CPU baseline leg")
        out["cpu_baseline"] = cpu_baseline(args, w)
    print(json.dumps(out))


def multibox_leg(args):
    """The multi-box form: --multibox-boxes independent periodic boxes (EdS
    Zel'dovich lattices, sides drawn from {16,24,32,48,64,96}, spectral
    index from [-2.5,-1.5], RMS 0.01 box, seed 3) packed into one context
    (apc_create_packed), so that every kernel of a step runs over all boxes
    in one launch.  --multibox-steps steps of a=0.02..0.1 in the measured
    precision (FP32 multigrid in FP64 GMRES for the Poisson solve) after one
    warmup step, from the state after the half drift; every box's solve meets
    rtol on its own residual.  The first step is also run in FP64 from the
    same state for the force error.  Beside it: the same boxes as separate
    contexts, one after another, for two steps (the form the packing
    replaces)."""
    import apcloud
    from apcloud import ic

    specs = ic.multibox_specs(nbox=args.multibox_boxes, seed=3)
    boxes = [ic.multibox_box(s) for s in specs]
    counts = np.array([len(b[0]) for b in boxes], np.int32)
    N = int(counts.sum())
    log("multibox leg: %d boxes, %d particles" % (len(boxes), N))
    params = apcloud.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27,
                                 leaf_capacity=8)
    P = apcloud.PackedContext(params, counts)
    P.set_particles(*[np.concatenate([b[q] for b in boxes]) for q in range(6)])
    P.set_masses(1.0 / counts)
    nsteps = args.multibox_steps
    a_start, a_end = 0.02, 0.1
    da = (a_end - a_start) / nsteps
    sched = schedule(a_start, a_end, da)
    _, _, cx0, _ = eds(a_start)
    P.drift(cx0, 0.5 * da)
    P.checkpoint(0)

    def opts(precision):
        rho = apcloud.ksp(rtol=args.rtol, restart=args.restart, max_it=args.max_it,
                          precision=precision, pc="jacobi")
        phi = apcloud.ksp(rtol=args.rtol, restart=args.restart, max_it=args.max_it,
                          precision=precision, pc=args.pc)
        return rho, phi

    def step(ctx, j, o, cd=1.0):
        t, dt = sched[j]
        cpot, cf, _, cv = eds(t)
        _, _, cxh, _ = eds(t + dt / 2)
        return ctx.step(cd, cpot, cf, cv, cxh, dt, o[0], o[1])[2]

    # force error of the first step: FP64 against the measured precision
    step(P, 0, opts(0))
    ref = P.particles()
    P.checkpoint(0, restore=True)
    step(P, 0, opts(args.precision))
    got = P.particles()
    d = np.stack([got[k] - ref[k] for k in (6, 7, 8)])
    err = float(np.sqrt((d ** 2).sum() / (np.stack(ref[6:9]) ** 2).sum()))
    P.checkpoint(0, restore=True)
    o = opts(args.precision)
    tuned = None
    if not args.no_tune:
        from apcloud import tune
        log("multibox leg: autotuning launch and cycle shapes")
        step(P, 0, o)
        tuned = tune.autotune(P, lambda: step(P, 1, o))
        P.checkpoint(0, restore=True)
    step(P, 0, o)  # warmup
    P.checkpoint(0, restore=True)
    P.sync()
    its = []
    t0 = time.perf_counter()
    for j in range(len(sched)):
        info = step(P, j, o)
        its.append((info.its_rho, info.its_phi))
        if info.reason_rho <= 0 or info.reason_phi <= 0:
            raise RuntimeError("bench: multibox step %d did not converge: %r" % (j, its[-1]))
    P.sync()
    wall = time.perf_counter() - t0
    rows = int(P.rows()[0][-1])
    its_box = P.solve_info()[1].tolist()  # the last step's Poisson iterations per box
    P.close()
    # the same boxes one context each, stepped one after another
    log("multibox leg: separate contexts")
    ctxs = []
    for b in boxes:
        g = apcloud.Context(params)
        g.set_particles(*b)
        g.drift(cx0, 0.5 * da)
        ctxs.append(g)
    for g, n in zip(ctxs, counts):
        if tuned:  # the packed context's tuned shapes (the tile grid has no use here)
            for k, v in tuned["choices"].items():
                g.set_knob(k, v)
        step(g, 0, o, 1.0 / n)  # warmup (first-use allocations)
    for g in ctxs:
        g.sync()
    t0 = time.perf_counter()
    for j in (1, 2):
        for g, n in zip(ctxs, counts):
            step(g, j, o, 1.0 / n)
    for g in ctxs:
        g.sync()
    wall_sep = (time.perf_counter() - t0) / 2
    for g in ctxs:
        g.close()
    value = N * len(sched) / wall
    return {"value": round(value, 1), "unit": "particle-steps/s",
            "ms_per_step": round(1e3 * wall / len(sched), 2), "boxes": len(boxes),
            "particles": N, "rows": rows, "sides": [s["n_side"] for s in specs],
            "steps": len(sched), "gmres_its": its, "last_step_poisson_its_per_box": its_box,
            "first_step_rel_rms_force_err_vs_fp64": err,
            "autotune": tuned,
            "separate_contexts_ms_per_step": round(1e3 * wall_sep, 2),
            "speedup_vs_separate_contexts": round(1e3 * wall_sep / (1e3 * wall / len(sched)), 3),
            "note": "one packed context (apc_create_packed): each kernel of a step over all "
                    "boxes in one launch; %d steps a=0.02..0.1 after one warmup step, every "
                    "box's solve to rtol %g on its own residual; separate contexts: the same "
                    "boxes one context each, stepped in turn (2 steps)" % (len(sched), args.rtol)}


def schedule(a_start, a_end, da):
    """(t, dt) of every step of the reference's loop after the initial half
    drift (controller.cpp:106-208): t += dt accumulates, the last dt is cut
    to end at a_end."""
    out, t = [], a_start + 0.5 * da
    while t < a_end:
        dt = da
        if t + dt > a_end:
            dt = a_end - t
        out.append((t, dt))
        t += dt
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, w):
    """The oracle (the FP64 CPU restatement, OpenMP where the reference has
    it) timed on this host on a bounded sample: one FP64 step of the same
    workload generator at --cpu-side^3 particles, same solver settings, on
    this process's share of the host's cores and on one thread."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po

    n_side = args.cpu_side
    x, y, z, vx, vy, vz = make_ic(w, n_side)
    N = len(x)
    host = os.cpu_count() or 1
    cores = min(16, host)  # a one-GPU job's share of the host

    def run(threads):
        p = po.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27, units=2,
                           omega=1.0, mass_per_particle=1.0 / N, rtol=args.rtol,
                           restart=args.restart, max_it=args.max_it, nthreads=threads,
                           leaf_capacity=8)
        o = po.Oracle(p)
        o.set_particles(x, y, z, vx, vy, vz)
        a = 0.02 + 0.004
        t0 = time.perf_counter()
        o.build_tree()
        S = o.select()
        t1 = time.perf_counter()
        o.set_constants(1.0 / N, 1.5 / a, -1.0)
        o.solve()
        t2 = time.perf_counter()
        o.move_velocity(1.0, 0.008)
        o.move_position(1.0, 0.008)
        t3 = time.perf_counter()
        return S, t3 - t0, (t1 - t0, t2 - t1, t3 - t2)

    S, wall, ph = run(cores)
    _, wall1, ph1 = run(1)
    return {"value": round(N / wall, 1), "unit": "particle-steps/s", "cores": cores,
            "kind": "port",
            "sample": "1 FP64 step (Jacobi GMRES(%d), max %d iterations: the reference's solver) "
                      "of %d^3 particles from the %s generator (S=%d nodes): %.1f s on %d threads "
                      "(select %.2f, solve %.2f, move %.2f s)"
                      % (args.restart, args.max_it, n_side, args.workload, S, wall, cores, *ph),
            "single_thread": {"value": round(N / wall1, 1), "cores": 1, "wall_s": round(wall1, 2),
                              "phase_s": [round(v, 3) for v in ph1]},
            "host_cpus": host, "cpu_model": cpu_model()}


if __name__ == "__main__":
    main()
