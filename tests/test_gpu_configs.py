"""The two plain-FP64 workload configurations, run as specified.

configs[0] (uniform ball): 20,000 equal-mass particles uniform in a sphere of
radius 0.5 centred in the unit box (seed 0), non-periodic, Dirichlet
boundary from the analytic uniform-sphere potential on the box faces, leaf
capacity 8, 27 GFD neighbours, one solve with GMRES rtol 1e-10.  Checked
against the oracle (same node set, density, potential and field to stated
tolerances) and against the analytic potential (the bounds of the oracle's
own pin in test_oracle_pins.py).

configs[1] (128^3): Zel'dovich lattice (P(k) ~ k^-2, RMS 0.005, seed 1), EdS,
periodic, 20 leapfrog steps from a = 0.02 to 0.1 (Delta_A = 0.004, the last
step shortened), GMRES rtol 1e-8 with the reference's Jacobi preconditioner.
Step 0 is compared with the oracle at full size; every step is checked
through size-independent properties: neighbour lists (distinct, never the
row, stored distances equal to the recomputed periodic distances),
quadratic reproduction of the Laplacian weights, and both solves meeting
their tolerance on the systems built (residuals recomputed on every row)."""
import os

import numpy as np
import pytest

import apcloud
import pyoracle as po
from apcloud import ic
from helpers import relabel

pytestmark = pytest.mark.gpu

BB = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def test_config0_uniform_ball_exact():
    n = 20000
    x, y, z = ic.uniform_ball(n, 0)
    kw = dict(bbox=BB, depth=20, bc=0, c=0.0, lm=27, leaf_capacity=8, bv_kind=1,
              boundary_at_faces=1, bv_mass=1.0, bv_radius=0.5)
    print("case: configs[0] ball n=%d" % n, flush=True)
    g = apcloud.Context(apcloud.make_params(**kw))
    g.set_particles(x, y, z)
    o = po.Oracle(po.make_params(rtol=1e-10, restart=200, max_it=3000, **kw))
    o.set_particles(x, y, z)
    g.octree_build()
    o.build_tree()
    assert g.select() == o.select()
    gc, oc = g.comp(), o.comp()
    p = relabel(gc, oc)
    M = gc["nc"] + gc["nv"]
    for s in (g, o):
        s.set_constants(1.0 / n, 4 * np.pi, -1.0)
    gi = g.solve(apcloud.ksp(rtol=1e-10, restart=200, max_it=3000))
    oi = o.solve()
    assert gi.reason_rho > 0 and gi.reason_phi > 0
    assert oi.reason_rho > 0 and oi.reason_phi > 0
    gf, of = g.fields(), o.fields()
    pm = p[:M]
    for f, tol in (("rho", 1e-8), ("phi", 1e-7), ("g", 1e-6), ("H", 1e-5)):
        ref = of[f][pm]
        err = np.abs(gf[f] - ref).max() / (np.abs(ref).max() + 1e-300)
        assert err < tol, (f, err)
    r = np.sqrt((gc["x"][:M] - 0.5) ** 2 + (gc["y"][:M] - 0.5) ** 2 + (gc["z"][:M] - 0.5) ** 2)
    exact = np.where(r < 0.5, -(3 * 0.25 - r * r) / (2 * 0.125), -1.0 / np.maximum(r, 1e-300))
    err = np.abs(gf["phi"] - exact) / np.abs(exact).max()
    assert err.max() < 0.1 and np.sqrt(np.mean(err ** 2)) < 0.04, (err.max(), np.sqrt(np.mean(err ** 2)))
    g.close()


def _wrap(d):
    d = np.where(d >= 0.5, d - 1.0, d)
    return np.where(d < -0.5, d + 1.0, d)


def _apply(diag, off, cols, v):
    out = np.empty_like(v)
    for a in range(0, len(v), 1 << 20):
        b = min(len(v), a + (1 << 20))
        out[a:b] = diag[a:b] * v[a:b] + (off[a:b] * v[cols[a:b]]).sum(axis=1)
    return out


def _check_step(g, sysm, nbr, dist, f, cpot, rtol, rng):
    comp = g.comp()
    M = comp["nc"] + comp["nv"]
    x, y, z, h = comp["x"], comp["y"], comp["z"], comp["h"]
    rows = rng.choice(M, size=min(20000, M), replace=False)
    nb, d = nbr[rows], dist[rows]
    assert nb.min() >= 0 and nb.max() < M and not (nb == rows[:, None]).any()
    assert (np.diff(np.sort(nb, axis=1), axis=1) != 0).all()
    ex, ey, ez = (_wrap(c[nb] - c[rows][:, None]) for c in (x, y, z))
    np.testing.assert_allclose(d, np.sqrt(ex * ex + ey * ey + ez * ez), rtol=1e-14, atol=0)
    cols, a1 = sysm["cols"][rows], sysm["a1"][rows]
    hi = h[rows][:, None]
    p1, p2, p3 = (_wrap(c[cols] - c[rows][:, None]) / hi for c in (x, y, z))
    basis = np.stack([p1, p2, p3, p1 * p1 / 2, p2 * p2 / 2, p3 * p3 / 2, p1 * p2, p1 * p3,
                      p2 * p3], axis=2)
    s = 24.0 * a1
    mom = np.einsum("rj,rjk->rk", s, basis)
    scale = np.einsum("rj,rjk->rk", np.abs(s), np.abs(basis)).max(axis=1, keepdims=True)
    err = np.abs(mom - np.array([0, 0, 0, 1, 1, 1, 0, 0, 0.0])) / scale
    assert err.max() < 1e-9, err.max()
    C = sysm["cols"]
    r1 = sysm["rhs1"] - _apply(sysm["d1"], sysm["a1"], C, f["rho"])
    rel1 = np.linalg.norm(r1 / sysm["d1"]) / np.linalg.norm(sysm["rhs1"] / sysm["d1"])
    assert rel1 < 2 * rtol, rel1
    h3 = h[:M] ** 3
    rhs2 = (f["rho"] - (f["rho"] * h3).sum() / h3.sum()) * cpot
    # the pinned row 0 keeps (rho_0 - mean) C_pot as its right-hand side (solver.cpp:244-259)
    r2 = rhs2 - _apply(sysm["d2"], sysm["a2"], C, f["phi"])
    # left Jacobi: the preconditioned residual is what the solve tests
    rel2 = np.linalg.norm(r2 / sysm["d2"]) / np.linalg.norm(rhs2 / sysm["d2"])
    assert rel2 < 2 * rtol, rel2
    return M


def test_config1_128_cubed_twenty_steps():
    x, y, z, vx, vy, vz = ic.zeldovich(128, seed=1, rms=0.005)
    n = len(x)
    rtol = 1e-8
    kw = dict(bbox=BB, depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8)
    opts = apcloud.ksp(rtol=rtol, restart=200, max_it=3000)
    g = apcloud.Context(apcloud.make_params(**kw))
    g.set_particles(x, y, z, vx, vy, vz)
    a0, a1, da = 0.02, 0.1, 0.004
    g.drift(np.sqrt(a0), 0.5 * da)  # the initial half drift (controller.cpp:106-111)
    P0 = g.particles()
    sched, t = [], a0 + 0.5 * da
    while t < a1:
        dt = da if t + da <= a1 else a1 - t
        sched.append((t, dt))
        t += dt
    assert len(sched) == 20
    rng = np.random.default_rng(1)
    for k, (t, dt) in enumerate(sched):
        print("case: configs[1] 128^3 step %d a=%.4f" % (k, t), flush=True)
        cd, cpot, cf = 1.0 / n, 1.5 / t, -1.0
        cv, cxh = np.sqrt(t) / t / t, np.sqrt(t + dt / 2)
        g.octree_build()
        g.select()
        g.set_constants(cd, cpot, cf)
        sysm = g.build_system()
        nbr, dist = g.neighbors()
        info = g.solve(opts)
        assert info.reason_rho > 0 and info.reason_phi > 0, (k, info.its_rho, info.its_phi)
        f = g.fields()
        M = _check_step(g, sysm, nbr, dist, f, cpot, rtol, rng)
        del sysm, nbr, dist
        if k == 0:
            o = po.Oracle(po.make_params(rtol=rtol, restart=200, max_it=3000,
                                         nthreads=min(16, os.cpu_count() or 1), **kw))
            o.set_particles(*P0[:6])
            o.build_tree()
            o.select()
            oc = o.comp()
            assert oc["nc"] + oc["nv"] == M and oc["nb"] == 0
            o.set_constants(cd, cpot, cf)
            oi = o.solve()
            assert oi.reason_rho > 0 and oi.reason_phi > 0
            o.move_velocity(cv, dt)
            o.move_position(cxh, dt)
            Q = o.particles()
        g.interp_kick(cv, dt)
        g.drift(cxh, dt)
        if k == 0:
            P = g.particles()
            acc_g, acc_o = np.stack(P[6:9]), np.stack(Q[6:9])
            rel = np.linalg.norm(acc_g - acc_o) / np.linalg.norm(acc_o)
            assert rel < 1e-6, rel
            dx = max(np.abs(_wrap(P[q] - Q[q])).max() for q in range(3))
            assert dx < 1e-9, dx
            print("configs[1] step 0 vs oracle: rel acc %.3e, max pos %.3e, its %d/%d"
                  % (rel, dx, info.its_rho, info.its_phi), flush=True)
    P = g.particles()
    assert all(np.isfinite(P[q]).all() and P[q].min() >= 0 and P[q].max() < 1 for q in range(3))
    g.close()
