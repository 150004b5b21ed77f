"""Packed boxes (apc_create_packed): independent periodic boxes in one
context, checked against one context per box.

Per box, at the first step: the same computational rows in the same order
(positions, cell lengths, particle counts, types; particle ranges shifted by
the box's first particle), Neigh24 and the neighbour lists bit for bit
(indices shifted by the box's first row), and the system rows bit for bit
(the pinned first row of every box included).  Over two leapfrog steps in
each of the four solver forms (Jacobi or multigrid, FP64 or an FP32 inner
solve in FP64 refinement): every box's solve takes the iterations of the
box alone (each box keeps its own Krylov space and convergence test),
accelerations within 1e-6 relative and the positions' difference within
1e-6 of the displacement over the run.  A tree that the key budget's lowered depth would cut is
refused."""
import numpy as np
import pytest

import apcloud
from apcloud import ic

pytestmark = pytest.mark.gpu

PARAMS = dict(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8)


def make_boxes(nbox, edges, seed):
    return [ic.multibox_box(s) for s in ic.multibox_specs(nbox=nbox, seed=seed, edges=edges)]


def packed_for(bx, **kw):
    counts = np.array([len(b[0]) for b in bx], np.int32)
    P = apcloud.PackedContext(apcloud.make_params(**dict(PARAMS, **kw)), counts)
    P.set_particles(*[np.concatenate([b[q] for b in bx]) for q in range(6)])
    P.set_masses(1.0 / counts)
    return P


def _wrap(d):
    d = np.where(d >= 0.5, d - 1.0, d)
    return np.where(d < -0.5, d + 1.0, d)


def eds(a):
    return 1.5 / a, -1.0, np.sqrt(a) / a / a


def test_packed_rows_lists_and_systems_match_each_box_alone():
    # 20 boxes: 5 box bits lower the depth from 20 to 19
    bx = make_boxes(20, (8, 12, 16), 5)
    print("case: packed structure, 20 boxes, sides",
          [round(len(b[0]) ** (1 / 3)) for b in bx], flush=True)
    P = packed_for(bx)
    P.octree_build()
    P.select()
    pc = P.comp()
    rs, ncb = P.rows()
    assert pc["nb"] == 0 and rs[0] == 0 and rs[-1] == pc["nc"] + pc["nv"]
    cpot, cf, _ = eds(0.05)
    P.set_constants(1.0, cpot, cf)
    sp = P.build_system()
    nbr_p, dist_p = P.neighbors()
    for b, box in enumerate(bx):
        g = apcloud.Context(apcloud.make_params(**PARAMS))
        g.set_particles(*box)
        g.octree_build()
        g.select()
        gc = g.comp()
        M = gc["nc"] + gc["nv"]
        lo, hi = rs[b], rs[b + 1]
        assert hi - lo == M and ncb[b] == gc["nc"], (b, hi - lo, M, ncb[b], gc["nc"])
        for k in ("x", "y", "z", "h", "np", "type"):
            assert np.array_equal(pc[k][lo:hi], gc[k][:M]), (b, k)
        assert np.array_equal(pc["first"][lo:hi], gc["first"][:M] + P.pstart[b]), b
        assert np.array_equal(pc["nneigh"][lo:hi], gc["nneigh"][:M]), b
        f0, cnt = pc["fneigh"][lo], int(gc["nneigh"][:M].sum())
        assert np.array_equal(pc["neigh24"][f0:f0 + cnt], gc["neigh24"][:cnt] + lo), b
        g.set_constants(1.0 / len(box[0]), cpot, cf)
        sg = g.build_system()
        nbr, dist = g.neighbors()
        assert np.array_equal(nbr_p[lo:hi], nbr + lo), b
        assert np.array_equal(dist_p[lo:hi], dist), b
        assert np.array_equal(sp["cols"][lo:hi], sg["cols"] + lo), b
        for k in ("a1", "a2", "d1", "d2", "rhs1", "pbnd"):
            assert np.array_equal(sp[k][lo:hi], sg[k]), (b, k)
        assert sg["d2"][0] == 1.0 and sp["d2"][lo] == 1.0  # every box pins its first row
        g.close()
    P.close()


FORMS = {  # the four GMRES forms: Jacobi (left) or multigrid (right), FP64 or FP32 inner
    "jacobi_fp64": dict(restart=200, max_it=3000),
    "jacobi_fp32ir": dict(restart=200, max_it=3000, precision=1),
    "amg_fp64": dict(restart=60, max_it=500, pc="amg"),
    "amg_fp32ir": dict(restart=60, max_it=500, precision=1, pc="amg"),
}


@pytest.mark.parametrize("form", sorted(FORMS))
def test_packed_steps_match_each_box_alone(form):
    bx = make_boxes(6, (12, 16, 24), 7)
    print("case: packed steps %s, sides" % form,
          [round(len(b[0]) ** (1 / 3)) for b in bx], flush=True)
    opts = apcloud.ksp(rtol=1e-8, **FORMS[form])
    P = packed_for(bx)
    singles = []
    for box in bx:
        g = apcloud.Context(apcloud.make_params(**PARAMS))
        g.set_particles(*box)
        singles.append(g)
    t, dt = 0.022, 0.004
    for k in range(2):
        cpot, cf, cv = eds(t)
        cx = np.sqrt(t + dt / 2)
        _, _, info, _ = P.step(1.0, cpot, cf, cv, cx, dt, opts)
        assert info.reason_rho > 0 and info.reason_phi > 0, (k, info.its_rho, info.its_phi)
        its_rho, its_phi, rr, rp, _, _ = P.solve_info()
        assert (rr > 0).all() and (rp > 0).all()
        assert info.its_phi == its_phi.max() and info.its_rho == its_rho.max()
        for b, g in enumerate(singles):
            # every box's solve is its own GMRES: the iterations of the box alone
            # (a reduction order apart).  A box of uniform density (every row
            # a leaf of the same size and count; see below) has a potential
            # right-hand side of rounding noise, exactly zero in one form and
            # not in the other: its potential solve's count is not compared
            g.octree_build()
            g.select()
            gc = g.comp()
            M = gc["nc"] + gc["nv"]
            uniform = np.ptp(gc["np"][:M] / gc["h"][:M] ** 3) == 0
            _, _, gi, _ = g.step(1.0 / len(bx[b][0]), cpot, cf, cv, cx, dt, opts)
            assert gi.reason_rho > 0 and gi.reason_phi > 0
            assert abs(its_rho[b] - gi.its_rho) <= 1, (b, its_rho[b], gi.its_rho)
            assert uniform or abs(its_phi[b] - gi.its_phi) <= 1, \
                (b, its_phi[b], gi.its_phi)
        print("step %d: packed its %d/%d" % (k, info.its_rho, info.its_phi), flush=True)
        t += dt
    Q = P.particles()
    Rs = [g.particles() for g in singles]
    refs = [np.linalg.norm(np.stack(R[6:9])) / np.sqrt(len(R[0])) for R in Rs]
    for b, g in enumerate(singles):
        lo, hi = P.pstart[b], P.pstart[b + 1]
        R = Rs[b]
        acc_p = np.stack([Q[q][lo:hi] for q in (6, 7, 8)])
        acc_s = np.stack(R[6:9])
        # RMS error against the box's RMS field; a box of uniform density (a
        # 16^3 lattice fills every level-3 leaf with exactly 8 particles) has
        # none of its own, and its density solve meets rtol jointly with the
        # other boxes' under a different Krylov polynomial: its field is held
        # to the batch's median field scale instead
        err = np.linalg.norm(acc_p - acc_s) / np.sqrt(hi - lo)
        scale = max(refs[b], np.median(refs))
        assert err <= 1e-6 * scale, (b, err, refs[b], np.median(refs))
        rel = err / scale
        # the positions differ by the kicks' difference times Cx dt: relative
        # to the displacement over the run, the acceleration's tolerance
        dx = np.stack([_wrap(Q[q][lo:hi] - R[q]) for q in range(3)])
        moved = np.stack([_wrap(R[q] - bx[b][q]) for q in range(3)])
        relx = np.linalg.norm(dx) / np.linalg.norm(moved)
        print("box %d: rms acc %.3e, rel acc %.2e, rel displacement %.2e"
              % (b, refs[b], rel, relx), flush=True)
        assert relx < 1e-6, (b, relx)
        g.close()
    P.close()


def test_packed_refuses_a_tree_the_lowered_depth_cuts():
    # depth 21 with 3 boxes (2 box bits) is lowered to 20; ten particles within
    # 1e-7 of one point fill a level-20 cell beyond the leaf capacity, which
    # the caller's depth 21 would open
    bx = make_boxes(3, (8,), 9)
    x, y, z, vx, vy, vz = bx[1]
    k = np.arange(10) * 1e-8
    bx[1] = (np.concatenate([x, 0.3 + k]), np.concatenate([y, 0.3 + k]),
             np.concatenate([z, 0.3 + k]), *[np.concatenate([v, np.zeros(10)]) for v in (vx, vy, vz)])
    print("case: packed cut refusal", flush=True)
    P = packed_for(bx, depth=21)
    P.octree_build()
    with pytest.raises(apcloud.ApcError) as e:
        P.select()
    assert "cut" in str(e.value)
    # the same boxes at depth 20 (no lowering) select
    Q = packed_for(bx, depth=20)
    Q.octree_build()
    Q.select()
    P.close()
    Q.close()
