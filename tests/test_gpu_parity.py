"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
inputs.  Bit-exact for keys, octree levels, selection, Neigh24, neighbour
lists and the mover; stated tolerances for the least-squares stencils, the
GMRES solves and the gradient."""
import numpy as np
import pytest

import apcloud
import pyoracle as po
from helpers import cloud, neigh_lists, relabel

pytestmark = pytest.mark.gpu

BB = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def pair(x, y, z, bc=1, depth=20, c=0.05, lm=17, vx=None, vy=None, vz=None, **kw):
    print("case: n=%d bc=%d depth=%d c=%g lm=%d %s" % (len(x), bc, depth, c, lm, kw), flush=True)
    g = apcloud.Context(apcloud.make_params(bbox=BB, depth=depth, bc=bc, c=c, lm=lm, **kw))
    g.set_particles(x, y, z, vx, vy, vz)
    o = po.Oracle(po.make_params(bbox=BB, depth=depth, bc=bc, c=c, lm=lm, **kw))
    o.set_particles(x, y, z, vx, vy, vz)
    return g, o


CASES = [
    ("uniform", 3000, 1, 12, 2.0),
    ("clustered", 6000, 1, 20, 0.05),
    ("clustered", 6000, 0, 20, 0.05),
    ("plummer", 20000, 1, 20, 0.01),
    ("plummer", 20000, 0, 20, 0.02),
    ("lattice", 4096, 1, 10, 0.05),
    ("jitter", 32768, 1, 20, 0.01),
    ("ball", 20000, 0, 20, 0.02),
]


def test_library_loads_and_sees_a_gpu():
    assert apcloud.device_count() >= 1
    assert b"gfx950" in apcloud.lib().apc_version()


@pytest.mark.parametrize("kind,n,bc,depth,c", CASES)
def test_keys_and_sort(kind, n, bc, depth, c):
    x, y, z = cloud(kind, n, 11)
    g, o = pair(x, y, z, bc=bc, depth=depth, c=c)
    g.octree_build()
    o.build_tree()
    gk, gi = g.sorted()
    ok, oi = o.tree_sorted()
    assert np.array_equal(gk, ok)
    # first sort from caller order: ties keep caller order, as the oracle's (key, index)
    assert np.array_equal(gi, oi)


def test_keys_on_bisection_planes():
    # coordinates exactly on the FP64 midpoints and the box faces
    m = 16
    g1 = np.arange(m + 1) / m
    X, Y, Z = np.meshgrid(g1, g1, g1, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    g, o = pair(x, y, z, depth=8)
    g.octree_build()
    o.build_tree()
    assert np.array_equal(g.sorted()[0], o.tree_sorted()[0])


@pytest.mark.parametrize("kind,n,bc,depth,c", CASES[:4])
def test_octree_levels(kind, n, bc, depth, c):
    x, y, z = cloud(kind, n, 12)
    g, o = pair(x, y, z, bc=bc, depth=depth, c=c)
    g.octree_build()
    o.build_tree()
    t = o.tree_nodes()
    base = np.concatenate([[0], np.cumsum(t["numnodes"])])
    for lv in range(depth + 1):
        k, f, npp = g.octree_level(lv)
        a, b = base[lv], base[lv + 1]
        assert len(k) == t["numnodes"][lv]
        assert np.array_equal(k, t["key"][a:b])
        assert np.array_equal(f, t["first"][a:b])
        assert np.array_equal(npp, t["count"][a:b])


def _selected(g, o, c):
    g.octree_build()
    o.build_tree()
    Sg = g.select(c)
    So = o.select(c)
    assert Sg == So
    gc, oc = g.comp(), o.comp()
    assert (gc["nc"], gc["nv"], gc["nb"]) == (oc["nc"], oc["nv"], oc["nb"])
    return gc, oc


@pytest.mark.parametrize("kind,n,bc,depth,c", CASES)
def test_selection_and_neigh24(kind, n, bc, depth, c):
    x, y, z = cloud(kind, n, 13)
    g, o = pair(x, y, z, bc=bc, depth=depth, c=c)
    gc, oc = _selected(g, o, c)
    nc = oc["nc"]
    # computational leaves: the reference's own order, field by field
    for f in ("x", "y", "z", "h", "np", "type"):
        assert np.array_equal(gc[f][:nc], oc[f][:nc]), f
    m = oc["np"][:nc] > 0
    assert np.array_equal(gc["first"][:nc][m], oc["first"][:nc][m])
    # everything else by relabelling
    p = relabel(gc, oc)
    assert np.array_equal(p[:nc], np.arange(nc))
    for f in ("np", "type", "nneigh"):
        assert np.array_equal(gc[f], oc[f][p]), f
    mm = gc["np"] > 0
    assert np.array_equal(gc["first"][mm], oc["first"][p][mm])
    gl, ol = neigh_lists(gc), neigh_lists(oc)
    for i in range(len(gl)):
        assert np.array_equal(p[gl[i]], ol[p[i]]), "Neigh24 of node %d" % i


@pytest.mark.parametrize("kind,n,bc,depth,c,lm", [
    ("clustered", 6000, 1, 20, 0.05, 17),
    ("plummer", 20000, 1, 20, 0.01, 27),
    ("plummer", 20000, 0, 20, 0.02, 27),
    ("jitter", 32768, 1, 20, 0.01, 27),
    ("lattice", 4096, 1, 10, 0.05, 27),
    ("ball", 20000, 0, 20, 0.02, 17),
])
def test_neighbor_lists(kind, n, bc, depth, c, lm):
    x, y, z = cloud(kind, n, 14)
    g, o = pair(x, y, z, bc=bc, depth=depth, c=c, lm=lm)
    gc, oc = _selected(g, o, c)
    p = relabel(gc, oc)
    gn, gd = g.neighbors()
    on, od = o.neighbors()
    M = gc["nc"] + gc["nv"]
    for i in range(M):
        assert np.array_equal(p[gn[i]], on[p[i]]), "neighbours of node %d" % i
        assert np.array_equal(gd[i], od[p[i]]), "distances of node %d" % i


@pytest.mark.parametrize("kind,n,bc,c,lm", [
    ("clustered", 6000, 1, 0.05, 17),
    ("plummer", 20000, 0, 0.02, 27),
])
def test_neighbor_lists_second_pass(monkeypatch, kind, n, bc, c, lm):
    """A 32-entry first-pass list sends most nodes through the second,
    full-bound pass; the lists must not change."""
    monkeypatch.setenv("APC_NEIGH_FAST_CAP", "32")
    x, y, z = cloud(kind, n, 14)
    g, o = pair(x, y, z, bc=bc, depth=20, c=c, lm=lm)
    gc, oc = _selected(g, o, c)
    p = relabel(gc, oc)
    g.set_timing(True)
    gn, gd = g.neighbors()
    assert g.kernel_times().get("neighbor_search_long", (0, 0, 0))[1] > 0
    on, od = o.neighbors()
    M = gn.shape[0]
    assert np.array_equal(p[gn], on[p[:M]])
    assert np.array_equal(gd, od[p[:M]])


@pytest.mark.parametrize("kind,n,bc,c,lm", [
    ("plummer", 20000, 1, 0.01, 27),
    ("ball", 20000, 0, 0.02, 17),
    ("jitter", 32768, 1, 0.01, 17),
])
def test_stencils_and_system(kind, n, bc, c, lm):
    x, y, z = cloud(kind, n, 15)
    g, o = pair(x, y, z, bc=bc, c=c, lm=lm, bv_kind=1 if bc == 0 else 0, boundary_at_faces=1)
    gc, oc = _selected(g, o, c)
    p = relabel(gc, oc)
    M = gc["nc"] + gc["nv"]
    for s in (g, o):
        s.set_constants(1.0 / n, 1.5 / 0.02, -1.0)
    gs, os_ = g.build_system(), o.build_system()
    pm = p[:M]
    # the least-squares weights: QR (oracle, dgels-like) vs the Cholesky
    # factor of the Gram matrix + semi-normal solve (GPU); agreement to 1e-9
    # of the row scale
    for f in ("a1", "a2"):
        scale = np.abs(os_[f]).max(axis=1, keepdims=True)[pm] + 1e-300
        err = np.abs(gs[f] - os_[f][pm]) / scale
        assert err.max() < 1e-9, (f, err.max())
    for f in ("d1", "d2", "rhs1", "pbnd"):
        ref = os_[f][pm]
        err = np.abs(gs[f] - ref) / (np.abs(ref).max() + 1e-300)
        assert err.max() < 1e-9, (f, err.max())


@pytest.mark.parametrize("kind,n,bc,c,lm", [
    ("plummer", 20000, 1, 0.01, 27),
    ("ball", 20000, 0, 0.02, 27),
])
def test_solve_matches_oracle(kind, n, bc, c, lm):
    x, y, z = cloud(kind, n, 16)
    kw = dict(bv_kind=1, boundary_at_faces=1, bv_mass=1.0, bv_radius=0.5) if bc == 0 else {}
    g, o = pair(x, y, z, bc=bc, c=c, lm=lm, rtol=1e-12, restart=60, max_it=2000, **kw)
    gc, oc = _selected(g, o, c)
    p = relabel(gc, oc)
    M = gc["nc"] + gc["nv"]
    cpot = 4 * np.pi if bc == 0 else 1.5 / 0.02
    for s in (g, o):
        s.set_constants(1.0 / n, cpot, -1.0)
    gi = g.solve(apcloud.ksp(rtol=1e-12, restart=60, max_it=2000))
    oi = o.solve()
    assert gi.reason_rho > 0 and gi.reason_phi > 0
    assert oi.reason_rho > 0 and oi.reason_phi > 0
    gf, of = g.fields(), o.fields()
    pm = p[:M]
    for f, tol in (("rho", 1e-9), ("phi", 1e-7), ("g", 1e-6), ("H", 1e-5)):
        ref = of[f][pm]
        err = np.abs(gf[f] - ref).max() / (np.abs(ref).max() + 1e-300)
        assert err < tol, (f, err)


@pytest.mark.parametrize("kind,n,bc", [("plummer", 20000, 1), ("ball", 20000, 0)])
def test_mover_bit_exact(kind, n, bc):
    x, y, z = cloud(kind, n, 17)
    rng = np.random.default_rng(5)
    vx, vy, vz = rng.standard_normal((3, n)) * 0.1
    g, o = pair(x, y, z, bc=bc, c=0.02, vx=vx, vy=vy, vz=vz)
    gc, oc = _selected(g, o, 0.02)
    p = relabel(gc, oc)
    M = gc["nc"] + gc["nv"]
    G = rng.standard_normal((M, 3))
    H = rng.standard_normal((M, 6))
    g.set_gradients(G, H)
    Go = np.empty_like(G)
    Ho = np.empty_like(H)
    Go[p[:M]] = G
    Ho[p[:M]] = H
    o.set_gradients(Go, Ho)
    dv_g = g.interp_kick(0.7, 0.013)
    dx_g = g.drift(1.3, 0.013)
    dv_o = o.move_velocity(0.7, 0.013)
    dx_o = o.move_position(1.3, 0.013)
    assert dv_g == dv_o and dx_g == dx_o
    for a, b in zip(g.particles(), o.particles()):
        assert np.array_equal(a, b)


def test_full_step_against_oracle():
    from apcloud import ic
    x, y, z, vx, vy, vz = ic.zeldovich(32, seed=1, rms=0.01)
    n = len(x)
    g, o = pair(x, y, z, c=0.01, lm=27, vx=vx, vy=vy, vz=vz, rtol=1e-12, restart=60, max_it=3000)
    a = 0.02
    cd, cp, cf, cx, cv = 1.0 / n, 1.5 / a, -1.0, np.sqrt(a), np.sqrt(a) / a / a
    o.build_tree()
    o.select()
    o.set_constants(cd, cp, cf)
    o.solve()
    dv_o = o.move_velocity(cv, 0.004)
    dx_o = o.move_position(cx, 0.004)
    dv_g, dx_g, info, _ = g.step(cd, cp, cf, cv, cx, 0.004, apcloud.ksp(rtol=1e-12, restart=60, max_it=3000))
    assert info.reason_rho > 0 and info.reason_phi > 0
    P_g, P_o = g.particles(), o.particles()
    acc_g = np.stack(P_g[6:9])
    acc_o = np.stack(P_o[6:9])
    rel = np.linalg.norm(acc_g - acc_o) / np.linalg.norm(acc_o)
    assert rel < 1e-6, rel
    for a_, b_ in zip(P_g[:3], P_o[:3]):
        assert np.abs(a_ - b_).max() < 1e-12
    assert abs(dv_g - dv_o) <= 1e-6 * dv_o


@pytest.mark.parametrize("kind,n,bc", [("plummer", 20000, 1), ("ball", 20000, 0)])
def test_mixed_precision_gmres_matches_fp64(kind, n, bc):
    # FP32 cycles inside FP64 iterative refinement converge to the FP64 answer
    x, y, z = cloud(kind, n, 18)
    kw = dict(bv_kind=1, boundary_at_faces=1) if bc == 0 else {}
    res = []
    for prec in (0, 1):
        g, _ = pair(x, y, z, bc=bc, c=0.01, lm=27, **kw)
        g.octree_build()
        g.select()
        g.set_constants(1.0 / n, 4 * np.pi if bc == 0 else 75.0, -1.0)
        info = g.solve(apcloud.ksp(rtol=1e-11, restart=60, max_it=4000, precision=prec))
        assert info.reason_rho > 0 and info.reason_phi > 0, (prec, info.reason_rho, info.reason_phi)
        res.append(g.fields())
        if prec == 1:
            assert info.refine_phi >= 2
    for f, tol in (("rho", 1e-9), ("phi", 1e-8), ("g", 1e-7)):
        ref = res[0][f]
        err = np.abs(res[1][f] - ref).max() / np.abs(ref).max()
        assert err < tol, (f, err)
