"""The multigrid preconditioner (pc_type 1): converged solutions equal the
oracle's Jacobi-GMRES solutions, FP32 multigrid inside FP64 flexible GMRES
equals the FP64 form, results are deterministic, and it needs far fewer
Krylov steps than Jacobi on a clustered cloud."""
import numpy as np
import pytest

import apcloud
from helpers import cloud, relabel
from test_gpu_parity import BB, pair, _selected

pytestmark = pytest.mark.gpu


def _solve_pair(kind, n, bc, seed, **kw):
    x, y, z = cloud(kind, n, seed)
    extra = dict(bv_kind=1, boundary_at_faces=1, bv_mass=1.0, bv_radius=0.5) if bc == 0 else {}
    g, o = pair(x, y, z, bc=bc, c=0.01 if bc else 0.02, lm=27, rtol=1e-12, restart=60,
                max_it=2000, **extra)
    gc, oc = _selected(g, o, 0.01 if bc else 0.02)
    cpot = 4 * np.pi if bc == 0 else 1.5 / 0.02
    for s in (g, o):
        s.set_constants(1.0 / n, cpot, -1.0)
    return g, o, gc, oc


@pytest.mark.parametrize("kind,n,bc", [("plummer", 20000, 1), ("ball", 20000, 0)])
@pytest.mark.parametrize("prec", [0, 1])
def test_amg_solution_matches_oracle(kind, n, bc, prec):
    g, o, gc, oc = _solve_pair(kind, n, bc, 21)
    p = relabel(gc, oc)
    M = gc["nc"] + gc["nv"]
    gi = g.solve(apcloud.ksp(rtol=1e-12, restart=60, max_it=2000, precision=prec, pc="amg"))
    oi = o.solve()
    assert gi.reason_rho > 0 and gi.reason_phi > 0, (gi.reason_rho, gi.reason_phi)
    assert oi.reason_rho > 0 and oi.reason_phi > 0
    levels = g.amg_levels()
    assert len(levels) >= 2 and levels[0][0] == M
    assert all(levels[k + 1][0] < levels[k][0] for k in range(len(levels) - 1))
    gf, of = g.fields(), o.fields()
    pm = p[:M]
    for f, tol in (("rho", 1e-9), ("phi", 1e-7), ("g", 1e-6), ("H", 1e-5)):
        ref = of[f][pm]
        err = np.abs(gf[f] - ref).max() / (np.abs(ref).max() + 1e-300)
        assert err < tol, (f, err)


def _spread3(v):
    out = np.zeros(len(v), np.uint64)
    for b in range(21):
        out |= ((v.astype(np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


@pytest.mark.parametrize("kind,bc", [("plummer", 1), ("ball", 0)])
def test_level0_aggregates_are_parent_cells(kind, bc):
    """Level 0's aggregates are the octree's parent cells: two rows share an
    aggregate exactly when their cells have the same parent (cells from the
    rows' centres and widths, recomputed here), and the aggregates are
    numbered in increasing order of the parent's Morton code (corner at 19
    bits per axis) and level."""
    g, _, gc, _ = _solve_pair(kind, 20000, bc, 29)
    info = g.solve(apcloud.ksp(rtol=1e-10, restart=60, max_it=2000, pc="amg"))
    assert info.reason_phi > 0
    _, agg = g.amg_level(0)
    M = gc["nc"] + gc["nv"]
    assert len(agg) == M
    lo = np.array(BB[0::2]), np.array(BB[1::2])
    ext = lo[1] - lo[0]
    L = np.rint(np.log2(ext[0] / gc["h"][:M])).astype(np.int64)
    pc = []
    for a, c in enumerate("xyz"):
        i = np.floor((gc[c][:M] - lo[0][a]) / ext[a] * 2.0 ** L).astype(np.int64)
        pc.append(i >> 1)  # the parent's coordinates at level L - 1
    la = L - 1
    m = sum(_spread3(pc[a] << (19 - la)) << np.uint64(a) for a in range(3))
    key = (m << np.uint64(5)) | la.astype(np.uint64)
    # the same partition: one aggregate per parent key and one key per aggregate
    uk, ki = np.unique(key, return_inverse=True)
    pairs = np.unique(np.stack([ki, agg]), axis=1)
    assert pairs.shape[1] == len(uk) == agg.max() + 1, (pairs.shape, len(uk), agg.max())
    # numbered along the keys
    order = np.empty(len(uk), np.int64)
    order[pairs[1]] = pairs[0]
    assert np.all(np.diff(order) > 0)
    g.close()


def test_amg_fp32_flexible_matches_fp64():
    res = []
    for prec in (0, 1):
        g, _, _, _ = _solve_pair("plummer", 30000, 1, 22)
        info = g.solve(apcloud.ksp(rtol=1e-11, restart=60, max_it=2000, precision=prec, pc="amg"))
        assert info.reason_rho > 0 and info.reason_phi > 0
        res.append((g.fields(), info))
    (f0, i0), (f1, i1) = res
    for f, tol in (("rho", 1e-9), ("phi", 1e-8), ("g", 1e-7)):
        err = np.abs(f1[f] - f0[f]).max() / np.abs(f0[f]).max()
        assert err < tol, (f, err)
    # flexible GMRES keeps the FP64 convergence: no restarts at an FP32 floor
    assert i1.its_phi <= i0.its_phi + 5, (i0.its_phi, i1.its_phi)


def test_amg_deterministic():
    out = []
    for _ in range(2):
        g, _, _, _ = _solve_pair("plummer", 20000, 1, 23)
        g.solve(apcloud.ksp(rtol=1e-10, restart=60, max_it=2000, pc="amg"))
        out.append(g.fields()["phi"])
    assert np.array_equal(out[0], out[1])


def test_amg_needs_fewer_krylov_steps_than_jacobi():
    its = {}
    for pc in ("jacobi", "amg"):
        g, _, _, _ = _solve_pair("plummer", 60000, 1, 24)
        info = g.solve(apcloud.ksp(rtol=1e-10, restart=200, max_it=3000, pc=pc))
        assert info.reason_phi > 0, (pc, info.reason_phi)
        its[pc] = info.its_phi
    assert its["amg"] * 3 <= its["jacobi"], its


def test_amg_rejects_unknown_pc_type():
    g, _, _, _ = _solve_pair("plummer", 5000, 1, 25)
    o = apcloud.ksp(rtol=1e-8)
    o.pc_type = 7
    with pytest.raises(RuntimeError):
        g.solve(o)


@pytest.mark.parametrize("bc", [1, 0])
@pytest.mark.parametrize("sorted_path", [0, 1])
def test_coarse_operators_are_galerkin_products(bc, sorted_path, monkeypatch):
    """Every coarse operator equals P^T A P of the level above it, with P the
    piecewise-constant prolongation of the level's aggregate map (scipy), for
    the coarse-row product and the global-sort product alike."""
    import scipy.sparse as sp

    if sorted_path:
        monkeypatch.setenv("APC_AMG_GALERKIN_SORTED", "1")
    g, _, gc, _ = _solve_pair("plummer" if bc else "ball", 30000, bc, 24)
    info = g.solve(apcloud.ksp(rtol=1e-10, restart=60, max_it=2000, pc="amg"))
    assert info.reason_phi > 0
    sysm = g.build_system()
    M = gc["nc"] + gc["nv"]
    cols, vals = sysm["cols"], sysm["a2"]
    rows = np.repeat(np.arange(M), cols.shape[1])
    m = cols.ravel() < M
    prev = (sp.csr_matrix((vals.ravel()[m], (rows[m], cols.ravel()[m])), shape=(M, M))
            + sp.diags(sysm["d2"])).tocsr()
    levels = g.amg_levels()
    assert len(levels) >= 3, levels
    for l in range(1, len(levels)):
        _, agg = g.amg_level(l - 1)
        Ac, _ = g.amg_level(l)
        n, nc = prev.shape[0], Ac.shape[0]
        assert agg.min() == 0 and agg.max() == nc - 1
        P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
        ref = (P.T @ prev @ P).tocsr()
        err = abs(Ac - ref).max() / abs(ref).max()
        assert err <= 1e-13, (l, err)
        # rows sorted by column, no duplicate columns
        for r in range(0, nc, max(1, nc // 50)):
            cs = Ac.indices[Ac.indptr[r]:Ac.indptr[r + 1]]
            assert np.all(np.diff(cs) > 0)
        prev = Ac


@pytest.mark.parametrize("prec", [0, 1])
def test_ap_cycle_equals_the_plain_cycle(prec, monkeypatch):
    """Level 0's residual updates through A P (r -= alpha (A P) e after each
    coarse correction, and the first post-smoothing step on that residual)
    are the plain cycle in exact arithmetic: after one preconditioned
    Krylov step (x = y0 M^-1 r0) the two potentials agree to rounding, and
    the converged solves agree to their tolerance in as many steps (+-1)."""
    one, full = {}, {}
    rho_opts = apcloud.ksp(rtol=1e-12, restart=60, max_it=2000)
    monkeypatch.setenv("APC_AMG_NOPRE", "0")  # the plain cycle has its pre-smoothing step
    for ap in (0, 1):
        monkeypatch.setenv("APC_AMG_AP", str(ap))
        g, _, _, _ = _solve_pair("plummer", 30000, 1, 26)
        info = g.solve(rho_opts, apcloud.ksp(rtol=1e-14, restart=60, max_it=1, precision=prec,
                                             pc="amg"))
        assert info.its_phi == 1, info.its_phi
        one[ap] = g.fields()["phi"]
        info = g.solve(rho_opts, apcloud.ksp(rtol=1e-11, restart=60, max_it=2000, precision=prec,
                                             pc="amg"))
        assert info.reason_phi > 0, info.reason_phi
        full[ap] = (g.fields()["phi"], info.its_phi)
        g.close()
    err = np.abs(one[1] - one[0]).max() / np.abs(one[0]).max()
    assert err < (1e-12 if prec == 0 else 1e-5), err
    assert np.abs(one[0]).max() > 0
    err = np.abs(full[1][0] - full[0][0]).max() / np.abs(full[0][0]).max()
    assert err < 1e-8, err
    assert abs(full[1][1] - full[0][1]) <= 1, (full[0][1], full[1][1])


@pytest.mark.parametrize("prec", [0, 1])
def test_cycle_without_level0_presmoothing_converges_to_the_same_solution(prec, monkeypatch):
    """Level 0 without its pre-smoothing step (the default) is a different
    preconditioner, not a different answer: the converged potential equals
    the one of the cycle with pre-smoothing to the solves' tolerance, and
    both take a similar number of Krylov steps."""
    out = {}
    rho_opts = apcloud.ksp(rtol=1e-12, restart=60, max_it=2000)
    for nopre in (0, 1):
        monkeypatch.setenv("APC_AMG_NOPRE", str(nopre))
        g, _, _, _ = _solve_pair("plummer", 30000, 1, 27)
        info = g.solve(rho_opts, apcloud.ksp(rtol=1e-11, restart=60, max_it=2000, precision=prec,
                                             pc="amg"))
        assert info.reason_phi > 0, info.reason_phi
        out[nopre] = (g.fields()["phi"], info.its_phi)
        g.close()
    err = np.abs(out[1][0] - out[0][0]).max() / np.abs(out[0][0]).max()
    assert err < 1e-8, err
    assert out[1][1] <= out[0][1] + 6, (out[0][1], out[1][1])
