"""Pins for the oracle's pieces that the reference build cannot run here
(solver.cpp needs PETSc and LAPACKE): the least-squares replacement of
LAPACKE_dgels against scipy's LAPACK dgels, the quicksort against hand
traces, GMRES against a direct solve and PETSc's stopping rules, and the GFD
stencils against what the mathematics fixes (exactness on quadratics, row
sums, uniform density, the uniform-ball potential, leapfrog closed forms,
Einstein-de Sitter coefficients)."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pyoracle as po
from apcloud import ic
from helpers import cloud


# ------------------------------------------------------------- keys ------
def test_key_closed_form_on_dyadic_cells():
    L = 6
    rng = np.random.default_rng(0)
    for _ in range(200):
        ix, iy, iz = rng.integers(0, 2 ** L, 3)
        x, y, z = (ix + 0.5) / 2 ** L, (iy + 0.5) / 2 ** L, (iz + 0.5) / 2 ** L
        want = 0
        for lv in range(L - 1, -1, -1):
            want = (want << 3) | (((ix >> lv) & 1) << 2) | (((iy >> lv) & 1) << 1) | ((iz >> lv) & 1)
        assert po.key64(x, y, z, [0, 1, 0, 1, 0, 1], L) == want


def test_key_midpoint_goes_up_and_nan_is_unshifted():
    assert po.key64(0.5, 0.25, 0.0, [0, 1, 0, 1, 0, 1], 1) == 4
    assert po.key64(0.5, 0.5, 0.5, [0, 1, 0, 1, 0, 1], 2) == 0o70
    assert po.key64(float("nan"), 0.2, 0.2, [0, 1, 0, 1, 0, 1], 5) == 0


# ------------------------------------------------------ least squares ----
@pytest.mark.parametrize("m,nrhs", [(17, 17), (27, 27), (27, 1), (40, 3)])
def test_wlsq_matches_lapack_dgels(m, nrhs):
    rng = np.random.default_rng(m + nrhs)
    B = rng.standard_normal((m, 9))
    b = rng.standard_normal((m, nrhs))
    info, X = po.wlsq(B, b)
    assert info == 0
    _, ref, info2 = lapack.dgels(B, b)
    assert info2 == 0
    np.testing.assert_allclose(X, ref[:9], rtol=1e-11, atol=1e-12)


def test_wlsq_reports_rank_deficiency():
    rng = np.random.default_rng(1)
    B = rng.standard_normal((20, 9))
    B[:, 4] = 0.0
    info, _ = po.wlsq(B, rng.standard_normal((20, 1)))
    assert info == 5


# -------------------------------------------------------- quicksort ------
def test_quicksort_hand_traces():
    # traced by hand through solver.cpp:600-632: ties are not kept stable
    a, b = po.quicksort([2.0, 1.0, 2.0, 1.0], [0, 1, 2, 3])
    assert a.tolist() == [1, 1, 2, 2] and b.tolist() == [1, 3, 2, 0]
    a, b = po.quicksort([3.0, 3.0, 3.0], [0, 1, 2])
    assert b.tolist() == [1, 2, 0]


def test_quicksort_sorts_and_permutes():
    rng = np.random.default_rng(2)
    for n in (1, 2, 5, 40, 300):
        v = rng.integers(0, 6, n).astype(float)
        a, b = po.quicksort(v, np.arange(n))
        assert np.all(np.diff(a) >= 0)
        assert np.array_equal(v[b], a)


# ------------------------------------------------------------ GMRES ------
def _csr(A):
    A = sp.csr_matrix(A)
    return A.indptr, A.indices, A.data


def test_gmres_matches_direct_solve():
    rng = np.random.default_rng(3)
    n = 400
    A = sp.random(n, n, density=0.02, random_state=3) + sp.diags(4.0 + rng.random(n))
    b = rng.standard_normal(n)
    x, its, reason, rn = po.gmres_csr(*_csr(A), b, rtol=1e-12, restart=30)
    assert reason in (2, 3)
    ref = spla.spsolve(sp.csc_matrix(A), b)
    np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-11)


def test_gmres_petsc_stopping_rules():
    n = 50
    A = sp.diags(np.linspace(1, 9, n))
    b = np.ones(n)
    # Jacobi turns a diagonal system into the identity: one iteration
    _, its, reason, _ = po.gmres_csr(*_csr(A), b)
    assert its == 1 and reason == 2
    # max_it caps the count and reports KSP_DIVERGED_ITS
    L = sp.diags([-1, 2.0001, -1], [-1, 0, 1], shape=(n, n))
    _, its, reason, _ = po.gmres_csr(*_csr(L), b, rtol=1e-14, restart=5, max_it=7)
    assert its == 7 and reason == -3
    # zero right-hand side converges at once (absolute)
    _, its, reason, _ = po.gmres_csr(*_csr(A), np.zeros(n))
    assert its == 0 and reason == 3


# ----------------------------------------------- GFD stencils / solve ----
def _oracle(x, y, z, lm=27, bc=1, c=0.01, **kw):
    o = po.Oracle(po.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=bc, c=c, lm=lm, **kw))
    o.set_particles(x, y, z)
    o.build_tree()
    o.select()
    return o


@pytest.mark.parametrize("lm", [17, 27])
def test_stencils_exact_on_quadratics(lm):
    x, y, z = cloud("plummer", 12000, 4)
    o = _oracle(x, y, z, lm=lm, c=0.02)
    o.set_constants(1.0, 1.0, -1.0)
    s = o.build_system()
    cp = o.comp()
    nbr, _ = o.neighbors()
    M = cp["nc"] + cp["nv"]
    X = np.stack([cp["x"], cp["y"], cp["z"]], 1)
    Q = np.array([[1.3, 0.2, -0.4], [0.2, -0.7, 0.5], [-0.4, 0.5, 0.9]])
    lin = np.array([0.3, -1.1, 0.6])

    def u(P, c0):
        d = P - c0
        return 0.5 * np.einsum("...i,ij,...j->...", d, Q, d) + d @ lin

    lap = np.trace(Q)
    ok = 0
    for i in range(M):
        P = X[nbr[i]]
        if np.abs(P - X[i]).max() > 0.4:  # skip rows whose stencil wraps
            continue
        c0 = X[i]
        ui, uj = u(X[i], c0), u(P, c0)
        got2 = s["a2"][i] @ uj + s["d2"][i] * ui
        if i == 0:  # periodic Dirichlet pin (solver.cpp:147-152)
            continue
        assert abs(got2 - lap) < 1e-8 * (1 + abs(lap) / cp["h"][i] ** 0), (i, got2, lap)
        got1 = s["a1"][i] @ uj + s["d1"][i] * ui
        assert abs(got1 - (ui + cp["h"][i] ** 2 / 24 * lap)) < 1e-10
        ok += 1
    assert ok > M // 4


def test_row_sums():
    x, y, z = cloud("plummer", 12000, 5)
    o = _oracle(x, y, z, c=0.02)
    o.set_constants(1.0, 1.0, -1.0)
    s = o.build_system()
    np.testing.assert_allclose(s["a1"].sum(1) + s["d1"], 1.0, atol=1e-12)
    r2 = s["a2"].sum(1) + s["d2"]
    scale = np.abs(s["d2"])
    assert np.all(np.abs(r2[1:]) <= 1e-10 * scale[1:])
    assert s["d2"][0] == 1.0 and np.all(s["a2"][0] == 0)


def test_gradient_exact_on_quadratics():
    x, y, z = cloud("plummer", 12000, 6)
    o = _oracle(x, y, z, c=0.02)
    o.set_constants(1.0, 1.0, -1.0)
    cp = o.comp()
    M = cp["nc"] + cp["nv"]
    X = np.stack([cp["x"], cp["y"], cp["z"]], 1)[:M]
    Q = np.array([[0.8, 0.1, 0.0], [0.1, -0.5, 0.3], [0.0, 0.3, 0.2]])
    lin = np.array([0.5, 0.2, -0.7])
    c0 = np.array([0.5, 0.5, 0.5])
    d = X - c0
    phi = 0.5 * np.einsum("ni,ij,nj->n", d, Q, d) + d @ lin
    o.L.orc_set_potential(o.h, po._p(np.ascontiguousarray(phi)))
    assert o.L.orc_differentiate(o.h) == 0
    f = o.fields()
    nbr, _ = o.neighbors()
    grad = d @ Q + lin
    for i in range(M):
        if np.abs(X[nbr[i]] - X[i]).max() > 0.4:
            continue
        np.testing.assert_allclose(f["g"][i], -grad[i], atol=1e-8)
        H = f["H"][i]
        np.testing.assert_allclose(H, -np.array([Q[0, 0], Q[1, 1], Q[2, 2], Q[0, 1], Q[0, 2], Q[1, 2]]),
                                   atol=1e-6)


def test_uniform_lattice_density_is_exact():
    x, y, z = ic.lattice(16)
    o = _oracle(x, y, z, c=0.05, rtol=1e-12)
    o.set_constants(1.0 / len(x), 10.0, -1.0)
    o.solve()
    f = o.fields()
    np.testing.assert_allclose(f["rho"], 1.0, rtol=1e-12)
    assert np.abs(f["g"]).max() < 1e-10


def test_uniform_ball_potential():
    # the reference's uniform-sphere case: Dirichlet boundary with the analytic
    # potential on the box faces, C_dens = m = 1/N, C_pot = 4 pi G
    n = 20000
    x, y, z = ic.uniform_ball(n, 0)
    o = _oracle(x, y, z, lm=27, bc=0, c=0.0, leaf_capacity=8, bv_kind=1, boundary_at_faces=1,
                bv_mass=1.0, bv_radius=0.5, rtol=1e-10, restart=200, max_it=3000)
    o.set_constants(1.0 / n, 4 * np.pi, -1.0)
    info = o.solve()
    assert info.reason_rho > 0 and info.reason_phi > 0
    cp, f = o.comp(), o.fields()
    M = cp["nc"] + cp["nv"]
    r = np.sqrt((cp["x"][:M] - 0.5) ** 2 + (cp["y"][:M] - 0.5) ** 2 + (cp["z"][:M] - 0.5) ** 2)
    exact = np.where(r < 0.5, -(3 * 0.25 - r * r) / (2 * 0.125), -1.0 / np.maximum(r, 1e-300))
    # 8 particles per cell: Monte-Carlo density noise dominates; a wrong sign,
    # constant or boundary term gives O(1) errors
    err = np.abs(f["phi"] - exact) / np.abs(exact).max()
    assert err.max() < 0.1 and np.sqrt(np.mean(err ** 2)) < 0.04, (err.max(), np.sqrt(np.mean(err ** 2)))


def test_leapfrog_closed_form_constant_acceleration():
    x, y, z = ic.lattice(8)
    n = len(x)
    o = _oracle(x, y, z, bc=1, c=1e9)  # one computational particle: the root
    cp = o.comp()
    M = cp["nc"] + cp["nv"]
    g = np.tile([1e-3, -2e-3, 0.5e-3], (M, 1))
    o.set_gradients(g, np.zeros((M, 6)))
    dt, steps = 0.01, 7
    for _ in range(steps):
        o.move_velocity(1.0, dt)
        o.move_position(1.0, dt)
    P = o.particles()
    k = steps
    for q, p0 in enumerate((x, y, z)):
        want = np.mod(p0 + g[0, q] * dt * dt * k * (k + 1) / 2, 1.0)
        np.testing.assert_allclose(P[q], want, atol=1e-14)


def test_eds_coefficients():
    p = po.make_params(units=2, omega=1.0, omega_lambda=0.0, mass_per_particle=0.25)
    for a in (0.02, 0.05, 0.1):
        cd, cpot, cf, cx, cv = po.coefficients(p, a)
        assert cd == 0.25 and cf == -1.0
        np.testing.assert_allclose([cpot, cx, cv], [1.5 / a, np.sqrt(a), a ** -1.5], rtol=1e-15)
    p0 = po.make_params(units=0)
    assert po.coefficients(p0, 0.3) == [1.0, 1000.0, -1.0, 1.0, 1.0]


def adaptive_lattice_case():
    """An 8^3 lattice at rest, leaf capacity 8: every leaf holds 8 particles
    in a cell of side 1/4, so max n_p/h^3 = 512 on every step (the lattice
    does not move: the density is uniform, so the force is zero).  The
    controller's step (controller.cpp:146-162) is then dt = 0.05 Delta_A /
    512^(1/2) — the reference takes a square root, not a cube root — below
    Delta_A.  From a = 0.02 (t starts at a + Delta_A / 2) to 0.0254 with
    Delta_A = 0.01 that is 18 full steps and a shortened 19th."""
    x, y, z = ic.lattice(8)
    n = len(x)
    da, start, end = 0.01, 0.02, 0.0254
    dt = 0.05 * da / np.sqrt(512.0)
    steps = int(np.ceil((end - (start + 0.5 * da)) / dt))
    return x, y, z, n, da, start, end, dt, steps


def test_adaptive_time_step_closed_form_on_a_lattice():
    x, y, z, n, da, start, end, dt, steps = adaptive_lattice_case()
    assert steps == 19
    # a cube root would give 7 steps, no adaptive step 1: the count pins it
    assert int(np.ceil((end - start - 0.5 * da) / (0.05 * da / 8.0))) == 7
    zero = np.zeros(n)
    o = po.Oracle(po.make_params(depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8, units=2,
                                 omega=1.0, mass_per_particle=1.0 / n, rtol=1e-10, restart=60,
                                 max_it=500, adaptive_dt=1))
    o.set_particles(x, y, z, zero, zero, zero)
    _, done = o.run(start, end, da)
    assert done == steps
    P = o.particles()
    assert max(np.abs(P[k] - p0).max() for k, p0 in enumerate((x, y, z))) < 1e-12
