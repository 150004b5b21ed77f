"""The 256^3 workload the bench measures, checked through properties that do
not depend on size.  The workload is 16.8M particles: a Zel'dovich lattice
with 64 Plummer spheres, periodic, lm = 27, about 7.5M computational nodes.
Rows are sampled at random.

* Neighbour lists (solver.cpp:456-570): lm distinct nodes per row, never the
  row itself.  Every stored distance equals the periodic distance recomputed
  from the node positions.
* Laplacian weights (solver.cpp:24-180): the weights s_j of a row are
  w_j^2 b_j . y with (B^T W^2 B) y = e3 + e4 + e5, so sum_j s_j b_j =
  e3 + e4 + e5 exactly (up to rounding).  That is: the stencil reproduces
  every quadratic, with zero first and mixed moments and unit pure second
  moments.  A wrong index, sign, scale or basis entry breaks it.
* The density and Poisson rows carry the same weights (s/24 and s/h^2).
"""
import numpy as np
import pytest

import apcloud
from apcloud import ic

pytestmark = pytest.mark.gpu

SAMPLE = 100_000


def _wrap(d):
    d = np.where(d >= 0.5, d - 1.0, d)
    return np.where(d < -0.5, d + 1.0, d)


@pytest.fixture(scope="module")
def full():
    x, y, z, vx, vy, vz = ic.zeldovich(256, seed=2, rms=0.04)
    x, y, z, vx, vy, vz = ic.plummer_clusters(x, y, z, vx, vy, vz, 2, n_spheres=64,
                                              scale=0.005, fraction=0.2)
    print("particles", len(x), flush=True)
    params = apcloud.make_params(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27,
                                 leaf_capacity=8)
    g = apcloud.Context(params)
    g.set_particles(x, y, z, vx, vy, vz)
    g.octree_build()
    g.select()
    g.set_constants(1.0 / len(x), 1.5 / 0.02, -1.0)
    sysm = g.build_system()
    nbr, dist = g.neighbors()
    comp = g.comp()
    M = comp["nc"] + comp["nv"]
    print("rows", M, flush=True)
    rows = np.random.default_rng(0).choice(M, size=min(SAMPLE, M), replace=False)
    out = dict(M=M, rows=rows, nbr=nbr[rows], dist=dist[rows], cols=sysm["cols"][rows],
               a1=sysm["a1"][rows], a2=sysm["a2"][rows], x=comp["x"], y=comp["y"],
               z=comp["z"], h=comp["h"])
    del sysm, nbr, dist
    g.close()
    return out


def test_full_size_neighbour_lists(full):
    rows, nb, d = full["rows"], full["nbr"], full["dist"]
    assert full["M"] > 7_000_000
    assert nb.min() >= 0 and nb.max() < full["M"]  # periodic: every neighbour is interior
    assert not (nb == rows[:, None]).any()
    s = np.sort(nb, axis=1)
    assert (np.diff(s, axis=1) != 0).all()
    x, y, z = full["x"], full["y"], full["z"]
    ex = _wrap(x[nb] - x[rows][:, None])
    ey = _wrap(y[nb] - y[rows][:, None])
    ez = _wrap(z[nb] - z[rows][:, None])
    dd = np.sqrt(ex * ex + ey * ey + ez * ez)
    np.testing.assert_allclose(d, dd, rtol=1e-14, atol=0)


def test_full_size_stencils_reproduce_quadratics(full):
    rows, cols = full["rows"], full["cols"]
    x, y, z, h = full["x"], full["y"], full["z"], full["h"]
    hi = h[rows][:, None]
    p1 = _wrap(x[cols] - x[rows][:, None]) / hi
    p2 = _wrap(y[cols] - y[rows][:, None]) / hi
    p3 = _wrap(z[cols] - z[rows][:, None]) / hi
    basis = np.stack([p1, p2, p3, p1 * p1 / 2, p2 * p2 / 2, p3 * p3 / 2, p1 * p2, p1 * p3,
                      p2 * p3], axis=2)
    s = 24.0 * full["a1"]  # the weights s_j themselves
    mom = np.einsum("rj,rjk->rk", s, basis)
    scale = np.einsum("rj,rjk->rk", np.abs(s), np.abs(basis)).max(axis=1, keepdims=True)
    want = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=float)
    err = np.abs(mom - want) / scale
    assert err.max() < 1e-9, err.max()
    # the Poisson row carries s / h^2 (zero on the pinned row)
    a2 = full["a2"]
    live = a2.any(axis=1)
    assert live.sum() >= len(rows) - 1
    np.testing.assert_allclose(a2[live] * hi[live] ** 2, s[live], rtol=1e-13, atol=0)
