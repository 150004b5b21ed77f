"""The least-squares fits' conditioning fallback (dgels, solver.cpp:106,342):
rows whose condition estimate exceeds the threshold (default 1e5) are fitted
by Householder QR on the GPU.  Checked against the oracle, whose fits are
Householder QR throughout:
* a flattened 20-point node (z offsets scaled by 1e-3, cond(B) ~ 1e6):
  every row is queued and matches the oracle's stencil weights and field
  fits to 1e-9 of the row scale;
* every row sent through QR (threshold 0) on a clustered cloud: the whole
  system and the solution match the oracle at the tolerances of the
  Cholesky path."""
import numpy as np
import pytest

import apcloud
import pyoracle as po
from helpers import cloud, relabel
from test_weight_closed_form import star

pytestmark = pytest.mark.gpu


def flat_star(eps):
    x, y, z, neigh, cls = star()
    z = 8.0 + (z - 8.0) * eps
    return x, y, z, neigh


def row_err(a, b):
    scale = np.abs(b).max(axis=1, keepdims=True) + 1e-300
    return (np.abs(a - b) / scale).max()


def test_ill_conditioned_rows_take_the_qr_fit():
    x, y, z, neigh = flat_star(1e-3)
    kw = dict(bbox=(0, 16, 0, 16, 0, 16), depth=20, bc=0, c=0.0, lm=20)
    print("case: flattened star", flush=True)
    g = apcloud.Context(apcloud.make_params(**kw))
    g.set_comp(21, 0, 0, x, y, z, np.ones(21), np.ones(21, np.int32), neigh)
    o = po.Oracle(po.make_params(**kw))
    o.set_comp(21, 0, 0, x, y, z, np.ones(21), np.ones(21, np.int32), neigh)
    for s in (g, o):
        s.set_constants(1.0, 1.0, -1.0)
    gs, os_ = g.build_system(), o.build_system()
    cond, nq, mx = g.fit_conditions()
    print("condition estimates %.3e..%.3e, %d rows by QR" % (cond.min(), cond.max(), nq), flush=True)
    assert cond.min() > 1e5 and nq == 21
    assert np.array_equal(gs["cols"], os_["cols"])
    assert row_err(gs["a2"], os_["a2"]) < 1e-9
    for f in ("d1", "d2"):
        np.testing.assert_allclose(gs[f], os_[f], rtol=1e-9)
    phi = (x - 8.0) ** 3 + 0.5 * (y - 8.0) ** 2 * (z - 8.0) * 1e3
    g.differentiate(phi)
    o.differentiate(phi)
    gf, of = g.fields(), o.fields()
    # against the field's own scale (row 0 of H is zero up to rounding)
    for f in ("g", "H"):
        err = np.abs(gf[f] - of[f]).max() / np.abs(of[f]).max()
        assert err < 1e-9, (f, err)
    # the same rows through the Cholesky path: reported for comparison
    g.set_fit_threshold(1e300)
    gc = g.build_system()
    print("A2 error vs oracle: QR %.3e, Cholesky %.3e"
          % (row_err(gs["a2"], os_["a2"]), row_err(gc["a2"], os_["a2"])), flush=True)
    g.close()


@pytest.mark.parametrize("kind,n,bc,c,lm", [("plummer", 20000, 1, 0.01, 27),
                                            ("ball", 20000, 0, 0.02, 27)])
def test_every_row_by_qr_matches_the_oracle(kind, n, bc, c, lm):
    x, y, z = cloud(kind, n, 16)
    kw = dict(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=bc, c=c, lm=lm)
    if bc == 0:
        kw.update(bv_kind=1, boundary_at_faces=1, bv_mass=1.0, bv_radius=0.5)
    print("case: all rows by QR, %s n=%d" % (kind, n), flush=True)
    g = apcloud.Context(apcloud.make_params(**kw))
    g.set_particles(x, y, z)
    o = po.Oracle(po.make_params(rtol=1e-12, restart=60, max_it=2000, **kw))
    o.set_particles(x, y, z)
    g.set_fit_threshold(0.0)
    g.octree_build()
    o.build_tree()
    assert g.select() == o.select()
    gc, oc = g.comp(), o.comp()
    p = relabel(gc, oc)
    M = gc["nc"] + gc["nv"]
    cpot = 4 * np.pi if bc == 0 else 1.5 / 0.02
    for s in (g, o):
        s.set_constants(1.0 / n, cpot, -1.0)
    gs, os_ = g.build_system(), o.build_system()
    _, nq, _ = g.fit_conditions()
    assert nq == M
    pm = p[:M]
    assert np.array_equal(p[gs["cols"]], os_["cols"][pm])
    for f in ("a1", "a2"):
        assert row_err(gs[f], os_[f][pm]) < 1e-9, f
    gi = g.solve(apcloud.ksp(rtol=1e-12, restart=60, max_it=2000))
    assert gi.fit_qr_rows == M
    oi = o.solve()
    assert gi.reason_phi > 0 and oi.reason_phi > 0
    gf, of = g.fields(), o.fields()
    for f, tol in (("rho", 1e-9), ("phi", 1e-7), ("g", 1e-6), ("H", 1e-5)):
        ref = of[f][pm]
        err = np.abs(gf[f] - ref).max() / (np.abs(ref).max() + 1e-300)
        assert err < tol, (f, err)
    g.close()
