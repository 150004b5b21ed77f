"""The oracle's restatement of keys, octree, selection (2:1 balance, vacuum
and boundary nodes, Neigh24) and the mover, pinned bit for bit against the
reference's own octree.cpp / mover.cpp: through the committed golden fixtures
(always) and through oracle/_ref/libref.so on fresh inputs (when built)."""
import glob
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import cloud

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_matches_reference_fixture(path):
    F = load(path)
    depth, bc, c = int(F["depth"]), int(F["bc"]), float(F["c"])
    o = po.Oracle(po.make_params(bbox=tuple(F["bbox"]), depth=depth, bc=bc, c=c))
    o.set_particles(F["x"], F["y"], F["z"], F["vx"], F["vy"], F["vz"])
    o.build_tree()
    k, _ = o.tree_sorted()
    assert np.array_equal(k, F["skey"])
    assert np.array_equal(o.tree_nodes()["numnodes"], F["numnodes"])
    assert o.select() == int(F["S"])
    cp = o.comp()
    assert (cp["nc"], cp["nv"], cp["nb"]) == (int(F["nc"]), int(F["nv"]), int(F["nb"]))
    for a, b in (("x", "cx"), ("y", "cy"), ("z", "cz"), ("h", "ch"), ("np", "cnp"),
                 ("type", "ctype"), ("fneigh", "fneigh"), ("nneigh", "nneigh"),
                 ("neigh24", "neigh24")):
        assert np.array_equal(cp[a], F[b]), a
    m = cp["np"] > 0
    assert np.array_equal(cp["first"][m], F["cfirst"][m])
    o.set_gradients(F["g"], F["H"])
    mdv = o.move_velocity(float(F["cv"]), float(F["dt"]))
    mdx = o.move_position(float(F["cxm"]), float(F["dt"]))
    assert mdv == float(F["max_dv"]) and mdx == float(F["max_dx"])
    P = o.particles()
    for a, b in zip(P, ("px", "py", "pz", "pvx", "pvy", "pvz", "pax", "pay", "paz")):
        assert np.array_equal(a, F[b]), b


needs_ref = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built here")


@needs_ref
@pytest.mark.parametrize("kind,n,bc,depth,c", [
    ("uniform", 3000, 1, 10, 2.0),
    ("clustered", 6000, 1, 20, 0.05),
    ("clustered", 6000, 0, 20, 0.05),
    ("plummer", 12000, 1, 20, 0.02),
    ("plummer", 12000, 0, 20, 0.01),
    ("jitter", 4096, 1, 20, 0.005),
])
def test_oracle_matches_reference_build(kind, n, bc, depth, c):
    x, y, z = cloud(kind, n, 21)
    bb = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    rt = po.ref_tree(x, y, z, bb, depth)
    o = po.Oracle(po.make_params(bbox=bb, depth=depth, bc=bc, c=c))
    o.set_particles(x, y, z)
    o.build_tree()
    tn = o.tree_nodes()
    for f in ("numnodes", "key", "depth", "first", "count", "fchild", "nchild", "lo", "hi"):
        assert np.array_equal(tn[f], rt[f]), f
    k, i = o.tree_sorted()
    assert np.array_equal(k, rt["skey"])
    assert sorted(zip(k.tolist(), i.tolist())) == sorted(zip(rt["skey"].tolist(), rt["sidx"].tolist()))
    rm = po.RefMover(x, y, z, x * 0, x * 0, x * 0, bb, depth, bc)
    assert o.select() == rm.select(c)
    oc, rc = o.comp(), rm.comp()
    for f in oc:
        if f == "first":
            m = oc["np"] > 0
            assert np.array_equal(oc[f][m], rc[f][m])
        else:
            assert np.array_equal(np.asarray(oc[f]), np.asarray(rc[f])), f


@needs_ref
def test_reference_nmax_retry_gives_same_result():
    # mover.cpp:84-88 doubles nmax when selection runs out of room; the
    # oracle grows its arrays instead — the outputs must agree
    x, y, z = cloud("plummer", 8000, 3)
    bb = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    rm = po.RefMover(x, y, z, x * 0, x * 0, x * 0, bb, 20, 1)
    S1 = rm.select(0.01, maxnc=64)
    o = po.Oracle(po.make_params(bbox=bb, depth=20, bc=1, c=0.01))
    o.set_particles(x, y, z)
    o.build_tree()
    assert o.select() == S1
    assert np.array_equal(o.comp()["neigh24"], rm.comp()["neigh24"])
