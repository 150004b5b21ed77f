"""The Viewer's ASCII VTK snapshots (viewer.cpp:3-168) byte for byte against
the reference's own viewer.cpp (compiled unmodified into oracle/_ref) on the
same particles: file names from the time (%f), headers, the %f positions and
%e fields, the limit snapshot, and the computational-particle blocks.  This
build's Viewer runs in tests/viewer_harness.cpp, compiled here with g++
(it needs no GPU)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "ap-cloud-cosmology_amd", "host")

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built here")


def build_harness(tmp):
    exe = os.path.join(tmp, "viewer_harness")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(REPO, "include"), "-I", HOST,
                    os.path.join(REPO, "tests", "viewer_harness.cpp"),
                    os.path.join(HOST, "viewer.cpp"), os.path.join(HOST, "parameter.cpp"),
                    "-o", exe], check=True)
    return exe


def data(seed):
    rng = np.random.default_rng(seed)
    n, S = 300, 57
    # magnitudes from 1e-12 to 1e7, signs, exact zeros and a negative zero
    pp = [rng.random(n) for _ in range(3)]
    pp += [rng.standard_normal(n) * 10.0 ** rng.integers(-12, 7, n) for _ in range(6)]
    pp[3][:3] = [0.0, -0.0, 1e-300]
    cp = dict(x=rng.random(S), y=rng.random(S), z=rng.random(S),
              np=rng.integers(0, 9, S), type=rng.choice([0, -1, 1], S),
              first=rng.integers(-1, n, S))
    for k in ("density", "potential", "p_x", "p_y", "p_z"):
        cp[k] = rng.standard_normal(S) * 10.0 ** rng.integers(-8, 8, S)
    return pp, cp


@needs_ref
@pytest.mark.parametrize("seed,t0,t1", [(0, 0.02, 0.0245678), (1, 0.1, 1.0000004)])
def test_viewer_matches_the_reference_viewer(tmp_path, seed, t0, t1):
    exe = build_harness(str(tmp_path))
    pp, cp = data(seed)
    bbox = (0.0, 1.0, -0.5, 0.5, 2.0, 3.25)
    inp = os.path.join(tmp_path, "in.bin")
    with open(inp, "wb") as f:
        np.array([len(pp[0]), len(cp["x"])], np.int32).tofile(f)
        np.array([t0, t1] + list(bbox), np.float64).tofile(f)
        for a in pp:
            np.asarray(a, np.float64).tofile(f)
        for k in ("x", "y", "z"):
            np.asarray(cp[k], np.float64).tofile(f)
        for k in ("np", "type", "first"):
            np.asarray(cp[k], np.int32).tofile(f)
        for k in ("density", "potential", "p_x", "p_y", "p_z"):
            np.asarray(cp[k], np.float64).tofile(f)
    ours, theirs = os.path.join(tmp_path, "ours"), os.path.join(tmp_path, "ref")
    os.makedirs(ours)
    os.makedirs(theirs)
    subprocess.run([exe, inp, ours], check=True)
    assert po.ref_view(theirs, t0, t1, bbox, pp, cp) == 0
    names = sorted(os.listdir(theirs))
    assert names == sorted(os.listdir(ours))
    assert names == sorted(["physical_%f.vtk" % t0, "physical_%f.vtk" % t1,
                            "computational_%f.vtk" % t1])
    for nm in names:
        a = open(os.path.join(ours, nm), "rb").read()
        b = open(os.path.join(theirs, nm), "rb").read()
        assert a == b, nm
