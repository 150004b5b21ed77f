"""Kick and drift with more particles than one pass of their bounded grid
covers (2048 blocks of 256 threads = 524,288): every thread moves several
particles and folds their maxima before its block's one atomic.  The returned
maxima are checked bit for bit against the same formulas evaluated in numpy
on the particles the library returns (mover.cpp:51-153: the square root of
max |a|^2 Cv^2 dt^2 over the particles for the kick, of max |v|^2 Cx^2 dt^2
for the drift; the kernels are built without FMA contraction, so the
products round as numpy's do)."""
import numpy as np
import pytest

import apcloud
from helpers import cloud

pytestmark = pytest.mark.gpu

BB = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


@pytest.mark.parametrize("bc", [1, 0])
def test_kick_and_drift_maxima_over_a_strided_grid(bc):
    n = 600_000
    x, y, z = cloud("plummer", n, 21)
    rng = np.random.default_rng(8)
    vx, vy, vz = rng.standard_normal((3, n)) * 0.1
    print("case: n=%d bc=%d" % (n, bc), flush=True)
    g = apcloud.Context(apcloud.make_params(bbox=BB, depth=20, bc=bc, c=0.02, lm=17))
    g.set_particles(x, y, z, vx, vy, vz)
    g.octree_build()
    g.select(0.02)
    comp = g.comp()
    M = comp["nc"] + comp["nv"]
    g.set_gradients(rng.standard_normal((M, 3)), rng.standard_normal((M, 6)))
    cv, cx, dt = 0.7, 1.3, 0.013
    dv = g.interp_kick(cv, dt)
    p = g.particles()
    ax, ay, az = p[6], p[7], p[8]
    assert dv == np.sqrt(((ax * ax + ay * ay + az * az) * cv * cv * dt * dt).max())
    dx = g.drift(cx, dt)
    p = g.particles()
    ux, uy, uz = p[3], p[4], p[5]
    assert dx == np.sqrt(((ux * ux + uy * uy + uz * uz) * cx * cx * dt * dt).max())
    g.close()
