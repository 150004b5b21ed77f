"""The autotuner on the GPU: it tunes a 32^3 step (FP32 multigrid in FP64
GMRES), keeps candidate values, leaves the state where it found it, and the
tuned knobs change the step's results by no more than the solve tolerance
(launch shapes change only the summation order, the cycle shapes the
preconditioner)."""
import numpy as np
import pytest

import apcloud
from apcloud import ic, tune

pytestmark = pytest.mark.gpu

PARAMS = dict(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8)


def test_autotune_a_small_step():
    x, y, z, vx, vy, vz = ic.zeldovich(32, seed=5, rms=0.02)
    n = len(x)
    o1 = apcloud.ksp(rtol=1e-8, restart=200, max_it=1000, precision=1, pc="jacobi")
    o2 = apcloud.ksp(rtol=1e-8, restart=200, max_it=1000, precision=1, pc="amg")
    a, dt = 0.03, 0.004
    args = (1.0 / n, 1.5 / a, -1.0, np.sqrt(a) / a / a, np.sqrt(a + dt / 2), dt, o1, o2)
    print("case: autotune 32^3", flush=True)
    g = apcloud.Context(apcloud.make_params(**PARAMS))
    g.set_particles(x, y, z, vx, vy, vz)
    before = g.particles()
    res = tune.autotune(g, lambda: g.step(*args))
    print(res, flush=True)
    for k, v in res["choices"].items():
        assert v == res["defaults"][k] or v in tune.CANDIDATES.get(k, tune.WALL.get(k, [])), (k, v)
    assert "amg_nopre" in res["trials"] and "amg_wlevels" in res["trials"]
    assert "seg_blocks" not in res["trials"]  # a single box never launches the tile kernels
    assert "search_chunk" in res["trials"] and "amg_g8_rows" in res["trials"]
    after = g.particles()
    assert all(np.array_equal(p, q) for p, q in zip(before[:6], after[:6]))
    g.step(*args)
    tuned = g.particles()
    h = apcloud.Context(apcloud.make_params(**PARAMS))
    h.set_particles(x, y, z, vx, vy, vz)
    h.step(*args)
    ref = h.particles()
    acc_t, acc_r = np.stack(tuned[6:9]), np.stack(ref[6:9])
    rel = np.linalg.norm(acc_t - acc_r) / np.linalg.norm(acc_r)
    assert rel < 1e-6, rel
    g.close()
    h.close()
