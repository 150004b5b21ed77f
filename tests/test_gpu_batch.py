"""Multi-box form (apc_batch_run): every box of a batch ends where stepping it
alone ends, bit for bit; one batched box against the oracle; per-box errors."""
import numpy as np
import pytest

import apcloud
from apcloud import ic

pytestmark = pytest.mark.gpu

PARAMS = dict(bbox=(0, 1, 0, 1, 0, 1), depth=20, bc=1, c=0.0, lm=27, leaf_capacity=8)


def eds(a):
    cx = np.sqrt(a)
    return 1.5 / a, -1.0, cx, cx / a / a


def schedule(steps, total=20, a0=0.02, a1=0.1):
    """The first `steps` rows (c_pot, c_field, Cv, Cx, dt) of a `total`-step
    run from a0 to a1 after the initial half drift (controller.cpp:119-208)."""
    da = (a1 - a0) / total
    t = a0 + 0.5 * da
    rows = []
    for _ in range(steps):
        dt = min(da, a1 - t) if t + da > a1 else da
        cpot, cf, _, cv = eds(t)
        _, _, cxh, _ = eds(t + dt / 2)
        rows.append((cpot, cf, cv, cxh, dt))
        t += dt
    return np.array(rows)


def test_batch_equals_each_box_alone():
    specs = ic.multibox_specs(nbox=7, seed=3, edges=(16, 24, 32, 48))
    boxes = [ic.multibox_box(s) for s in specs]
    print("sides", [s["n_side"] for s in specs], flush=True)
    params = apcloud.make_params(**PARAMS)
    o = apcloud.ksp(rtol=1e-8, restart=200, max_it=1000)
    coef = schedule(2)
    cd = np.array([1.0 / len(b[0]) for b in boxes])
    B = apcloud.Batch(params, len(boxes), nthreads=4)
    for i, b in enumerate(boxes):
        B.set_particles(i, *b)
    info, status = B.run(coef, cd, o)
    assert np.all(status == 0)
    assert all(inf.reason_rho > 0 and inf.reason_phi > 0 for inf in info)
    for i, b in enumerate(boxes):
        g = apcloud.Context(params)
        g.set_particles(*b)
        for row in coef:
            g.step(cd[i], *row, o)
        want = np.array(g.particles())
        got = np.array(B.boxes[i].particles())
        assert np.array_equal(got, want), (i, np.abs(got - want).max())
        g.close()
    B.close()


def test_batched_box_matches_oracle():
    import pyoracle as po

    spec = dict(n_side=32, spectral_index=-2.0, seed=11, rms=0.02)
    x, y, z, vx, vy, vz = ic.multibox_box(spec)
    n = len(x)
    cpot, cf, cv, cx, dt = schedule(1)[0]
    params = apcloud.make_params(**PARAMS)
    B = apcloud.Batch(params, 2, nthreads=2)
    B.set_particles(0, x, y, z, vx, vy, vz)
    B.set_particles(1, *ic.multibox_box(dict(spec, seed=12)))
    o = apcloud.ksp(rtol=1e-12, restart=60, max_it=4000)
    B.run([(cpot, cf, cv, cx, dt)], [1.0 / n, 1.0 / n], o)
    P = B.boxes[0].particles()
    orc = po.Oracle(po.make_params(rtol=1e-12, restart=60, max_it=4000, **PARAMS))
    orc.set_particles(x, y, z, vx, vy, vz)
    orc.build_tree()
    orc.select()
    orc.set_constants(1.0 / n, cpot, cf)
    orc.solve()
    orc.move_velocity(cv, dt)
    orc.move_position(cx, dt)
    Q = orc.particles()
    acc_g, acc_o = np.stack(P[6:9]), np.stack(Q[6:9])
    assert np.linalg.norm(acc_o) > 1.0
    rel = np.linalg.norm(acc_g - acc_o) / np.linalg.norm(acc_o)
    assert rel < 1e-6, rel
    assert max(np.abs(a - b).max() for a, b in zip(P[:3], Q[:3])) < 1e-11
    B.close()


def test_batch_reports_the_failing_box():
    params = apcloud.make_params(**PARAMS)
    B = apcloud.Batch(params, 3)
    x, y, z, vx, vy, vz = ic.zeldovich(16, seed=1, rms=0.02)
    B.set_particles(0, x, y, z, vx, vy, vz)
    B.set_particles(2, x, y, z, vx, vy, vz)
    with pytest.raises(apcloud.ApcError, match="box 1"):
        B.run(schedule(1), [1.0 / len(x)] * 3, apcloud.ksp(rtol=1e-8))
    B.close()


def test_batch_under_load_is_repeatable():
    """48 boxes (16^3 to 96^3) on 16 host threads, multigrid on the Poisson
    solve: two runs end bit-identical and every solve converges.  Concurrent
    boxes desynchronise the waves of a workgroup, which exposed an intra-block
    race in the coarsest-level Gauss-Jordan pivot scaling (some boxes' solves
    then ran to max_it)."""
    specs = ic.multibox_specs()
    boxes = [ic.multibox_box(s) for s in specs]
    params = apcloud.make_params(**PARAMS)
    o1 = apcloud.ksp(rtol=1e-8, restart=200, max_it=1000)
    o2 = apcloud.ksp(rtol=1e-8, restart=200, max_it=1000, pc="amg")
    coef = schedule(2)
    cd = np.array([1.0 / len(b[0]) for b in boxes])
    runs = []
    for _ in range(2):
        B = apcloud.Batch(params, len(boxes), nthreads=16)
        for i, b in enumerate(boxes):
            B.set_particles(i, *b)
        info, status = B.run(coef, cd, o1, o2)
        assert np.all(status == 0)
        bad = [i for i, x in enumerate(info) if not (x.reason_rho > 0 and x.reason_phi > 0)]
        assert not bad, bad
        runs.append([np.array(B.boxes[i].particles()) for i in range(len(boxes))])
        B.close()
    diff = [i for i in range(len(boxes)) if not np.array_equal(runs[0][i], runs[1][i])]
    assert not diff, diff
