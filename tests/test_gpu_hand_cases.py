"""The hand-traced neighbour-search cases (tests/golden/hand_cases.json) and
the weight-function closed forms on the GPU, through apc_set_comp (the
Solver over a caller-filled Computational_particle): traced rows and error
codes as traced; every row bit for bit against the oracle on the same
inputs; the symmetric 20-point node's Laplacian weights, gradient and
Hessian fits against their closed forms to 1e-9 (the GPU forms the fits by
Cholesky of the Gram matrix, the oracle by Householder QR)."""
import json
import os

import numpy as np
import pytest

import apcloud
from test_hand_cases import CASES, oracle_for
from test_weight_closed_form import check_fields, check_row0, closed_forms, star

pytestmark = pytest.mark.gpu


def gpu_for(case):
    g = apcloud.Context(apcloud.make_params(bbox=tuple(case["bbox"]), depth=20, bc=case["bc"],
                                            c=0.0, lm=case["lm"]))
    g.set_comp(case["nc"], case["nv"], case["nb"], case["x"], case["y"], case["z"], case["h"],
               case["np"], case["neigh"])
    return g


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_search_matches_the_hand_trace_and_the_oracle(case):
    print("case: hand trace %s" % case["name"], flush=True)
    g = gpu_for(case)
    if case["error"]:
        with pytest.raises(apcloud.ApcError) as e:
            g.neighbors()
        assert case["error"] in str(e.value), str(e.value)
        g.close()
        return
    nbr, dist = g.neighbors()
    g.close()
    for row, want in case["rows"].items():
        r = int(row)
        assert list(nbr[r]) == want["nbr"], (list(nbr[r]), want["nbr"])
        np.testing.assert_array_equal(dist[r], np.sqrt(np.array(want["d2"], float)))
    st, onbr, odist = oracle_for(case).neighbors_status()
    assert st == 0
    assert np.array_equal(nbr, onbr) and np.array_equal(dist, odist)


def test_gpu_weights_match_the_closed_forms():
    x, y, z, neigh, cls = star()
    print("case: 20-point star", flush=True)
    g = apcloud.Context(apcloud.make_params(bbox=(0, 16, 0, 16, 0, 16), depth=20, bc=0, c=0.0,
                                            lm=20))
    g.set_comp(21, 0, 0, x, y, z, np.ones(21), np.ones(21, np.int32), neigh)
    g.set_constants(1.0, 1.0, -1.0)
    cf = closed_forms()
    check_row0(g.build_system(), cls, cf, 1e-9)
    g.differentiate((x - 8.0) ** 3)
    f3 = g.fields()
    g.differentiate((x - 8.0) ** 4)
    f4 = g.fields()
    check_fields(f3, f4, cf, 1e-9)
    g.close()


def test_set_comp_rejects_bad_sets_and_the_mover_refuses():
    case = CASES[0]
    g = gpu_for(case)
    with pytest.raises(apcloud.ApcError):
        g.interp_kick(1.0, 0.01)  # no physical particles belong to these nodes
    bad = dict(case)
    bad["neigh"] = [list(range(1, 13)) * 3] + case["neigh"][1:]  # 36 > 24 entries
    with pytest.raises(apcloud.ApcError):
        gpu_for(bad)
    g.close()
