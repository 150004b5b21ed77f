"""The GFD weight function (solver.cpp:77-85: maxdis = 0.7 max d,
w = (exp(-d^2/maxdis^2) - e^-4) / (1 - e^-4)) pinned by closed forms.

A node with the 20 neighbours (+-1,0,0)..., (+-2,0,0)... and (+-1,+-1,+-1)
(h = 1, so the scaled offsets p are these integers) is symmetric under
reflections and axis permutations.  The weighted least-squares fit of
dgels with rows scaled by w (solver.cpp:87-106) minimises sum W_j (B_j c -
f_j)^2 with W = w^2, and its normal matrix splits into blocks:
  gradient terms p_k alone: N_x = sum W p1^2 = 2 W1 + 8 W2 + 8 W3;
  the p_k^2/2 terms: N = a I + b J with a = W1/2 + 8 W2, b = 2 W3;
  the cross terms alone.
(W1, W2, W3: the weights at d = 1, 2, sqrt 3; max d = 2, maxdis = 1.4.)
Hence, by hand:
  Laplacian weights s_j = W_j (|p_j|^2 / 2) / (a + 3 b)   (A2 = s / h^2,
  A1 = s / 24, diagonals -sum s and 1 - sum s / 24, solver.cpp:116-170);
  for phi = x^3: g_x = c_field (2 W1 + 32 W2 + 8 W3) / N_x, g_y = g_z = 0,
  H = 0 (solver.cpp:295-362);
  for phi = x^4: with r = (W1 + 64 W2 + 4 W3, 4 W3, 4 W3), c = (r - b sum(r)
  / (a + 3 b)) / a, H_xx = c_field c_1, H_yy = H_zz = c_field c_2, g = 0.
These hold for any weights; the values pin 0.7, the e^-4 shift and the
squaring, none of which a quadratic-exactness check can see.  The GPU runs
the same node in test_gpu_hand_cases.py."""
import numpy as np
import pytest

import pyoracle as po

GAMMA, SHIFT = 0.7, np.exp(-4.0)


def star():
    """(x, y, z, neigh, cls): node 0 at (8, 8, 8) and its 20 neighbours;
    cls[j] = 1, 2, 3 for the d = 1, d = 2 and corner neighbours.  Every
    neighbour lists only node 0 (its search reaches the others through it)."""
    offs, cls = [], []
    for ax in range(3):
        for sgn in (1, -1):
            for d, c in ((1, 1), (2, 2)):
                o = [0, 0, 0]
                o[ax] = sgn * d
                offs.append(o)
                cls.append(c)
    for sx in (1, -1):
        for sy in (1, -1):
            for sz in (1, -1):
                offs.append([sx, sy, sz])
                cls.append(3)
    p = np.array([[8.0, 8.0, 8.0]] + [[8.0 + a, 8.0 + b, 8.0 + c] for a, b, c in offs])
    neigh = [list(range(1, 21))] + [[0] for _ in range(20)]
    return p[:, 0], p[:, 1], p[:, 2], neigh, np.array([0] + cls)


def closed_forms(gamma=GAMMA, shift=SHIFT, square=True, c_field=-1.0):
    dmax = gamma * 2.0

    def w(d):
        return (np.exp(-d * d / dmax / dmax) - shift) / (1.0 - shift)
    W1, W2, W3 = (w(1.0), w(2.0), w(np.sqrt(3.0)))
    if square:
        W1, W2, W3 = W1 * W1, W2 * W2, W3 * W3
    a, b = W1 / 2 + 8 * W2, 2 * W3
    s = {1: W1 * 0.5 / (a + 3 * b), 2: W2 * 2.0 / (a + 3 * b), 3: W3 * 1.5 / (a + 3 * b)}
    gx = c_field * (2 * W1 + 32 * W2 + 8 * W3) / (2 * W1 + 8 * W2 + 8 * W3)
    r = np.array([W1 + 64 * W2 + 4 * W3, 4 * W3, 4 * W3])
    c = (r - b * r.sum() / (a + 3 * b)) / a
    return dict(s=s, gx3=gx, hxx4=c_field * c[0], hyy4=c_field * c[1])


def oracle_star():
    x, y, z, neigh, cls = star()
    o = po.Oracle(po.make_params(bbox=(0, 16, 0, 16, 0, 16), depth=20, bc=0, c=0.0, lm=20))
    o.set_comp(21, 0, 0, x, y, z, np.ones(21), np.ones(21, np.int32), neigh)
    o.set_constants(1.0, 1.0, -1.0)
    return o, x, cls


def check_row0(system, cls, cf, tol):
    cols, a1, a2 = system["cols"][0], system["a1"][0], system["a2"][0]
    assert sorted(cols.tolist()) == list(range(1, 21))
    s = np.array([cf["s"][cls[j]] for j in cols])
    np.testing.assert_allclose(a2, s, rtol=tol, atol=0)
    np.testing.assert_allclose(a1, s / 24.0, rtol=tol, atol=0)
    np.testing.assert_allclose(system["d2"][0], -s.sum(), rtol=tol)
    np.testing.assert_allclose(system["d1"][0], 1.0 - s.sum() / 24.0, rtol=tol)


def check_fields(f3, f4, cf, tol):
    g3, H3, g4, H4 = f3["g"][0], f3["H"][0], f4["g"][0], f4["H"][0]
    scale = abs(cf["gx3"])
    assert abs(g3[0] - cf["gx3"]) <= tol * scale, (g3[0], cf["gx3"])
    assert np.abs(g3[1:]).max() <= tol * scale and np.abs(H3).max() <= tol * scale
    hs = abs(cf["hxx4"])
    assert abs(H4[0] - cf["hxx4"]) <= tol * hs, (H4[0], cf["hxx4"])
    assert abs(H4[1] - cf["hyy4"]) <= tol * hs and abs(H4[2] - cf["hyy4"]) <= tol * hs
    assert np.abs(H4[3:]).max() <= tol * hs and np.abs(g4).max() <= tol * hs


def test_oracle_laplacian_weights_match_the_closed_form():
    o, _, cls = oracle_star()
    check_row0(o.build_system(), cls, closed_forms(), 1e-12)


def test_oracle_gradient_fit_of_a_cubic_and_a_quartic_matches_the_closed_form():
    o, x, cls = oracle_star()
    o.build_system()
    o.differentiate((x - 8.0) ** 3)
    f3 = o.fields()
    o.differentiate((x - 8.0) ** 4)
    f4 = o.fields()
    check_fields(f3, f4, closed_forms(), 1e-11)


@pytest.mark.parametrize("slip", [dict(gamma=0.75), dict(gamma=0.65), dict(shift=0.0),
                                  dict(square=False)])
def test_a_wrong_weight_function_is_caught(slip):
    """A neighbouring gamma, a dropped e^-4 shift or unsquared weights move the
    closed forms far beyond the tolerances above."""
    good, bad = closed_forms(), closed_forms(**slip)
    rel = max(abs(bad["s"][k] / good["s"][k] - 1) for k in (1, 2, 3))
    rel = max(rel, abs(bad["gx3"] / good["gx3"] - 1), abs(bad["hxx4"] / good["hxx4"] - 1))
    assert rel > 1e-3, (slip, rel)
