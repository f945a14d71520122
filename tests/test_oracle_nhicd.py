"""The oracle's bright-field NHICD schedule (BrightField/BFReconstructionEngine.cpp:595-750,
BrightField/BFReconstructionEngine_Extra.cpp:476-1273), pinned from outside the restated
loop:

* the homogeneous calls (even calls) take consecutive twentieths of one random list of all
  lines, the last twentieth longer, so every 20 of them cover each line exactly once;
* a non-homogeneous call's lines all lie within two lines (z and x) of a line updated
  before (the 5 x 5 filter's support), and the list holds the n/20 - 1 largest
  filtered magnitudes when they are distinct (the order statistic of a full sort);
* the stop test fires only on calls that are multiples of 20;
* with one outer iteration (threshold 5e10: no rejections, a convex problem) NHICD only
  reorders updates and must reach the minimum that full random-order sweeps reach.
"""
import numpy as np
import pytest

import oracle_py as O
from problems import bf_problem


def bf_params(**kw):
    base = dict(bf_mode=1, mrf_c=0.01, sigma_x=0.01, mrf_q=1.2, bragg_T=3.0, bragg_delta=0.5,
                zero_skipping=0)
    base.update(kw)
    return O.params(**base)


def _run(nhicd, calls, stop=0.0, seed=12345, **kw):
    g, ang, vol, counts = bf_problem(bragg_frac=0.03)
    offs = np.full(len(ang), -np.log(2000.0))
    eng = O.Engine(g, ang, counts, bf_params(bf_nhicd=nhicd, num_outer=1, num_inner=calls,
                                             stop_threshold=stop, seed=seed, **kw),
                   offsets=offs)
    costs, n = eng.run()
    return g, eng, costs, n


def test_homogeneous_calls_cover_every_line_once_per_twenty():
    g, eng, costs, n = _run(1, 80)
    L = g.N_z * g.N_x
    lists = eng.nhicd_lists()
    assert n == 80 and len(lists) == 80
    hom = lists[0::2]
    for c in range(19):
        assert len(hom[c]) == L // 20
    assert len(hom[19]) == L - 19 * (L // 20)
    first = np.concatenate(hom[:20])
    np.testing.assert_array_equal(np.sort(first), np.arange(L))
    # the next twenty homogeneous calls walk the same list again
    for a, b in zip(hom[:20], hom[20:40]):
        np.testing.assert_array_equal(a, b)


def test_non_homogeneous_lists_follow_the_filtered_magnitudes():
    g, eng, costs, n = _run(1, 6)
    lists = eng.nhicd_lists()
    L = g.N_z * g.N_x
    # call 1: the filtered magnitudes are non-zero only near the lines of call 0
    visited = np.zeros((g.N_z, g.N_x), bool)
    visited.ravel()[lists[0]] = True
    near = np.zeros_like(visited)
    for dz in range(-2, 3):
        for dx in range(-2, 3):
            near |= np.roll(np.roll(visited, dz, 0), dx, 1)
    z, x = np.divmod(lists[1], g.N_x)
    assert np.all(near[z, x])
    # at most n / 20 - 1 lines lie above the (n - n / 20)-th order statistic
    for c in (1, 3, 5):
        assert 0 < len(lists[c]) <= L // 20 - 1
        assert len(np.unique(lists[c])) == len(lists[c])


def test_filter_and_threshold_helpers():
    from scipy.signal import convolve2d
    rng = np.random.default_rng(3)
    mag = rng.exponential(1.0, (11, 14))
    filt = np.zeros_like(mag)
    O.lib().orc_nhicd_filter(O.dp(np.ascontiguousarray(mag)), 11, 14, O.dp(filt))
    from test_nhicd import WINDOW  # the reference's 4-digit table, itself checked there
    np.testing.assert_allclose(filt, convolve2d(mag, WINDOW, mode="same"), rtol=1e-12)
    v = rng.normal(size=997)
    assert O.lib().orc_nhicd_threshold(O.dp(v), 997) == np.sort(v)[997 - 997 // 20]


def test_stop_test_fires_on_multiples_of_twenty():
    g, eng, costs, n = _run(1, 2000, stop=2e-2)
    assert n < 2000 and (n - 1) % 20 == 0, n


def test_nhicd_reaches_the_full_sweep_minimum():
    # 400 calls = 10 equivalent sweeps of the homogeneous lists plus as many
    # non-homogeneous calls on the lines that change most
    g, eng, costs, n = _run(1, 600)
    c = np.asarray(costs)
    assert np.all(np.diff(c) <= 1e-9 * np.abs(c[1:]))  # no rejections: cost never rises
    _, ref, rc, _ = _run(0, 40)
    assert c[-1] == pytest.approx(rc[-1], rel=1e-4)
    x, xr = eng.x(), ref.x()
    assert np.sqrt(np.mean((x - xr) ** 2)) <= 1e-2 * np.sqrt(np.mean(xr ** 2))
