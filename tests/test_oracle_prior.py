"""The oracle's prior re-estimation (ESTIMATE_PRIOR): estimateSigmaX
(HAADF/HAADF_ReconstructionEngine.cpp:1827-1945) and updatePriorModel
(HAADF/HAADF_QGGMRFPriorModel.cpp:24-31), pinned from outside the restated loops:

* the estimate equals a vectorised numpy sum of the QGGMRF potential over the 13
  forward neighbour offsets (weights by the number of zero offsets: faces, edges,
  corners) for first voxels in the planes [int(0.25 N_z), int(0.75 N_z)), each
  difference scaled by SigmaX, over N_x N_y (zEnd - zStart + 1);
* the new SigmaX is the closed form gamma (estimate q)^(1/q), gamma = 5;
* the engine's prior cost term afterwards is the potential with the new SigmaX (a
  stale SIGMA_X_* constant breaks it);
* the run loop re-estimates once per outer iteration that does not end the run.
"""
import itertools

import numpy as np
import pytest

import oracle_py as O
from problems import small_problem


def _rho(d, sigma, q, c):
    a = np.abs(d)
    return (a ** 2 / sigma ** 2) / (c + a ** (2 - q) / sigma ** (2 - q))


def _forward_offsets():
    """The 13 of the 26 neighbour offsets that come after the voxel in [z][x][y] order."""
    out = []
    for d in itertools.product((-1, 0, 1), repeat=3):
        if d > (0, 0, 0):
            out.append(d)
    assert len(out) == 13
    return out


def _pair_sum(x, z0, z1, scale, sigma, q, c):
    nz, nx, ny = x.shape
    total = 0.0
    for di, dj, dk in _forward_offsets():
        zeros = (di == 0) + (dj == 0) + (dk == 0)
        f = {2: 0.0523, 1: 0.0370, 0: 0.0302}[zeros]
        i = np.arange(z0, z1)
        i = i[(i + di >= 0) & (i + di < nz)]
        j = np.arange(nx)
        j = j[(j + dj >= 0) & (j + dj < nx)]
        k = np.arange(ny)
        k = k[(k + dk >= 0) & (k + dk < ny)]
        a = x[np.ix_(i, j, k)]
        b = x[np.ix_(i + di, j + dj, k + dk)]
        total += f * np.sum(_rho((a - b) * scale, sigma, q, c))
    return total


@pytest.mark.parametrize("q,sigma", [(1.2, 0.1), (2.0, 0.05), (1.5, 0.3)])
def test_sigma_x_estimate_and_prior_update(q, sigma):
    g, ang, vol, y = small_problem(seed=21)
    rng = np.random.default_rng(4)
    x = vol + 0.05 * rng.random(vol.shape)
    c = 1e-3
    eng = O.Engine(g, ang, y, O.params(sigma_x=sigma, mrf_q=q, mrf_c=c), x0=x)
    nz = g.N_z
    z0, z1 = int(0.25 * nz), int(0.75 * nz)
    expect = _pair_sum(x, z0, z1, sigma, sigma, q, c) / (g.N_x * g.N_y * (z1 - z0 + 1))
    est, sx = eng.estimate_sigma_x()
    assert est == pytest.approx(expect, rel=1e-12)
    assert sx == pytest.approx(5.0 * (expect * q) ** (1.0 / q), rel=1e-12)
    assert eng.sigma_x() == sx
    # the prior term of the cost now uses the new SigmaX over the whole volume
    _, t = eng.cost(terms=True)
    assert t[1] == pytest.approx(_pair_sum(x, 0, nz, 1.0, sx, q, c), rel=1e-11)


def test_run_loop_reestimates_after_each_outer_iteration():
    g, ang, vol, y = small_problem(seed=22)
    kw = dict(num_outer=3, num_inner=2, stop_threshold=0.0, sigma_x=0.1)
    a = O.Engine(g, ang, y, O.params(estimate_prior=1, **kw))
    a.run()
    # the same loop by hand: sweeps, then an estimate after every outer iteration
    b = O.Engine(g, ang, y, O.params(**kw))
    n_iter = 2
    for outer in range(3):
        for inner in range(n_iter):
            b.sweep(outer, inner)
        n_iter = 1
        b.estimate_sigma_x()
    assert a.sigma_x() == b.sigma_x() and a.sigma_x() != 0.1
    np.testing.assert_array_equal(a.x(), b.x())
