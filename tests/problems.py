"""Seeded test problems shared by the oracle and GPU tests."""
import numpy as np

import oracle_py as O
from mbirtem import phantom as P


def small_problem(n_r=32, n_t=8, n_views=31, max_deg=60.0, thickness=8, seed=5, noise=0.02,
                  gains=None, offsets=None):
    """A small HAADF problem: geometry, angles, phantom, noisy sinogram (FP64 oracle projection)."""
    ang = P.tilt_series(n_views, max_deg)
    g = O.geometry(n_r, n_t, ang, thickness=thickness)
    shape = (g.N_z, g.N_x, g.N_y)
    vol = P.random_ellipsoids(shape, n=3, seed=seed).astype(np.float64)
    prm = O.params()
    eng = O.Engine(g, ang, np.zeros((n_views, n_r, n_t)), prm, x0=vol)
    proj = eng.yest()
    y = P.noisy_sinogram(proj, noise, seed + 1, gains=gains, offsets=offsets)
    return g, ang, vol, y


def bf_problem(n_r=32, n_t=6, n_views=25, max_deg=60.0, thickness=8, dose=2000.0, seed=3,
               bragg_frac=0.0):
    """A small bright-field problem: transmitted counts Poisson(dose exp(-A x)) (FP64
    oracle projection), optionally a fraction of entries with far fewer counts than the
    model predicts (Bragg diffraction)."""
    ang = P.tilt_series(n_views, max_deg)
    g = O.geometry(n_r, n_t, ang, thickness=thickness)
    vol = 0.02 * P.random_ellipsoids((g.N_z, g.N_x, g.N_y), n=3, seed=seed).astype(np.float64)
    eng = O.Engine(g, ang, np.zeros((n_views, n_r, n_t)), O.params(), x0=vol)
    rng = np.random.default_rng(seed + 1)
    counts = rng.poisson(dose * np.exp(-eng.yest())).astype(np.float64)
    if bragg_frac > 0:
        m = rng.uniform(size=counts.shape) < bragg_frac
        counts[m] *= 0.3
    return g, ang, vol, np.maximum(counts, 1.0)
