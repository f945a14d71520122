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
