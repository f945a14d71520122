"""The oracle's HAADF mode with a bright-field sinogram (the reference's BF_Flag,
HAADF/HAADF_ReconstructionEngine_Extra.cpp:89-175, 636-669, 718-727 and
HAADF_ReconstructionEngine_UpdateVoxels.cpp:354-362, 438-446), pinned by identities:

* a scale that is constant within each view, s = c_theta, is the plain model with the
  gains multiplied by c_theta: same error sinogram, same sweeps (a missing factor in
  theta, in the error update or in the initial error, or a wrong view index, breaks it);
* after sweeps with a scale that varies per element, e still equals y - s Yest(x) - mu
  for the current volume (a wrong element index in the error update breaks it);
* with bright-field data joint estimation moves each view's weighted least-squares
  offset of e (the minimiser of sum w (e - a)^2, solved by numpy's least squares) from
  e into mu and leaves the gains alone.
"""
import numpy as np

import oracle_py as O
from problems import small_problem

BF_OFFSET = 32768.0


def _engine(g, ang, y, gains, offsets, x0=None, **kw):
    # a target gain above every gain mean: the initial gains are taken as given (the
    # engine rescales them to the target when their mean exceeds it)
    return O.Engine(g, ang, y, O.params(**dict(dict(sigma_x=0.1, mrf_q=1.2, num_inner=2,
                                                    num_outer=1, target_gain=1e6), **kw)),
                    gains=gains, offsets=offsets, x0=x0)


def test_per_view_scale_equals_scaled_gains():
    g, ang, vol, y = small_problem(n_r=24, n_t=6, n_views=17, seed=8)
    nt = len(ang)
    rng = np.random.default_rng(1)
    gains = rng.uniform(0.9, 1.1, nt)
    offs = rng.uniform(-0.02, 0.02, nt)
    c = rng.uniform(0.5, 2.0, nt)
    bf = np.broadcast_to((c - BF_OFFSET)[:, None, None], (nt, g.N_r, g.N_t)).copy()
    a = _engine(g, ang, y, gains, offs, seed=3)
    a.set_bf_sinogram(bf)
    b = _engine(g, ang, y, gains * c, offs, seed=3)
    np.testing.assert_allclose(a.e(), b.e(), rtol=0, atol=1e-12)
    for it in range(2):
        a.sweep(0, it)
        b.sweep(0, it)
    np.testing.assert_allclose(a.x(), b.x(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(a.e(), b.e(), rtol=1e-9, atol=1e-11)


def test_element_scale_keeps_e_consistent_with_x():
    g, ang, vol, y = small_problem(n_r=24, n_t=6, n_views=17, seed=9)
    nt = len(ang)
    rng = np.random.default_rng(2)
    gains = rng.uniform(0.9, 1.1, nt)
    offs = rng.uniform(-0.02, 0.02, nt)
    s = rng.uniform(0.5, 1.5, (nt, g.N_r, g.N_t))
    a = _engine(g, ang, y, gains, offs, seed=4)
    a.set_bf_sinogram(s - BF_OFFSET)
    for it in range(3):
        a.sweep(0, it)
    x = a.x()
    assert np.abs(x).max() > 0
    # the forward projection of the current volume (with the gains) from a fresh engine
    yest = _engine(g, ang, y, gains, offs, x0=x).yest()
    e_expect = y.reshape(nt, g.N_r, g.N_t) - s * yest - offs[:, None, None]
    np.testing.assert_allclose(a.e(), e_expect, rtol=0, atol=1e-10)


def test_offset_only_estimation_with_bright_field_data():
    g, ang, vol, y = small_problem(n_r=24, n_t=6, n_views=17, seed=10)
    nt = len(ang)
    rng = np.random.default_rng(3)
    gains = np.full(nt, 1.05)
    offs = rng.uniform(-0.02, 0.02, nt)
    # per-view variances: the weights differ between views
    var = rng.uniform(0.5, 2.0, nt)
    a = O.Engine(g, ang, y, O.params(sigma_x=0.1, mrf_q=1.2, num_inner=2, num_outer=1,
                                     target_gain=1e6, joint_estimation=1, noise_model=1,
                                     seed=5),
                 gains=gains, offsets=offs, variances=var)
    a.set_bf_sinogram(rng.uniform(0.5, 1.5, (nt, g.N_r, g.N_t)) - BF_OFFSET)
    a.sweep(0, 0)
    e0, w = a.e().copy(), a.w().copy()
    assert np.abs(e0.reshape(nt, -1).mean(axis=1)).max() > 1e-3  # a non-trivial step
    data0 = 0.5 * np.sum(e0 ** 2 * w)
    a.joint_estimation()
    I0, mu, al = a.nuisance()
    step = np.empty(nt)
    for t in range(nt):
        sw = np.sqrt(w[t].reshape(-1))
        step[t] = np.linalg.lstsq(sw[:, None], sw * e0[t].reshape(-1), rcond=None)[0][0]
    np.testing.assert_allclose(mu, offs + step, rtol=0, atol=1e-13)
    np.testing.assert_allclose(a.e(), e0 - step[:, None, None], rtol=0, atol=1e-13)
    np.testing.assert_array_equal(I0, gains)
    # the new offset minimises the weighted data term over each view's constant shift
    e1 = a.e()
    data1 = 0.5 * np.sum(e1 ** 2 * w)
    assert data1 <= data0
    for d in (1e-3, -1e-3):
        assert 0.5 * np.sum((e1 + d) ** 2 * w) > data1
