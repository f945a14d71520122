"""Pinning of the oracle's bright-field model (CPU): log counts and weights, the
Bragg rejector and its half-quadratic surrogate, theta against the Bragg data cost,
the offset and variance updates, the threshold schedule and monotone costs.

The checks come from outside the restated formulas: the delta-method variance of
log Poisson counts, the majorization properties any correct surrogate must have,
finite differences of the engine's own cost, and scipy minimisations.
"""
import numpy as np
import pytest
from scipy import optimize

import oracle_py as O
from problems import bf_problem


def bf_params(**kw):
    base = dict(bf_mode=1, mrf_c=0.01, sigma_x=0.01, mrf_q=1.2, bragg_T=3.0, bragg_delta=0.5)
    base.update(kw)
    return O.params(**base)


def test_log_counts_and_weights_match_poisson_statistics():
    g, ang, vol, counts = bf_problem()
    eng = O.Engine(g, ang, counts, bf_params(), x0=np.zeros((g.N_z, g.N_x, g.N_y)))
    # with x = 0 and mu = 0 the error sinogram is the measurement -log(counts)
    np.testing.assert_allclose(eng.e(), -np.log(counts), rtol=1e-14)
    # the weight is the inverse variance of -log(N) for N ~ Poisson(lambda): by the
    # delta method Var(-log N) ~ 1/lambda, so w ~ lambda; check against sampling
    lam = 2000.0
    smp = np.random.default_rng(0).poisson(lam, 200000)
    var = np.var(-np.log(smp))
    assert 1.0 / var == pytest.approx(lam, rel=2e-2)
    np.testing.assert_allclose(eng.w(), counts, rtol=1e-14)  # exp(+log N) = N, alpha = 1
    # offset added before the log, variance divides the weights
    eng2 = O.Engine(g, ang, counts, bf_params(bf_offset=5.0, noise_model=1),
                    x0=np.zeros((g.N_z, g.N_x, g.N_y)), variances=np.full(len(ang), 4.0))
    np.testing.assert_allclose(eng2.e(), -np.log(counts + 5.0), rtol=1e-14)
    np.testing.assert_allclose(eng2.w(), (counts + 5.0) / 4.0, rtol=1e-14)
    assert eng.sel().min() == 1  # every entry selected at setup


def test_bragg_surrogate_majorizes_the_rejector():
    T, d, w = 3.0, 0.5, 2.5
    es = np.linspace(-20, 20, 4001)
    beta = np.array([O.lib().orc_bf_beta(e, w, T, d) for e in es])
    # continuous at the threshold |e sqrt(w)| = T
    e_t = T / np.sqrt(w)
    lo = O.lib().orc_bf_beta(e_t * (1 - 1e-12), w, T, d)
    hi = O.lib().orc_bf_beta(e_t * (1 + 1e-12), w, T, d)
    assert lo == pytest.approx(hi, rel=1e-9)
    for e0 in (-9.0, -4.0, 2.5, 6.0, 15.0):  # rejected points
        assert e0 * e0 * w >= T * T
        qc = d * T / (abs(e0) * np.sqrt(w))  # computeTheta's weight of a rejected entry
        b0 = O.lib().orc_bf_beta(e0, w, T, d)
        q = qc * w * es ** 2 + (b0 - qc * w * e0 ** 2)
        assert np.all(q >= beta - 1e-9)  # majorizes everywhere
        # touches at e0 with the same slope
        h = 1e-6
        db = (O.lib().orc_bf_beta(e0 + h, w, T, d) - O.lib().orc_bf_beta(e0 - h, w, T, d)) / (2 * h)
        assert 2 * qc * w * e0 == pytest.approx(db, rel=1e-6)


def _data_cost(g, ang, counts, x, prm, T):
    eng = O.Engine(g, ang, counts, prm, x0=x)
    eng.bragg(T)
    eng.update_selector()
    c, t = eng.cost(terms=True)
    return t[0], eng


def test_theta_is_the_bragg_cost_derivative_and_majorizes_it():
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    prm = bf_params()
    eng = O.Engine(g, ang, counts, prm, x0=vol)
    eng.update_selector()
    assert eng.sel().mean() < 0.99  # some entries are rejected
    for (z, x, y) in ((4, 12, 2), (8, 16, 3), (10, 20, 1)):
        t1, t2 = eng.theta(z, x, y)
        D0, _ = _data_cost(g, ang, counts, vol, prm, 3.0)
        h = 1e-6
        xp, xm = vol.copy(), vol.copy()
        xp[z, x, y] += h
        xm[z, x, y] -= h
        Dp, _ = _data_cost(g, ang, counts, xp, prm, 3.0)
        Dm, _ = _data_cost(g, ang, counts, xm, prm, 3.0)
        # THETA1 = -sum e PE w (stored negated, :369) is the data cost's slope
        assert (Dp - Dm) / (2 * h) == pytest.approx(t1, rel=1e-4, abs=1e-6)
        for step in (-0.05, -0.01, 0.003, 0.02, 0.08):
            xs = vol.copy()
            xs[z, x, y] += step
            Ds, _ = _data_cost(g, ang, counts, xs, prm, 3.0)
            assert Ds <= D0 + t1 * step + 0.5 * t2 * step * step + 1e-8 * abs(D0)


def test_offset_step_minimises_the_surrogate_and_lowers_the_cost():
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    prm = bf_params(joint_estimation=1, noise_model=1)
    eng = O.Engine(g, ang, counts, prm, offsets=np.full(len(ang), -np.log(2000.0) + 0.05))
    for it in range(3):
        eng.sweep(0, it)
    eng.update_selector()
    e, w, sel = eng.e(), eng.w(), eng.sel()
    T, d = eng.bragg(), 0.5
    mu0 = eng.nuisance()[1].copy()
    c0 = eng.cost(terms=True)[1][0]
    eng.joint_estimation()
    mu1 = eng.nuisance()[1]
    for k in (0, 7, len(ang) - 1):
        ek, wk, sk = e[k].ravel(), w[k].ravel(), sel[k].ravel()
        qc = np.where(sk == 1, 1.0, d * T / (np.abs(ek) * np.sqrt(wk)))

        def surrogate(dm):
            return np.sum(qc * wk * (ek - dm) ** 2)

        r = optimize.minimize_scalar(surrogate, bracket=(-1, 1), tol=1e-12)
        assert mu1[k] - mu0[k] == pytest.approx(r.x, abs=1e-9)
    # majorize-minimize: the Bragg data cost did not rise (same selector)
    assert eng.cost(terms=True)[1][0] <= c0 + 1e-9 * abs(c0)


def test_shared_variance_is_the_maximum_likelihood_value_without_rejections():
    g, ang, vol, counts = bf_problem()
    prm = bf_params(noise_model=1, bragg_T=1e9)
    eng = O.Engine(g, ang, counts, prm, x0=vol)
    eng.bragg(1e9)
    eng.update_weights()
    alpha = eng.nuisance()[2]
    assert np.all(alpha == alpha[0])  # one value for every view
    # data + noise terms: 1/2 sum e^2 w_u / a + (M/2) log a, minimised at a = S / M
    e = eng.e()
    S = np.sum(e * e * counts)
    M = e.size
    assert alpha[0] == pytest.approx(S / M, rel=1e-12)
    f = lambda a: 0.5 * S / a + (M // 2) * np.log(a)
    r = optimize.minimize_scalar(f, bracket=(alpha[0] * 0.5, alpha[0] * 2), tol=1e-12)
    assert alpha[0] == pytest.approx(r.x, rel=1e-5)


def test_shared_variance_with_rejections_is_a_majorize_minimize_step():
    """With rejected entries the shared variance has no closed form.  The data and noise
    terms of the engine's cost, as a function of the variance a with e, the unscaled
    weights w_u and the selector fixed, are
        f(a) = 1/2 sum_sel e^2 w_u / a + delta T sum_rej |e| sqrt(w_u / a) + (M/2) log a.
    Bounding a^(-1/2) <= (sqrt(a0) / a + 1 / sqrt(a0)) / 2 (AM-GM, equal at a0) gives a
    majoriser whose minimiser is the reference's update
    (BrightField/BFForwardModel.cpp:447-511).  So each update must not raise f, and
    repeated updates must reach f's minimiser (scipy); a wrong power of a, a missing
    delta T or a dropped sum moves the fixed point."""
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    prm = bf_params(joint_estimation=1, noise_model=1)
    eng = O.Engine(g, ang, counts, prm, offsets=np.full(len(ang), -np.log(2000.0) + 0.05))
    for it in range(3):
        eng.sweep(0, it)
    eng.update_selector()
    sel = eng.sel()
    assert 1e-3 < 1.0 - sel.mean() < 0.5  # rejections take part
    e = eng.e()
    a0 = eng.nuisance()[2][0]
    wu = eng.w() * a0
    T, d = eng.bragg(), 0.5
    M = e.size
    assert M % 2 == 0  # the engine's noise term halves M as an integer
    rej = sel == 0
    S1 = np.sum(np.where(rej, 0.0, e * e * wu))
    S2 = np.sum(np.where(rej, np.abs(e) * np.sqrt(wu), 0.0))

    def f(a):
        return 0.5 * S1 / a + d * T * S2 / np.sqrt(a) + (M // 2) * np.log(a)

    # f is the engine's own data + noise cost up to a constant
    const = 0.5 * np.sum(rej) * T * T * (1 - 2 * d)
    _, t = eng.cost(terms=True)
    assert t[0] + t[2] == pytest.approx(f(a0) + const, rel=1e-12)
    r = optimize.minimize_scalar(f, bounds=(a0 * 1e-3, a0 * 1e3), method="bounded",
                                 options=dict(xatol=1e-12 * a0))
    fa = f(a0)
    for k in range(80):
        eng.update_weights()
        a = eng.nuisance()[2]
        assert np.all(a == a[0])
        assert f(a[0]) <= fa + 1e-12 * abs(fa), k  # majorize-minimize: f never rises
        fa = f(a[0])
    assert abs(a[0] - a0) > 1e-3 * a0  # the fixed point is away from the start
    assert a[0] == pytest.approx(r.x, rel=1e-6)
    _, t = eng.cost(terms=True)
    assert t[0] + t[2] == pytest.approx(f(a[0]) + const, rel=1e-12)


def test_threshold_schedule():
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    # NumIter > 1: the first outer iteration runs without rejection, then T applies
    eng = O.Engine(g, ang, counts, bf_params(num_outer=1, num_inner=3, stop_threshold=0))
    eng.run()
    assert eng.bragg() == 3.0
    assert eng.sel().min() == 1  # no selector refresh after the switch yet
    eng = O.Engine(g, ang, counts, bf_params(num_outer=3, num_inner=3, stop_threshold=0,
                                            joint_estimation=1))
    eng.run()
    assert eng.bragg() == 3.0 and eng.sel().min() == 0  # refreshed by the offset step
    # NumIter == 1: T from the start
    eng = O.Engine(g, ang, counts, bf_params(num_outer=2, num_inner=1, stop_threshold=0))
    eng.sweep(0, 0)
    assert eng.bragg() == 3.0


def test_bf_costs_never_increase():
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    prm = bf_params(num_outer=6, num_inner=4, stop_threshold=1e-7, joint_estimation=1,
                    noise_model=1)
    eng = O.Engine(g, ang, counts, prm, offsets=np.full(len(ang), -np.log(2000.0)))
    costs, nsw = eng.run()
    assert nsw == 4 + 5
    # cost 0 is at setup; costs 1-3 are the first outer iteration's sweeps (no
    # rejection; the threshold switches after the last one, whose cost is not
    # recorded); the rest come after the switch
    c = np.asarray(costs)
    assert np.all(np.diff(c[:4]) <= 1e-9 * np.abs(c[1:4]))
    # after the first nuisance step the variance term changes the objective;
    # compare sweeps and nuisance steps from the second outer iteration on
    assert np.all(np.diff(c[5:]) <= 1e-9 * np.abs(c[6:]))
