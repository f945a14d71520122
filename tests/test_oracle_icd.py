"""Pinning of the oracle's projector, QGGMRF prior, ICD update, cost and nuisance
estimation (CPU), with checks that do not retype the oracle's formulas:
adjoint identities, finite differences, brute-force 1-D minimisation, an
independent numpy cost, KKT conditions at convergence, and the constrained
least-squares system of the gain/offset estimate solved with numpy.linalg.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import optimize

import oracle_py as O
from problems import small_problem

FILTER_CLASS = {2: 0.0523, 1: 0.0370, 0: 0.0302}  # faces / edges / corners


def _qg(sigma=0.3, q=1.2, c=1e-3, alpha=None):
    v = O.QG()
    O.lib().orc_qggmrf_init(sigma, q, c, C.byref(v))
    if alpha is not None:
        v.MRF_ALPHA = alpha
    return v


def rho(d, sigma, q, c, p=2.0):
    """QGGMRF potential of the reference's prior, written from its definition."""
    a = np.abs(d)
    return (a ** p / sigma ** p) / (c + a ** (p - q) / sigma ** (p - q))


def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-seeded mt19937 is 4123659995
    mt = (C.c_uint32 * 625)()
    lib = O.lib()
    lib.orc_mt_seed.argtypes = [C.c_void_p, C.c_uint32]
    lib.orc_mt_next.argtypes = [C.c_void_p]
    lib.orc_mt_next.restype = C.c_uint32
    lib.orc_mt_seed(C.addressof(mt), 5489)
    for _ in range(9999):
        lib.orc_mt_next(C.addressof(mt))
    assert lib.orc_mt_next(C.addressof(mt)) == 4123659995
    # and the raw stream equals numpy's legacy MT19937 for an integer seed
    lib.orc_mt_seed(C.addressof(mt), 12345)
    ours = [lib.orc_mt_next(C.addressof(mt)) for _ in range(8)]
    theirs = np.random.RandomState(12345).randint(0, 2 ** 32, size=8, dtype=np.uint64)
    assert ours == [int(t) for t in theirs]


def test_qggmrf_derivative_and_curvature():
    q = _qg()
    for d in (-2.0, -0.3, -1e-3, 1e-3, 0.05, 0.7, 3.0):
        val = O.lib().orc_qggmrf_value(d, C.byref(q))
        assert val == pytest.approx(rho(d, 0.3, 1.2, 1e-3), rel=1e-12)
        h = 1e-6 * max(1.0, abs(d))
        fd = (rho(d + h, 0.3, 1.2, 1e-3) - rho(d - h, 0.3, 1.2, 1e-3)) / (2 * h)
        assert O.lib().orc_qggmrf_derivative(d, C.byref(q)) == pytest.approx(fd, rel=1e-6)
    # rho'(d)/d -> P / (sigma^P c) as d -> 0
    lim = O.lib().orc_qggmrf_derivative(1e-15, C.byref(q)) / 1e-15  # |d|^(p-q) -> 0 slowly
    assert O.lib().orc_qggmrf_second_derivative_at_zero(C.byref(q)) == pytest.approx(lim, rel=1e-5)


def test_qggmrf_surrogate_majorizes():
    """b = rho'(d0)/d0 gives the symmetric quadratic bound
    rho(d) <= rho(d0) + b/2 (d^2 - d0^2), tight at +-d0."""
    q = _qg(0.2, 1.2, 1e-3)
    d = np.linspace(-3, 3, 2001)
    for d0 in (-1.5, -0.2, 0.01, 0.4, 2.0):
        b = O.lib().orc_qggmrf_derivative(d0, C.byref(q)) / d0
        bound = rho(d0, 0.2, 1.2, 1e-3) + 0.5 * b * (d ** 2 - d0 ** 2)
        assert np.all(bound >= rho(d, 0.2, 1.2, 1e-3) - 1e-9 * (1 + np.abs(bound)))


def test_functional_substitution_minimises_the_surrogate():
    rng = np.random.default_rng(3)
    F = np.zeros(27)
    O.lib().orc_filter(O.dp(F))
    for i in range(27):
        ones = sum(1 for a in (i // 9, (i // 3) % 3, i % 3) if a == 1)
        assert F[i] == (0.0 if ones == 3 else FILTER_CLASS[ones])
    assert F.sum() == pytest.approx(0.9994)
    for trial in range(20):
        nb = rng.uniform(0, 1, 27)
        nb[13] = 0.0
        bflag = (rng.uniform(size=27) < 0.85).astype(np.uint8)
        v = rng.uniform(0, 1)
        t1, t2 = rng.normal(0, 2), rng.uniform(0.5, 5)
        sig, qq, cc = 0.3, 1.2, 1e-3
        # weights of the quadratic surrogate at the current value v
        qv = _qg(sig, qq, cc, alpha=1.0)
        b = np.array([O.lib().orc_qggmrf_derivative(v - nb[i], C.byref(qv)) / (v - nb[i])
                      if v != nb[i] else O.lib().orc_qggmrf_second_derivative_at_zero(C.byref(qv))
                      for i in range(27)])
        use = (bflag == 1) & (np.arange(27) != 13)

        def Q(u):
            return t1 * (u - v) + 0.5 * t2 * (u - v) ** 2 + np.sum(
                0.5 * F[use] * b[use] * (u - nb[use]) ** 2)

        ustar = optimize.minimize_scalar(Q, bracket=(-5, 5), tol=1e-12).x
        bf = bflag.ctypes.data_as(C.POINTER(C.c_uint8))
        got1 = O.lib().orc_qggmrf_functional_substitution(-1e9, 1e9, v, bf, O.dp(F), O.dp(nb), t1,
                                                           t2, C.byref(qv))
        assert got1 == pytest.approx(ustar, abs=1e-7)
        q15 = _qg(sig, qq, cc)  # reference over-relaxation 1.5, clipped to [lo, hi]
        lo, hi = 0.1, 0.6
        got = O.lib().orc_qggmrf_functional_substitution(lo, hi, v, bf, O.dp(F), O.dp(nb), t1, t2,
                                                          C.byref(q15))
        assert got == pytest.approx(min(max(v + 1.5 * (ustar - v), lo), hi), abs=1e-7)


def _numpy_prior(x, sigma, q, c):
    """Prior term as the half-sum over every voxel's 26 neighbours (each pair twice)."""
    nz, nx, ny = x.shape
    xp = np.pad(x, 1)
    tot = 0.0
    for dz in (-1, 0, 1):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                if dz == dx == dy == 0:
                    continue
                f = FILTER_CLASS[(dz == 0) + (dx == 0) + (dy == 0)]
                nbv = xp[1 + dz:1 + dz + nz, 1 + dx:1 + dx + nx, 1 + dy:1 + dy + ny]
                mask = np.pad(np.ones_like(x), 1)[1 + dz:1 + dz + nz, 1 + dx:1 + dx + nx,
                                                   1 + dy:1 + dy + ny]
                tot += np.sum(f * rho(x - nbv, sigma, q, c) * mask)
    return tot / 2


def test_projector_adjoint_and_cost():
    g, ang, vol, y = small_problem()
    eng = O.Engine(g, ang, y, O.params(sigma_x=0.1))
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (g.N_z, g.N_x, g.N_y))
    e = rng.normal(0, 1, (g.N_theta, g.N_r, g.N_t))
    Ax = eng.project(x)
    # A^T e through the ICD theta accumulation: with w = 1, I0 = 1, -theta1 = (A^T e)_i
    eng_e = O.Engine(g, ang, e, O.params())  # x = 0 so its error sinogram is e itself
    ATe = np.zeros_like(x)
    for z in range(g.N_z):
        for xx in range(g.N_x):
            for yy in range(g.N_y):
                ATe[z, xx, yy] = -eng_e.theta(z, xx, yy)[0]
    assert np.sum(Ax * e) == pytest.approx(np.sum(x * ATe), rel=1e-10)
    # theta2 is the squared norm of the voxel's projection
    one = np.zeros_like(x)
    one[5, 11, 3] = 1.0
    assert eng_e.theta(5, 11, 3)[1] == pytest.approx(np.sum(eng.project(one) ** 2), rel=1e-12)
    # cost = 1/2 |e|^2 + prior, with the prior from an independent numpy formulation
    eng.set_x(x)
    c, terms = eng.cost(terms=True)
    ex = eng.e()
    assert terms[0] == pytest.approx(0.5 * np.sum(ex ** 2), rel=1e-12)
    assert terms[1] == pytest.approx(_numpy_prior(x, 0.1, 1.2, 1e-3), rel=1e-10)


def test_icd_cost_monotone_and_kkt():
    g, ang, vol, y = small_problem(n_r=24, n_t=4, n_views=21, thickness=6, seed=8)
    prm = O.params(sigma_x=0.5, num_inner=200)
    eng = O.Engine(g, ang, y, prm)
    costs = [eng.cost()]
    for it in range(150):
        eng.sweep(0, it)
        costs.append(eng.cost())
    d = np.diff(costs)
    assert np.all(d <= 1e-9 * np.abs(np.array(costs[:-1])))
    # KKT of min_x>=0 C(x): dC/dx_i = 0 where x_i > 0, >= 0 where x_i = 0
    x = eng.x()

    def C_of(xv):
        e = y - eng.project(xv)
        return 0.5 * np.sum(e ** 2) + _numpy_prior(xv, 0.5, 1.2, 1e-3)

    rng = np.random.default_rng(1)
    pos = np.argwhere(x > 1e-3)
    zero = np.argwhere(x == 0)
    h = 1e-5
    for idx in pos[rng.choice(len(pos), min(8, len(pos)), replace=False)]:
        xp, xm = x.copy(), x.copy()
        xp[tuple(idx)] += h
        xm[tuple(idx)] -= h
        grad = (C_of(xp) - C_of(xm)) / (2 * h)
        assert abs(grad) < 2e-3, (idx, grad)
    for idx in zero[rng.choice(len(zero), min(8, len(zero)), replace=False)] if len(zero) else []:
        xp = x.copy()
        xp[tuple(idx)] += h
        assert (C_of(xp) - C_of(x)) / h > -2e-3


def test_stop_threshold():
    g, ang, vol, y = small_problem()
    eng = O.Engine(g, ang, y, O.params(stop_threshold=0.5))
    s0, st0 = eng.sweep(0, 0)
    assert s0 == 1 and st0[1] == 0  # nothing to compare against on the first sweep
    s1, st1 = eng.sweep(0, 1)
    assert (s1 == 0) == (st1[0] / st1[1] < 0.5)


def test_joint_estimation_is_the_constrained_least_squares():
    rng = np.random.default_rng(2)
    nv = 15
    g, ang, vol, y = small_problem(n_views=nv, gains=rng.uniform(0.8, 1.2, nv),
                                   offsets=rng.uniform(-0.1, 0.1, nv))
    eng = O.Engine(g, ang, y, O.params(joint_estimation=1, noise_model=1))
    for it in range(3):
        eng.sweep(0, it)
    I0, mu, _ = eng.nuisance()
    e = eng.e()
    w = eng.w()
    yhat = (y - e - mu[:, None, None]) / I0[:, None, None]  # normalised projection
    eng.joint_estimation()
    I_new, mu_new, _ = eng.nuisance()
    # min sum_k 1/2 sum w (y - I_k yhat - mu_k)^2  s.t.  sum_k I_k = N * target
    n = nv
    A = np.zeros((2 * n + 1, 2 * n + 1))
    b = np.zeros(2 * n + 1)
    for k in range(n):
        wk, yk, hk = w[k].ravel(), y[k].ravel(), yhat[k].ravel()
        A[k, k] = np.sum(wk * hk * hk)
        A[k, n + k] = A[n + k, k] = np.sum(wk * hk)
        A[n + k, n + k] = np.sum(wk)
        A[k, 2 * n] = A[2 * n, k] = 1.0
        b[k] = np.sum(wk * hk * yk)
        b[n + k] = np.sum(wk * yk)
    b[2 * n] = n * 1.0
    sol = np.linalg.solve(A, b)
    np.testing.assert_allclose(I_new, sol[:n], rtol=1e-8)
    np.testing.assert_allclose(mu_new, sol[n:2 * n], rtol=1e-8, atol=1e-10)
    # the error sinogram follows the new nuisance parameters
    np.testing.assert_allclose(eng.e(), y - mu_new[:, None, None] - I_new[:, None, None] * yhat,
                               atol=1e-9)
    # noise variance = mean squared error of each view; weights its inverse
    eng.update_weights()
    _, _, alpha = eng.nuisance()
    np.testing.assert_allclose(alpha, np.mean(eng.e() ** 2, axis=(1, 2)), rtol=1e-12)
    np.testing.assert_allclose(eng.w(), 1.0 / alpha[:, None, None] * np.ones_like(y), rtol=1e-12)
    # ... which is the maximum-likelihood variance: the cost's noise terms
    # 1/2 sum e^2/alpha + 1/2 sum log(2 pi alpha) grow when any view's alpha moves
    prm = O.params(joint_estimation=1, noise_model=1)
    x = eng.x()
    c0 = O.Engine(g, ang, y, prm, gains=I_new, offsets=mu_new, variances=alpha, x0=x).cost()
    assert c0 == pytest.approx(eng.cost(), rel=1e-12)
    for k in (0, 7, nv - 1):
        for f in (0.97, 1.03):
            a = alpha.copy()
            a[k] *= f
            c = O.Engine(g, ang, y, prm, gains=I_new, offsets=mu_new, variances=a, x0=x).cost()
            assert c > c0


def test_run_costs_never_increase():
    rng = np.random.default_rng(4)
    nv = 21
    g, ang, vol, y = small_problem(n_views=nv, gains=rng.uniform(0.9, 1.1, nv),
                                   offsets=rng.uniform(-0.02, 0.02, nv))
    eng = O.Engine(g, ang, y, O.params(num_outer=5, num_inner=4, joint_estimation=1,
                                       noise_model=1, stop_threshold=1e-6))
    costs, nsw = eng.run()
    assert nsw == 4 + 4 and len(costs) == 1 + nsw + 2 * 5
    # the first weight update changes the cost's noise term from its initial variance,
    # so compare within each constant-weight stretch and after each nuisance step
    assert np.all(np.diff(costs[1:]) <= 1e-8 * np.abs(costs[1:-1]))


def _sweep_line_order(g, seed):
    """The line order of one oracle sweep (UpdateYSlice's draw, ..._UpdateVoxels.cpp:
    213-231: Index = next() % remaining, Counter[Index] = Counter[last])."""
    raw = np.random.RandomState(seed).randint(0, 2 ** 32, size=g.N_z * g.N_x, dtype=np.uint64)
    counter = list(range(g.N_z * g.N_x))
    out = []
    for t in range(g.N_z * g.N_x):
        idx = int(raw[t]) % len(counter)
        out.append(counter[idx])
        counter[idx] = counter[-1]
        counter.pop()
    return out


def test_unit_order_replay_equals_the_sweep():
    """update_units replays a given order of voxel updates.  Units of one slice each, in
    a sweep's own order (lines as drawn, slices ascending), must reproduce the sweep bit
    for bit; and a unit of many slices is its even slices then its odd slices."""
    g, ang, vol, y = small_problem()
    a = O.Engine(g, ang, y, O.params(seed=99))
    b = O.Engine(g, ang, y, O.params(seed=99))
    a.sweep(0, 0)  # sweep_counter 0: seed 99
    order = _sweep_line_order(g, 99)
    b.update_units([(l, yy, 1) for l in order for yy in range(g.N_y)])
    np.testing.assert_array_equal(a.x(), b.x())
    np.testing.assert_array_equal(a.e(), b.e())
    c = O.Engine(g, ang, y, O.params(seed=99), default_recon=0.2)
    d = O.Engine(g, ang, y, O.params(seed=99), default_recon=0.2)
    y0, ny = 1, g.N_y - 3
    lines = order[:50]
    c.update_units([(l, y0, ny) for l in lines])
    d.update_units([(l, y0 + k, 1) for l in lines
                    for k in list(range(0, ny, 2)) + list(range(1, ny, 2))])
    np.testing.assert_array_equal(c.x(), d.x())
    x = c.x()
    touched = np.zeros(x.shape, bool)
    for l in lines:
        touched[l // g.N_x, l % g.N_x, y0:y0 + ny] = True
    assert np.all(x[~touched] == 0.2) and np.any(x[touched] != 0.2)


def test_sweep_slab_is_a_slab_of_the_sweep():
    """sweep_slab (the bounded CPU-baseline sample) runs a sweep's own line order over a
    y-range: over all slices it equals the sweep bit for bit; over [0, 4) it updates
    exactly those slices."""
    g, ang, vol, y = small_problem()
    a = O.Engine(g, ang, y, O.params(seed=5), default_recon=0.2)
    b = O.Engine(g, ang, y, O.params(seed=5), default_recon=0.2)
    a.sweep(0, 0)
    b.sweep_slab(0, 0, 0, g.N_y)
    np.testing.assert_array_equal(a.x(), b.x())
    c = O.Engine(g, ang, y, O.params(seed=5), default_recon=0.2)
    st = c.sweep_slab(0, 0, 0, 4)
    x = c.x()
    assert st[3] > 0 and np.all(x[:, :, 4:] == 0.2) and np.any(x[:, :, :4] != 0.2)
