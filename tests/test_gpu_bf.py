"""GPU: the bright-field path against the FP64 oracle.

From identical states (the same volume, counts and nuisance parameters) the GPU's
log counts, weights, selector, offset and variance steps and cost must match the
oracle to float32 precision; over sweeps the update order differs, so costs must
fall and the converged volumes agree within the oracle's own seed-to-seed spread.
"""
import numpy as np
import pytest

import oracle_py as O
from mbirtem import capi
from problems import bf_problem

pytestmark = pytest.mark.gpu

BF = dict(bf_mode=1, mrf_c=0.01, sigma_x=0.01, mrf_q=1.2, bragg_T=3.0, bragg_delta=0.5)


def _pair(ctx, g, ang, counts, x0=None, offsets=None, **kw):
    prm = dict(BF, **kw)
    eng = O.Engine(g, ang, counts, O.params(**prm), x0=x0, offsets=offsets)
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    job = capi.Job(ctx, geom, ang, counts.astype(np.float32), init_volume=x0,
                   init_offset=offsets, **prm)
    return eng, job  # both start from the variance 1 (the defaults)


def test_bf_setup_matches_oracle(gpu_ctx):
    g, ang, vol, counts = bf_problem()
    eng, job = _pair(gpu_ctx, g, ang, counts, x0=vol, noise_model=1)
    ref_e = eng.e()
    assert np.max(np.abs(job.error_sinogram() - ref_e)) <= 1e-5 * np.max(np.abs(ref_e))
    np.testing.assert_allclose(job.weights(), eng.w(), rtol=1e-6)
    assert job.selector().min() == 1
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)


def test_bf_offset_and_variance_steps_match_oracle(gpu_ctx):
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    offs = np.full(len(ang), -np.log(2000.0) + 0.03)
    eng, job = _pair(gpu_ctx, g, ang, counts, x0=vol, offsets=offs, joint_estimation=1,
                     noise_model=1)
    for step in range(3):  # the first step uses the all-selected setup selector
        eng.joint_estimation()
        eng.update_selector()
        eng.update_weights()
        eng.update_selector()
        job.nuisance_update()
        I0, mu, al = eng.nuisance()
        I0g, mug, alg = job.nuisance()
        np.testing.assert_allclose(mug, mu, rtol=0, atol=2e-6)
        np.testing.assert_allclose(alg, al, rtol=2e-5)
        np.testing.assert_array_equal(I0g, I0)  # the gain stays at the target
        sel_g, sel_r = job.selector(), eng.sel()
        # entries at the threshold may fall either side in float32
        assert np.mean(sel_g != sel_r) < 1e-3, np.mean(sel_g != sel_r)
    assert eng.sel().mean() < 0.99  # rejections took part
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-5)


def test_bf_sweeps_lower_the_cost_and_keep_e_consistent(gpu_ctx):
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    offs = np.full(len(ang), -np.log(2000.0))
    eng, job = _pair(gpu_ctx, g, ang, counts, offsets=offs)
    costs = [job.cost()]
    for it in range(8):
        job.sweep(0, it)
        costs.append(job.cost())
    assert all(b <= a * (1 + 1e-6) + 1e-9 for a, b in zip(costs, costs[1:]))
    # e = y - A x - mu for the GPU's own volume (oracle projection, FP64)
    y = -np.log(counts)
    e_ref = y - eng.project(job.volume().astype(np.float64)) - offs[:, None, None]
    assert np.max(np.abs(job.error_sinogram() - e_ref)) <= 2e-5 * np.max(np.abs(e_ref))
    # the selector is the rule |e sqrt(w)| < T for every refreshed entry
    e, w, sel = job.error_sinogram(), job.weights(), job.selector()
    rule = (np.abs(e * np.sqrt(w)) < 3.0).astype(np.uint8)
    assert np.mean(sel != rule) < 1e-3


def test_bf_run_matches_oracle(gpu_ctx):
    """BFReconstructionEngine's loop (NumIter 6 then 14 outer iterations with the
    offset and variance steps).  Oracle runs with other seeds lie within a few
    percent of each other; the GPU must lie within 3 of their scatters of their mean."""
    g, ang, vol, counts = bf_problem(bragg_frac=0.03)
    offs = np.full(len(ang), -np.log(2000.0))
    kw = dict(num_outer=15, num_inner=6, joint_estimation=1, noise_model=1, stop_threshold=1e-5)
    runs = []
    for seed in (12345, 1, 2, 3):
        e = O.Engine(g, ang, counts, O.params(**dict(BF, seed=seed, **kw)), offsets=offs)
        c, _ = e.run()
        runs.append((e.x(), c[-1], e.nuisance()))
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    job = capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs, seed=7,
                   **dict(BF, **kw))
    rs, cg = job.run()
    assert rs.sweeps == 6 + 14 or rs.status == 0

    def rms(a):
        return np.sqrt(np.mean(a ** 2))

    xbar = np.mean([r[0] for r in runs], axis=0)
    spread = max(rms(r[0] - xbar) for r in runs)
    xg = job.volume().astype(np.float64)
    assert rms(xg - xbar) <= max(3.0 * spread, 1e-2 * rms(xbar)), (rms(xg - xbar), spread)
    assert cg[-1] == pytest.approx(np.mean([r[1] for r in runs]), rel=2e-3)
    mu = np.mean([r[2][1] for r in runs], axis=0)
    al = np.mean([r[2][2] for r in runs], axis=0)
    _, mug, alg = job.nuisance()
    np.testing.assert_allclose(mug, mu, atol=2e-3)
    np.testing.assert_allclose(alg, al, rtol=5e-2)


def test_bf_nhicd_lists_and_stop_call_match_oracle(gpu_ctx):
    """The NHICD schedule on the GPU against the oracle's restatement, same seed.  The
    homogeneous calls (even calls) take the same lines: one list of all lines drawn by
    the reference's list algorithm from the same MT19937 seed.  A non-homogeneous call
    takes the lines above the (n - n/20)-th filtered update magnitude; the magnitudes
    differ between the two only through the update order inside a call and float32
    against FP64.  After the first homogeneous call (a twentieth of the lines, far apart)
    the magnitudes hardly depend on the order, so the first non-homogeneous lists must
    overlap by at least 90 %; later the magnitudes come from many interacting updates
    in different orders, and the lists must only have the oracle's size.  The stop test
    (every 20 calls) must fire at the same call."""
    g, ang, vol, counts = bf_problem(bragg_frac=0.03)
    offs = np.full(len(ang), -np.log(2000.0))
    kw = dict(BF, bf_nhicd=1, num_outer=1, num_inner=2000, stop_threshold=2e-2,
              zero_skipping=0, seed=77)
    eng = O.Engine(g, ang, counts, O.params(**kw), offsets=offs)
    _, n_r = eng.run()
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    job = capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs, **kw)
    job.set_nhicd_log(True)
    rs, _ = job.run()
    lr, lg = eng.nhicd_lists(), job.nhicd_lists()
    assert len(lr) == n_r and len(lg) == rs.sweeps
    assert n_r < 2000 and (n_r - 1) % 20 == 0
    assert rs.sweeps == n_r, (rs.sweeps, n_r)
    for c, (a, b) in enumerate(zip(lr, lg)):
        if c % 2 == 0:
            np.testing.assert_array_equal(np.sort(b), np.sort(a))
        else:
            assert len(a) > 0 and abs(len(b) - len(a)) <= 2, (c, len(a), len(b))
    a, b = lr[1], lg[1]
    assert len(np.intersect1d(a, b)) >= 0.9 * max(len(a), len(b)), (np.sort(a), np.sort(b))


def test_bf_nhicd_call_steps_are_conflict_free(gpu_ctx):
    """An NHICD call's lines are packed into few conflict-free steps, chains across rows.
    The steps of the last call (homogeneous and non-homogeneous calls alike) hold every
    line of the call once, and no two lines of a step share a detector bin (their
    exported columns) or are prior neighbours, so each step is still equivalent to any
    sequential order of its lines."""
    g, ang, vol, counts = bf_problem(n_r=64, n_t=6, n_views=31, bragg_frac=0.03)
    offs = np.full(len(ang), -np.log(2000.0))
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    for calls in (5, 6):  # the last call homogeneous (even), then non-homogeneous
        kw = dict(BF, bf_nhicd=1, num_outer=1, num_inner=calls, stop_threshold=0.0,
                  zero_skipping=0, seed=5)
        job = capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs, **kw)
        job.set_nhicd_log(True)
        rs, _ = job.run()
        assert rs.sweeps == calls
        steps = job.schedule()
        last = job.nhicd_lists()[-1]
        units = np.concatenate(steps)
        nx = geom.N_x
        cols = {}
        for l in np.unique(last):
            idx, _ = job.column(int(l) // nx, int(l) % nx)
            if len(idx):
                cols[int(l)] = set(idx.tolist())
        np.testing.assert_array_equal(np.sort(units[:, 0]), np.sort(list(cols)))
        assert len(steps) < len(units)
        for st in steps:
            ls = [int(l) for l in st[:, 0]]
            for a in range(len(ls)):
                for b in range(a + 1, len(ls)):
                    za, xa, zb, xb = ls[a] // nx, ls[a] % nx, ls[b] // nx, ls[b] % nx
                    assert abs(za - zb) > 1 or abs(xa - xb) > 1, (ls[a], ls[b])
                    assert not (cols[ls[a]] & cols[ls[b]]), (ls[a], ls[b])
        job.close()


def test_bf_nhicd_converges_to_the_full_sweep_minimum(gpu_ctx):
    """The NHICD schedule (homogeneous twentieths alternating with the filtered top-5 %
    lists) only reorders the updates.  With one outer iteration the Bragg threshold
    stays at its "regular" value 5e10, the problem is convex, and NHICD must reach the
    minimum the full random-order sweeps reach.  Zero skipping is off here: with it,
    NHICD from a zero volume (as the reference's, BFUpdateYSlice.cpp:228-232) skips
    from the second call on every zero voxel whose neighbours are zero, so a region
    that no line of the first twentieth reached stays at zero."""
    g, ang, vol, counts = bf_problem(bragg_frac=0.03)
    offs = np.full(len(ang), -np.log(2000.0))
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    c32 = counts.astype(np.float32)
    full = capi.Job(gpu_ctx, geom, ang, c32, init_offset=offs, seed=7, num_outer=1,
                    num_inner=40, stop_threshold=1e-9, zero_skipping=0, **BF)
    _, cf = full.run()
    nh = capi.Job(gpu_ctx, geom, ang, c32, init_offset=offs, seed=7, num_outer=1,
                  num_inner=1000, stop_threshold=1e-9, zero_skipping=0, bf_nhicd=1, **BF)
    rs, cn = nh.run()
    assert rs.sweeps == 1000
    assert all(b <= a * (1 + 1e-6) for a, b in zip(cn, cn[1:]))  # convex: never up
    xf, xn = full.volume().astype(np.float64), nh.volume().astype(np.float64)
    rel = np.sqrt(np.mean((xn - xf) ** 2)) / np.sqrt(np.mean(xf ** 2))
    assert rel < 1e-2, rel
    assert cn[-1] == pytest.approx(cf[-1], rel=1e-4)


def test_bf_nhicd_stop_rule_every_twenty_calls(gpu_ctx):
    """The NHICD stop test runs on calls 20, 40, ... (ROI update magnitude of the map
    against the ROI volume sum taken 20 calls before): a threshold above any ratio
    ends the run after the 21st call, a zero threshold never."""
    g, ang, vol, counts = bf_problem()
    offs = np.full(len(ang), -np.log(2000.0))
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    c32 = counts.astype(np.float32)
    for stop, calls in ((10.0, 21), (0.0, 45)):
        job = capi.Job(gpu_ctx, geom, ang, c32, init_offset=offs, seed=3, num_outer=1,
                       num_inner=45, stop_threshold=stop, bf_nhicd=1, **BF)
        rs, _ = job.run()
        assert rs.sweeps == calls, (stop, rs.sweeps)


def test_bf_nhicd_run_without_costs_equals_run_with_costs(gpu_ctx):
    """A run that records no costs (costs NULL, as the reference runs without its
    COST_CALCULATE debugging build) takes the same calls and leaves the same volume,
    offsets and variance, bit for bit, as one that records a cost after every call."""
    g, ang, vol, counts = bf_problem()
    offs = np.full(len(ang), -np.log(2000.0))
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    c32 = counts.astype(np.float32)
    out = []
    for cap in (4096, 0):
        job = capi.Job(gpu_ctx, geom, ang, c32, init_offset=offs, seed=3, num_outer=2,
                       num_inner=25, stop_threshold=0.0, bf_nhicd=1, **BF)
        rs, costs = job.run(cap=cap)
        assert (len(costs) == 0) if cap == 0 else (len(costs) > rs.sweeps), (cap, len(costs))
        out.append((rs.sweeps, job.volume(), job.nuisance()))
        job.close()
    (s0, v0, n0), (s1, v1, n1) = out
    assert s0 == s1 == 26
    assert np.array_equal(v0, v1)
    for a, b in zip(n0, n1):
        assert np.array_equal(a, b)


def test_bf_jobs_in_the_server_equal_alone(gpu_ctx):
    """Bright-field jobs in the request server (their own launches, beside HAADF jobs)
    follow the same loop as mbir_job_run: threshold schedule, stop test before the
    nuisance step, offset and shared-variance steps.  Each ends bit for bit as
    alone."""
    from problems import small_problem

    cases = []
    for i, (frac, n_r, n_t, nin, nout, je, nm) in enumerate(
            [(0.03, 32, 6, 6, 1, 0, 0), (0.05, 40, 8, 3, 3, 1, 1), (0.02, 24, 4, 1, 2, 1, 0)]):
        g, ang, vol, counts = bf_problem(n_r=n_r, n_t=n_t, seed=30 + i, bragg_frac=frac)
        geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
        kw = dict(BF, init_offset=np.full(len(ang), -np.log(2000.0)), seed=11 + i,
                  num_inner=nin, num_outer=nout, joint_estimation=je, noise_model=nm,
                  stop_threshold=1e-3)
        cases.append((geom, ang, counts.astype(np.float32), kw))
    g, ang, vol, y = small_problem(n_r=32, n_t=8, n_views=31, seed=41)
    cases.append((capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2), ang,
                  y.astype(np.float32), dict(seed=5, stop_threshold=2e-3, num_inner=20)))
    refs = []
    for geom, ang, c, kw in cases:
        job = capi.Job(gpu_ctx, geom, ang, c, **kw)
        rs, _ = job.run(cap=0)
        refs.append((job.volume(), job.error_sinogram(), job.nuisance(), rs.sweeps))
        job.close()
    srv = capi.Server(gpu_ctx, max_jobs=8)
    tickets = [srv.submit(geom, ang, c, **kw) for geom, ang, c, kw in cases]
    for t, (vol_r, err_r, (I0, mu, al), sweeps) in zip(tickets, refs):
        rep = srv.wait(t, timeout=60.0)
        assert rep is not None and rep.state == capi.JOB_DONE and rep.error == 0
        assert rep.sweeps == sweeps
        job = srv.job(t)
        np.testing.assert_array_equal(job.volume(), vol_r)
        np.testing.assert_array_equal(job.error_sinogram(), err_r)
        I0s, mus, als = job.nuisance()
        np.testing.assert_array_equal(mus, mu)
        np.testing.assert_array_equal(als, al)
        srv.release(t)
    st = srv.stats()
    assert st.jobs_done == len(cases) and st.jobs_failed == 0 and st.fatal_error == 0
    srv.close()


def test_bf_jobs_in_a_batched_sweep_equal_alone(gpu_ctx):
    """Bright-field jobs in mbir_batch_sweep (launches of their own, beside a HAADF
    job): every sweep ends bit for bit as the job's own sweep."""
    from problems import small_problem

    specs = []
    for i, (frac, n_r, n_t) in enumerate([(0.03, 32, 6), (0.05, 40, 8)]):
        g, ang, vol, counts = bf_problem(n_r=n_r, n_t=n_t, seed=50 + i, bragg_frac=frac)
        specs.append((capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2), ang,
                      counts.astype(np.float32),
                      dict(BF, init_offset=np.full(len(ang), -np.log(2000.0)), seed=21 + i)))
    g, ang, vol, y = small_problem(n_r=32, n_t=8, n_views=31, seed=42)
    specs.append((capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2), ang,
                  y.astype(np.float32), dict(seed=6)))
    alone = [capi.Job(gpu_ctx, *s[:3], **s[3]) for s in specs]
    batched = [capi.Job(gpu_ctx, *s[:3], **s[3]) for s in specs]
    for it in range(3):
        for j in alone:
            j.sweep(0, it)
        capi.batch_sweep(gpu_ctx, batched, outer=[0] * 3, inner=[it] * 3)
        for a, b in zip(alone, batched):
            np.testing.assert_array_equal(b.volume(), a.volume())
            np.testing.assert_array_equal(b.error_sinogram(), a.error_sinogram())
    for j in alone + batched:
        j.close()


@pytest.mark.parametrize("n_t,y_slabs", [(1, 1), (13, 1), (100, 1), (64, 2)])
def test_bf_line_lengths(gpu_ctx, n_t, y_slabs):
    """Bright-field lines of 1, 13 and 100 slices (100 as whole lines: two per lane) and
    64 slices in two y-slab units: the setup against the oracle, the cost never rises
    over sweeps, e stays y - A x - mu for the GPU's own volume, the selector follows its
    rule, and the same job in a batched sweep ends every sweep bit for bit as alone."""
    g, ang, vol, counts = bf_problem(n_r=24, n_t=n_t, n_views=21, seed=13, bragg_frac=0.05)
    assert g.N_y == n_t
    offs = np.full(len(ang), -np.log(2000.0))
    eng = O.Engine(g, ang, counts, O.params(**BF), offsets=offs)
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    jobs = [capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs,
                     y_slabs=y_slabs, seed=5, kernel_mode=m, **BF) for m in (0, 1)]
    job, bat = jobs
    ref_e = eng.e()
    assert np.max(np.abs(job.error_sinogram() - ref_e)) <= 1e-5 * np.max(np.abs(ref_e))
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)
    costs = [job.cost()]
    for it in range(6):
        job.sweep(0, it)
        capi.batch_sweep(gpu_ctx, [bat], outer=[0], inner=[it])
        np.testing.assert_array_equal(bat.volume(), job.volume())
        np.testing.assert_array_equal(bat.error_sinogram(), job.error_sinogram())
        costs.append(job.cost())
    assert all(b <= a * (1 + 1e-6) + 1e-9 for a, b in zip(costs, costs[1:])), costs
    y = -np.log(counts)
    e_ref = y - eng.project(job.volume().astype(np.float64)) - offs[:, None, None]
    assert np.max(np.abs(job.error_sinogram() - e_ref)) <= 2e-5 * np.max(np.abs(e_ref))
    e, w, sel = job.error_sinogram(), job.weights(), job.selector()
    rule = (np.abs(e * np.sqrt(w)) < 3.0).astype(np.uint8)
    assert np.mean(sel != rule) < 1e-3
    for j in jobs:
        j.close()


def test_bf_graph_sweeps_equal_per_step(gpu_ctx):
    """Bright-field sweeps replayed from a HIP graph (kernel_mode 3) end bit for bit as the
    per-step launches: e, weights with the selector and the volume, also when the Bragg
    threshold changes (the descriptor changes and the graph is captured again) and
    across offset and variance updates."""
    g, ang, vol, counts = bf_problem(bragg_frac=0.05)
    offs = np.full(len(ang), -np.log(2000.0))
    geom = capi.geometry_init(g.N_r, g.N_t, ang, g.LengthZ / 2)
    kw = dict(BF, joint_estimation=1, noise_model=1, seed=9)
    a, b = (capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs,
                     kernel_mode=m, **kw) for m in (0, 3))
    for outer in range(3):
        for j in (a, b):
            j.set_bragg(5e10 if outer == 0 else 3.0)
        for it in range(3):
            a.sweep(outer, it)
            b.sweep(outer, it)
            np.testing.assert_array_equal(a.volume(), b.volume())
            np.testing.assert_array_equal(a.error_sinogram(), b.error_sinogram())
            np.testing.assert_array_equal(a.weights(), b.weights())
        a.nuisance_update()
        b.nuisance_update()
    for u, v in zip(a.nuisance(), b.nuisance()):
        np.testing.assert_array_equal(u, v)
    a.close()
    b.close()


def test_bf_four_rows_per_slice_matches_oracle(gpu_ctx):
    """Bright field at four detector rows per slice (level 0 of three resolution levels,
    the reference's default for its bright-field reconstruction): setup against the
    oracle (e, weights, cost), then the cost never rises over sweeps and e stays
    y - A x - mu for the GPU's own volume."""
    _, ang, _, counts = bf_problem(n_r=64, n_t=16, bragg_frac=0.05)
    n_r, n_t = counts.shape[1], counts.shape[2]
    g = O.geometry(n_r, n_t, ang, thickness=16, delta_xz=4.0, interpolate_factor=4.0)
    offs = np.full(len(ang), -np.log(2000.0))
    prm = dict(BF, zero_skipping=0)
    eng = O.Engine(g, ang, counts, O.params(**prm), offsets=offs)
    geom = capi.geometry_init(n_r, n_t, ang, g.LengthZ / 2, delta_xz_in=4.0,
                              interpolate_factor=4.0)
    assert (geom.N_x, geom.N_y, geom.N_z) == (g.N_x, g.N_y, g.N_z)
    assert geom.N_y * 4 == n_t
    job = capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs, **prm)
    ref_e = eng.e()
    assert np.max(np.abs(job.error_sinogram() - ref_e)) <= 1e-5 * np.max(np.abs(ref_e))
    np.testing.assert_allclose(job.weights(), eng.w(), rtol=1e-6)
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)
    costs = [job.cost()]
    for it in range(12):
        job.sweep(0, it)
        costs.append(job.cost())
    assert all(b <= a * (1 + 1e-6) + 1e-9 for a, b in zip(costs, costs[1:])), costs
    y = -np.log(counts)
    e_ref = y - eng.project(job.volume().astype(np.float64)) - offs[:, None, None]
    assert np.max(np.abs(job.error_sinogram() - e_ref)) <= 2e-5 * np.max(np.abs(e_ref))


def test_bf_two_rows_per_slice_matches_oracle(gpu_ctx):
    """Voxels twice the detector pitch (a coarse level of the multi-resolution
    driver): every slice spans two detector rows.  Setup against the oracle, then
    sweeps: the cost never rises, e stays y - A x - mu for the GPU's own volume, and
    after 60 sweeps the volume and cost lie within 3 of the oracle's own scatters over
    update orders (seeds) of their means: the Bragg rejector makes the data term
    non-convex, so orders settle slightly apart."""
    _, ang, _, counts = bf_problem(n_t=8, bragg_frac=0.05)
    n_r, n_t = counts.shape[1], counts.shape[2]
    g = O.geometry(n_r, n_t, ang, thickness=8, delta_xz=2.0, interpolate_factor=2.0)
    offs = np.full(len(ang), -np.log(2000.0))
    prm = dict(BF, zero_skipping=0)
    eng = O.Engine(g, ang, counts, O.params(**prm), offsets=offs)
    geom = capi.geometry_init(n_r, n_t, ang, g.LengthZ / 2, delta_xz_in=2.0,
                              interpolate_factor=2.0)
    assert (geom.N_x, geom.N_y, geom.N_z) == (g.N_x, g.N_y, g.N_z)
    assert geom.N_y * 2 == n_t
    job = capi.Job(gpu_ctx, geom, ang, counts.astype(np.float32), init_offset=offs, **prm)
    ref_e = eng.e()
    assert np.max(np.abs(job.error_sinogram() - ref_e)) <= 1e-5 * np.max(np.abs(ref_e))
    np.testing.assert_allclose(job.weights(), eng.w(), rtol=1e-6)
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)
    costs = [job.cost()]
    for it in range(60):
        eng.sweep(0, it)
        job.sweep(0, it)
        if it < 8:
            costs.append(job.cost())
    assert all(b <= a * (1 + 1e-6) + 1e-9 for a, b in zip(costs, costs[1:])), costs
    y = -np.log(counts)
    e_ref = y - eng.project(job.volume().astype(np.float64)) - offs[:, None, None]
    assert np.max(np.abs(job.error_sinogram() - e_ref)) <= 2e-5 * np.max(np.abs(e_ref))
    runs = [(eng.x(), eng.cost())]
    for seed in (1, 2, 3):
        e2 = O.Engine(g, ang, counts, O.params(**dict(prm, seed=seed)), offsets=offs)
        for it in range(60):
            e2.sweep(0, it)
        runs.append((e2.x(), e2.cost()))

    def rms(a):
        return np.sqrt(np.mean(a ** 2))

    xbar = np.mean([r[0] for r in runs], axis=0)
    spread = max(rms(r[0] - xbar) for r in runs)
    xg = job.volume().astype(np.float64)
    assert rms(xg - xbar) <= max(3.0 * spread, 1e-2 * rms(xbar)), (rms(xg - xbar), spread)
    cbar = np.mean([r[1] for r in runs])
    cspread = max(abs(r[1] - cbar) for r in runs)
    assert abs(job.cost() - cbar) <= max(3.0 * cspread, 1e-5 * abs(cbar)), (job.cost(), cbar, cspread)
