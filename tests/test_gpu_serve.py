"""GPU: the request server (continuous batching of many jobs) against each job
run alone with mbir_job_run.

A job's steps touch only its own memory and run the same kernels in the same
order inside the server as alone, so volumes, error sinograms and nuisance
parameters must agree bit for bit, whatever else is resident and whenever the job
arrives.
"""
import time

import numpy as np
import pytest

from mbirtem import capi
from problems import small_problem

pytestmark = pytest.mark.gpu


def _geom(g, ang):
    return capi.geometry_init(g.N_r, g.N_t, ang, thickness=g.LengthZ / 2)


def _cases():
    rng = np.random.default_rng(3)
    out = []
    for i, (nr, nt, nv, deg, th) in enumerate([(32, 8, 31, 60.0, 8), (48, 16, 41, 70.0, 12),
                                                (24, 4, 21, 72.0, 6), (64, 32, 45, 66.0, 16),
                                                (40, 12, 37, 60.0, 10)]):
        joint = i % 2
        gains = rng.uniform(0.9, 1.1, nv) if joint else None
        offs = rng.uniform(-0.03, 0.03, nv) if joint else None
        g, ang, vol, y = small_problem(n_r=nr, n_t=nt, n_views=nv, max_deg=deg, thickness=th,
                                       seed=40 + i, gains=gains, offsets=offs)
        kw = dict(seed=100 + i, stop_threshold=2e-3, num_inner=12 if joint else 25,
                  num_outer=6 if joint else 1, joint_estimation=joint, noise_model=joint)
        out.append((g, ang, y, kw))
    return out


def _alone(ctx, case):
    g, ang, y, kw = case
    job = capi.Job(ctx, _geom(g, ang), ang, y, **kw)
    rs, _ = job.run(cap=0)
    res = (job.volume(), job.error_sinogram(), job.nuisance(), rs.sweeps)
    job.close()
    return res


def _check_same(job, ref, rep):
    vol, err, (I0, mu, al), sweeps = ref
    assert rep.state == capi.JOB_DONE and rep.error == 0
    assert rep.sweeps == sweeps
    np.testing.assert_array_equal(job.volume(), vol)
    np.testing.assert_array_equal(job.error_sinogram(), err)
    I0s, mus, als = job.nuisance()
    np.testing.assert_array_equal(I0s, I0)
    np.testing.assert_array_equal(mus, mu)
    np.testing.assert_array_equal(als, al)


def test_server_results_equal_alone(gpu_ctx):
    cases = _cases()
    refs = [_alone(gpu_ctx, c) for c in cases]
    srv = capi.Server(gpu_ctx, max_jobs=8)
    tickets = []
    for i, (g, ang, y, kw) in enumerate(cases):
        tickets.append(srv.submit(_geom(g, ang), ang, y, **kw))
        if i == 2:
            time.sleep(0.05)  # later arrivals join jobs already running
    for t, ref in zip(tickets, refs):
        rep = srv.wait(t, timeout=60.0)
        assert rep is not None, "job did not finish"
        _check_same(srv.job(t), ref, rep)
        assert rep.t_submit <= rep.t_admit <= rep.t_start <= rep.t_done
        assert rep.voxel_visits > 0
        srv.release(t)
    st = srv.stats()
    assert st.jobs_done == len(cases) and st.jobs_failed == 0 and st.fatal_error == 0
    assert st.peak_resident >= 2  # jobs really ran together
    assert st.entries > st.launches  # launches carried several jobs' steps
    srv.close()


def test_server_two_lanes_results_equal_alone(gpu_ctx):
    """Jobs on the second launch stream (long sweeps; here the first two admitted,
    by a threshold of one step) run beside the others and still end bit for bit as
    alone."""
    cases = _cases()
    refs = [_alone(gpu_ctx, c) for c in cases]
    srv = capi.Server(gpu_ctx, max_jobs=8, long_steps=1, long_jobs=2)
    tickets = [srv.submit(_geom(g, ang), ang, y, **kw) for (g, ang, y, kw) in cases]
    for t, ref in zip(tickets, refs):
        rep = srv.wait(t, timeout=60.0)
        assert rep is not None, "job did not finish"
        _check_same(srv.job(t), ref, rep)
        srv.release(t)
    st = srv.stats()
    assert st.jobs_done == len(cases) and st.jobs_failed == 0 and st.fatal_error == 0
    assert 1 <= st.long_lane_jobs <= len(cases)
    srv.close()
    # one stream for all (long_steps < 0): no job on the second
    srv = capi.Server(gpu_ctx, max_jobs=8, long_steps=-1)
    t = srv.submit(_geom(cases[0][0], cases[0][1]), cases[0][1], cases[0][2], **cases[0][3])
    rep = srv.wait(t, timeout=60.0)
    _check_same(srv.job(t), refs[0], rep)
    srv.release(t)
    assert srv.stats().long_lane_jobs == 0
    srv.close()


def test_server_line_lengths_equal_alone(gpu_ctx):
    """Lines of 1, 13 and 100 slices (100 as whole lines: two per lane) served together
    end bit for bit as each run alone."""
    cases = []
    for i, nt in enumerate((1, 13, 100)):
        g, ang, vol, y = small_problem(n_r=24, n_t=nt, n_views=21, thickness=8, seed=80 + i)
        cases.append((g, ang, y, dict(seed=90 + i, y_slabs=1, stop_threshold=2e-3,
                                      num_inner=15)))
    refs = [_alone(gpu_ctx, c) for c in cases]
    srv = capi.Server(gpu_ctx, max_jobs=4)
    tickets = [srv.submit(_geom(g, ang), ang, y, **kw) for g, ang, y, kw in cases]
    for t, ref in zip(tickets, refs):
        rep = srv.wait(t, timeout=120.0)
        _check_same(srv.job(t), ref, rep)
        srv.release(t)
    srv.close()


def test_server_budget_admission(gpu_ctx):
    cases = _cases()[:4]
    need = [capi.problem_device_bytes(_geom(g, ang), ang, y, **kw) for g, ang, y, kw in cases]
    budget = int(sum(sorted(need)[-2:]) * 1.01)  # room for the two largest at most
    srv = capi.Server(gpu_ctx, max_jobs=8, hbm_budget_bytes=budget)
    tickets = [srv.submit(_geom(g, ang), ang, y, **kw) for g, ang, y, kw in cases]
    done = []
    while len(done) < len(tickets):
        t = srv.wait_any(timeout=60.0)
        assert t is not None, "no job finished"
        rep = srv.wait(t, timeout=0.0)
        assert rep.state == capi.JOB_DONE, rep.error
        done.append(t)
        srv.release(t)  # frees its memory for the next admission
    st = srv.stats()
    assert st.peak_reserved_bytes <= budget
    assert st.jobs_done == len(cases)
    srv.close()


def test_server_rejects_and_cancels(gpu_ctx):
    g, ang, y, kw = _cases()[0]
    geom = _geom(g, ang)
    need = capi.problem_device_bytes(geom, ang, y, **kw)
    srv = capi.Server(gpu_ctx, max_jobs=1, hbm_budget_bytes=need - 1)
    t = srv.submit(geom, ang, y, **kw)
    rep = srv.wait(t, timeout=30.0)
    assert rep.state == capi.JOB_FAILED and rep.error == -4  # larger than the whole budget
    srv.release(t)
    srv.close()

    srv = capi.Server(gpu_ctx, max_jobs=1)
    kw_long = dict(kw, num_inner=400, stop_threshold=0.0)
    t1 = srv.submit(geom, ang, y, **kw_long)
    t2 = srv.submit(geom, ang, y, **kw)  # waits: one slot
    srv.cancel(t2)
    rep2 = srv.wait(t2, timeout=30.0)
    assert rep2.state == capi.JOB_FAILED and rep2.error == capi.MBIR_ERR_CANCELLED
    srv.cancel(t1)
    rep1 = srv.wait(t1, timeout=60.0)
    assert rep1.state == capi.JOB_FAILED and rep1.error == capi.MBIR_ERR_CANCELLED
    assert rep1.sweeps < 400
    srv.close()


def test_server_unsupported_shape_fails_cleanly(gpu_ctx):
    """Units of more than 128 slices are beyond the server's line kernels: a job of 160
    slices fails with -3 as whole lines, and runs (equal to alone) when cut into y slabs
    (the automatic choice for more than 64 slices)."""
    g, ang, vol, y = small_problem(n_r=16, n_t=160, n_views=11, thickness=4, seed=2)
    srv = capi.Server(gpu_ctx, max_jobs=2)
    t = srv.submit(_geom(g, ang), ang, y, num_inner=2, y_slabs=1)
    rep = srv.wait(t, timeout=60.0)
    assert rep.state == capi.JOB_FAILED and rep.error == -3
    kw = dict(num_inner=3, stop_threshold=0.0)
    t = srv.submit(_geom(g, ang), ang, y, **kw)
    rep = srv.wait(t, timeout=60.0)
    ref = _alone(gpu_ctx, (g, ang, y, kw))
    _check_same(srv.job(t), ref, rep)
    srv.release(t)
    srv.close()
