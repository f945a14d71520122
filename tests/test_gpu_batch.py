"""GPU: the batched step kernel (one job and many jobs) against the per-step kernel.

Lines of one step never share a detector bin, so a job's result depends only on
its own schedule: the same job alone, in the batched kernel, and in a mix of
other jobs must give bitwise-identical volumes.
"""
import numpy as np
import pytest

from mbirtem import capi
from problems import small_problem

pytestmark = pytest.mark.gpu


def _job(ctx, g, ang, y, **kw):
    geom = capi.geometry_init(g.N_r, g.N_t, ang, thickness=g.LengthZ / 2)
    return capi.Job(ctx, geom, ang, y, **kw)


def test_batched_kernel_equals_per_step(gpu_ctx):
    g, ang, vol, y = small_problem()
    a = _job(gpu_ctx, g, ang, y, kernel_mode=0, seed=3)
    b = _job(gpu_ctx, g, ang, y, kernel_mode=1, seed=3)
    for it in range(6):
        _, sa = a.sweep(0, it)
        _, sb = b.sweep(0, it)
        assert sa.updated > 0
        assert sa.updated == sb.updated and sa.zero_skipped == sb.zero_skipped
    np.testing.assert_array_equal(a.volume(), b.volume())
    np.testing.assert_array_equal(a.error_sinogram(), b.error_sinogram())


@pytest.mark.parametrize("n_t,y_slabs", [(8, 1), (16, 2), (100, 1)])
def test_graph_sweeps_equal_per_step(gpu_ctx, n_t, y_slabs):
    """Sweeps replayed from a HIP graph (kernel_mode 3) give the per-step launches'
    result bit for bit, replay after replay, also across nuisance updates (which
    change the job descriptor and make the graph be captured again); whole lines, y-slab
    units (the graph's step slots carry each unit's y-range) and lines of two slices
    per lane."""
    g, ang, vol, y = small_problem(n_t=n_t)
    kw = dict(seed=3, joint_estimation=1, noise_model=1, y_slabs=y_slabs)
    a = _job(gpu_ctx, g, ang, y, kernel_mode=0, **kw)
    b = _job(gpu_ctx, g, ang, y, kernel_mode=3, **kw)
    for outer in range(3):
        for it in range(4):
            _, sa = a.sweep(outer, it)
            _, sb = b.sweep(outer, it)
            assert (sa.updated, sa.zero_skipped, sa.status) == (sb.updated, sb.zero_skipped,
                                                                sb.status)
            assert sa.sum_abs_update == sb.sum_abs_update
            np.testing.assert_array_equal(a.volume(), b.volume())
            np.testing.assert_array_equal(a.error_sinogram(), b.error_sinogram())
        a.nuisance_update()
        b.nuisance_update()
    for u, v in zip(a.nuisance(), b.nuisance()):
        np.testing.assert_array_equal(u, v)


def test_batch_result_same_alone_and_in_mix(gpu_ctx):
    probs = [small_problem(n_r=32, n_t=8, seed=5),
             small_problem(n_r=48, n_t=16, n_views=41, thickness=12, seed=9),
             small_problem(n_r=24, n_t=4, n_views=21, max_deg=72.0, thickness=6, seed=11)]
    alone = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=1, seed=20 + i) for i, p in enumerate(probs)]
    mixed = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=1, seed=20 + i) for i, p in enumerate(probs)]
    for it in range(5):
        for j in alone:
            j.sweep(0, it)
        st = capi.batch_sweep(gpu_ctx, mixed, outer=[0] * 3, inner=[it] * 3)
        assert all(s.updated + s.zero_skipped > 0 for s in st)
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a.volume(), m.volume())
        np.testing.assert_array_equal(a.error_sinogram(), m.error_sinogram())


def test_batch_line_lengths_same_alone_and_in_mix(gpu_ctx):
    """Lines of 1, 13 and 100 slices (100: two per lane) and of 200 slices cut into y
    slabs by the automatic choice: each job alone on its own step kernel and all four
    in one batched sweep (three kernel shapes per launch) give the same bits."""
    shapes = [(1, 1), (13, 1), (100, 1), (200, 0)]  # (slices, y slabs; 0: automatic)
    probs = [small_problem(n_r=24, n_t=nt, n_views=21, thickness=8, seed=60 + i)
             for i, (nt, _) in enumerate(shapes)]
    alone = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=0, y_slabs=ys, seed=70 + i)
             for i, (p, (_, ys)) in enumerate(zip(probs, shapes))]
    mixed = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=1, y_slabs=ys, seed=70 + i)
             for i, (p, (_, ys)) in enumerate(zip(probs, shapes))]
    for it in range(3):
        for j in alone:
            j.sweep(0, it)
        st = capi.batch_sweep(gpu_ctx, mixed, outer=[0] * 4, inner=[it] * 4)
        assert all(s.updated + s.zero_skipped > 0 for s in st)
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a.volume(), m.volume())
        np.testing.assert_array_equal(a.error_sinogram(), m.error_sinogram())
        a.close()
        m.close()


def test_two_rows_per_slice_batched_and_served_equal_alone(gpu_ctx):
    """Voxels two detector pitches wide (a coarse level: two detector rows per slice, the
    batched kernels' NV = 2 forms) beside a job of one row per slice: each job alone on
    its own step kernel, in one batched sweep and in the request server end bit for bit
    alike."""
    specs = []
    for i, (n_t, ratio) in enumerate([(16, 2.0), (12, 2.0), (8, 1.0)]):
        g, ang, vol, y = small_problem(n_r=32, n_t=n_t, n_views=31, thickness=8, seed=90 + i)
        geom = capi.geometry_init(32, n_t, ang, g.LengthZ / 2, delta_xz_in=ratio,
                                  interpolate_factor=ratio)
        specs.append((geom, ang, y, dict(seed=30 + i, stop_threshold=0.0, num_inner=4)))
    assert specs[0][0].N_y * 2 == 16
    alone = [capi.Job(gpu_ctx, *sp[:3], kernel_mode=0, **sp[3]) for sp in specs]
    mixed = [capi.Job(gpu_ctx, *sp[:3], kernel_mode=1, **sp[3]) for sp in specs]
    for it in range(4):
        for j in alone:
            j.sweep(0, it)
        capi.batch_sweep(gpu_ctx, mixed, outer=[0] * 3, inner=[it] * 3)
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a.volume(), m.volume())
        np.testing.assert_array_equal(a.error_sinogram(), m.error_sinogram())
    for j in mixed:
        j.close()
    srv = capi.Server(gpu_ctx, max_jobs=4)
    tickets = [srv.submit(*sp[:3], **sp[3]) for sp in specs]
    for t, a in zip(tickets, alone):
        rep = srv.wait(t, timeout=120.0)
        assert rep.state == capi.JOB_DONE and rep.sweeps == 4, (rep.state, rep.error, rep.sweeps)
        np.testing.assert_array_equal(srv.job(t).volume(), a.volume())
        np.testing.assert_array_equal(srv.job(t).error_sinogram(), a.error_sinogram())
        srv.release(t)
    srv.close()
    for j in alone:
        j.close()


def test_autotune_keeps_results(gpu_ctx):
    """The autotuner times both forms of the batched line kernels and keeps the faster
    per shape and launch-size bucket; since the forms compute identical values, jobs
    swept during and after tuning end bit for bit where untuned jobs end."""
    probs = [small_problem(n_r=32, n_t=8, seed=5),
             small_problem(n_r=48, n_t=16, n_views=41, thickness=12, seed=9)]
    tuned = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=1, seed=40 + i) for i, p in enumerate(probs)]
    plain = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=1, seed=40 + i) for i, p in enumerate(probs)]
    saved = capi.get_tuning(gpu_ctx)
    try:
        rep = capi.autotune(gpu_ctx, tuned, inner=[0, 0], sweeps=1)
        measured = 0
        for a in range(2):
            for b in range(2):
                for k in range(5):
                    n0, n1 = rep.launches[a][b][k][0], rep.launches[a][b][k][1]
                    if n0 and n1:
                        measured += 1
                        u0, u1 = rep.us_per_wg[a][b][k][0], rep.us_per_wg[a][b][k][1]
                        assert u0 > 0 and u1 > 0
                        assert rep.prefetch[a][b][k] == (1 if u1 <= u0 else 0)
        assert measured >= 1
        assert (capi.get_tuning(gpu_ctx).table() == rep.table()).all()
        for _ in range(2):  # as the two tuning sweeps (first-sweep flags: inner 0)
            capi.batch_sweep(gpu_ctx, plain, inner=[0, 0])
        for it in range(2, 5):
            capi.batch_sweep(gpu_ctx, tuned, inner=[it, it])
            capi.batch_sweep(gpu_ctx, plain, inner=[it, it])
        for a, b in zip(tuned, plain):
            np.testing.assert_array_equal(a.volume(), b.volume())
            np.testing.assert_array_equal(a.error_sinogram(), b.error_sinogram())
    finally:
        capi.set_tuning(gpu_ctx, saved)


def test_batch_cost_monotone(gpu_ctx):
    probs = [small_problem(seed=s) for s in (31, 32, 33, 34)]
    jobs = [_job(gpu_ctx, *p[:2], p[3], kernel_mode=1, seed=s) for s, p in enumerate(probs)]
    last = [j.cost() for j in jobs]
    for it in range(6):
        capi.batch_sweep(gpu_ctx, jobs, inner=[it] * len(jobs))
        now = [j.cost() for j in jobs]
        for a, b in zip(last, now):
            assert b <= a * (1 + 1e-6)
        last = now


@pytest.mark.skip(reason="the persistent one-launch sweep (kernel_mode 2) did not finish on "
                         "the GPU twice; the path is switched off (mbir_job_create accepts the "
                         "mode but sweeps return -3) until the cause of the unfinished step "
                         "waits is shown")
def test_persistent_sweep_equals_per_step(gpu_ctx):
    """kernel_mode 2 runs each sweep in one persistent launch: the same steps in the
    same order, so the same bits as one launch per step."""
    for prob in (dict(), dict(n_r=64, n_t=16, n_views=45, max_deg=70.0, thickness=16, seed=8)):
        g, ang, vol, y = small_problem(**prob)
        a = _job(gpu_ctx, g, ang, y, kernel_mode=0, seed=4)
        b = _job(gpu_ctx, g, ang, y, kernel_mode=2, seed=4)
        for it in range(6):
            _, sa = a.sweep(0, it)
            _, sb = b.sweep(0, it)
            assert sa.updated > 0
            assert (sa.updated, sa.zero_skipped, sa.status) == (sb.updated, sb.zero_skipped, sb.status)
        np.testing.assert_array_equal(a.volume(), b.volume())
        np.testing.assert_array_equal(a.error_sinogram(), b.error_sinogram())


def test_persistent_sweep_is_refused(gpu_ctx):
    g, ang, vol, y = small_problem()
    b = _job(gpu_ctx, g, ang, y, kernel_mode=2, seed=4)
    with pytest.raises(capi.MbirError) as ei:
        b.sweep(0, 0)
    assert ei.value.code == -3
