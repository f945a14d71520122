"""GPU: the forward projection without allocation (mbir_job_forward_project_noalloc)
against mbir_job_forward_project, through the C-ABI: two calls in a row on one GPU (the
second with a smaller job and volume, so the shared scratch holds the first call's
larger contents), with the job's own volume and with an uploaded one, and the refusal of
a job beyond the scratch's stated maximum."""
import numpy as np
import pytest

from mbirtem import capi
from mbirtem import phantom as P
from problems import small_problem

pytestmark = pytest.mark.gpu


def _noalloc(job, volume=None):
    out = np.zeros(job.sshape)
    v = None if volume is None else np.ascontiguousarray(volume, np.float32)
    rc = capi.lib().mbir_job_forward_project_noalloc(job.h, capi._fptr(v), capi._dptr(out))
    return rc, out


def _job(ctx, n_r, n_t, n_views, thickness, seed):
    g, ang, vol, y = small_problem(n_r=n_r, n_t=n_t, n_views=n_views, thickness=thickness,
                                   seed=seed)
    geom = capi.geometry_init(g.N_r, g.N_t, ang, thickness=g.LengthZ / 2)
    return capi.Job(ctx, geom, ang, y, default_recon=0.05), vol


def test_noalloc_projection_equals_the_allocating_one(gpu_ctx):
    big, vb = _job(gpu_ctx, 96, 16, 45, 24, seed=31)
    small, vs = _job(gpu_ctx, 32, 8, 21, 8, seed=32)
    for job, vol in ((big, vb), (small, vs)):  # the second call is smaller
        for v in (None, vol):
            ref = job.forward_project(v)
            rc, got = _noalloc(job, v)
            assert rc == 0
            # FP64 sums of the same terms (atomic order may differ): 1e-12 of the peak
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()
            assert np.abs(ref).max() > 0
    big.close()
    small.close()


def test_noalloc_refuses_jobs_beyond_the_scratch(gpu_ctx):
    ang = P.tilt_series(65, 60.0)
    geom = capi.geometry_init(1024, 64, ang, thickness=4.0)
    job = capi.Job(gpu_ctx, geom, ang, np.zeros((65, 1024, 64), np.float32))
    assert 65 * 1024 * 64 > 4 << 20
    rc, _ = _noalloc(job)
    assert rc == -1
    job.close()
