"""GPU: entry points of the C-ABI that the larger tests do not reach: the job's
geometry and device footprint, replacing a job's volume, the library version."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from mbirtem import capi
from problems import small_problem

pytestmark = pytest.mark.gpu


def test_job_geometry_footprint_and_volume(gpu_ctx):
    """mbir_job_get_geometry returns the geometry the job was made from;
    mbir_job_device_bytes covers at least the volume and the two sinograms (y, e) and stays
    within the admission bound mbir_problem_device_bytes; mbir_job_set_volume replaces
    the volume and re-derives e = y - A x, so e and the cost equal the oracle's for that
    volume."""
    g, ang, vol, y = small_problem()
    geom = capi.geometry_init(g.N_r, g.N_t, ang, thickness=g.LengthZ / 2)
    job = capi.Job(gpu_ctx, geom, ang, y)
    out = capi.Geometry()
    assert capi.lib().mbir_job_get_geometry(job.h, C.byref(out)) == 0
    for name, _ in capi.Geometry._fields_:
        assert getattr(out, name) == getattr(geom, name), name
    used = job.device_bytes()
    nvox = g.N_z * g.N_x * g.N_y
    nent = len(ang) * g.N_r * g.N_t
    assert used >= 4 * nvox + 2 * 4 * nent, used
    assert used <= capi.problem_device_bytes(geom, ang, y), used
    x = vol.astype(np.float32)
    job.set_volume(x)
    np.testing.assert_array_equal(job.volume(), x)
    eng = O.Engine(g, ang, y, O.params(), x0=x.astype(np.float64))
    ref_e = eng.e()
    assert np.max(np.abs(job.error_sinogram() - ref_e)) <= 1e-5 * np.max(np.abs(ref_e))
    assert job.cost() == pytest.approx(eng.cost(), rel=1e-6)
    job.close()


def test_version(gpu_ctx):
    assert capi.lib().mbir_version() > 0
