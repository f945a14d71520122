"""ctypes binding of the C-ABI in include/mbir_tem.h (libmbir_tem.so).

This is the Python-side stub a maintainer would bind; it only marshals host
arrays.  There is no CPU fallback: a missing or unloadable library raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libmbir_tem.so")

MBIR_OK = 0
MBIR_ERR_CANCELLED = -999

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)


class MbirError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} failed with status {code}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("verbose", C.c_int32), ("hbm_budget_bytes", C.c_uint64)]


class Geometry(C.Structure):
    _fields_ = [
        ("N_theta", C.c_int32), ("N_r", C.c_int32), ("N_t", C.c_int32),
        ("N_x", C.c_int32), ("N_y", C.c_int32), ("N_z", C.c_int32),
        ("delta_r", C.c_double), ("delta_t", C.c_double),
        ("delta_xz", C.c_double), ("delta_xy", C.c_double),
        ("R0", C.c_double), ("RMax", C.c_double), ("T0", C.c_double), ("TMax", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("LengthX", C.c_double), ("LengthY", C.c_double), ("LengthZ", C.c_double),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("geom", Geometry),
        ("angles_deg", _D), ("counts", _F), ("init_volume", _F), ("default_recon", C.c_double),
        ("init_gain", _D), ("init_offset", _D), ("init_variance", _D),
        ("target_gain", C.c_double), ("default_offset", C.c_double),
        ("default_variance", C.c_double),
        ("sigma_x", C.c_double), ("mrf_q", C.c_double), ("mrf_c", C.c_double),
        ("stop_threshold", C.c_double),
        ("num_outer", C.c_int32), ("num_inner", C.c_int32),
        ("joint_estimation", C.c_int32), ("noise_model", C.c_int32),
        ("zero_skipping", C.c_int32), ("seed", C.c_uint32),
        ("schedule", C.c_int32), ("rows_per_step", C.c_int32), ("kernel_mode", C.c_int32),
        ("bf_mode", C.c_int32), ("bragg_T", C.c_double), ("bragg_delta", C.c_double),
        ("bf_offset", C.c_double), ("bf_nhicd", C.c_int32), ("pad_", C.c_int32),
        ("bf_sinogram", _F), ("estimate_prior", C.c_int32), ("y_slabs", C.c_int32),
    ]


class SweepStats(C.Structure):
    _fields_ = [
        ("sum_abs_update", C.c_double), ("sum_abs_value", C.c_double),
        ("zero_skipped", C.c_int64), ("updated", C.c_int64), ("nnz_updated", C.c_int64),
        ("kernel_ms", C.c_double), ("steps", C.c_int32), ("status", C.c_int32),
    ]


class RunStats(C.Structure):
    _fields_ = [("sweeps", C.c_int32), ("outer_iterations", C.c_int32), ("status", C.c_int32),
                ("seconds", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("max_jobs", C.c_int32), ("setup_threads", C.c_int32),
                ("hbm_budget_bytes", C.c_uint64), ("long_steps", C.c_int32),
                ("long_jobs", C.c_int32)]


class JobReport(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("sweeps", C.c_int32),
                ("outer_iterations", C.c_int32), ("status", C.c_int32), ("pad", C.c_int32),
                ("voxel_visits", C.c_int64), ("updated", C.c_int64), ("device_bytes", C.c_int64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("ticks", C.c_int64), ("launches", C.c_int64), ("entries", C.c_int64),
                ("workgroups", C.c_int64), ("jobs_submitted", C.c_int32),
                ("jobs_done", C.c_int32), ("jobs_failed", C.c_int32),
                ("peak_resident", C.c_int32), ("budget_bytes", C.c_uint64),
                ("reserved_bytes", C.c_uint64), ("peak_reserved_bytes", C.c_uint64),
                ("fatal_error", C.c_int32), ("long_lane_jobs", C.c_int32)]


JOB_QUEUED, JOB_SETUP, JOB_RUNNING, JOB_DONE, JOB_FAILED = range(5)

class MrcInfo(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("mode", C.c_int32),
                ("next", C.c_int32), ("nlabels", C.c_int32), ("fei", C.c_int32),
                ("amin", C.c_float), ("amax", C.c_float), ("amean", C.c_float)]


class TiltSelect(C.Structure):
    _fields_ = [("use_subvolume", C.c_int32), ("x_start", C.c_int32), ("x_end", C.c_int32),
                ("y_start", C.c_int32), ("y_end", C.c_int32), ("z_start", C.c_int32),
                ("z_end", C.c_int32), ("interpolate_factor", C.c_double),
                ("excluded_views", C.POINTER(C.c_int32)), ("n_excluded", C.c_int32),
                ("tilts", _F), ("n_tilts", C.c_int32)]


class TiltDims(C.Structure):
    _fields_ = [("N_theta", C.c_int32), ("N_r", C.c_int32), ("N_t", C.c_int32),
                ("x_start", C.c_int32), ("x_end", C.c_int32), ("y_start", C.c_int32),
                ("y_end", C.c_int32), ("delta_r", C.c_double), ("delta_t", C.c_double)]


_I32P = C.POINTER(C.c_int32)


class Tuning(C.Structure):
    _fields_ = [("prefetch", C.c_int32 * 5 * 2 * 2), ("us_per_wg", C.c_double * 2 * 5 * 2 * 2),
                ("launches", C.c_int32 * 2 * 5 * 2 * 2)]

    def table(self):
        """prefetch as an array [ypl>1][nv>1][bucket]"""
        return np.array([[[self.prefetch[a][b][k] for k in range(5)] for b in range(2)]
                         for a in range(2)])

_SIG = {
    "mbir_version": (C.c_int32, []),
    "mbir_nhicd_filter": (C.c_int32, [_F, C.c_int32, C.c_int32, _D]),
    "mbir_nhicd_threshold": (C.c_int32, [_D, C.c_int32, _D]),
    "mbir_autotune": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                  C.POINTER(Tuning)]),
    "mbir_get_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    "mbir_set_tuning": (C.c_int32, [C.c_void_p, C.POINTER(Tuning)]),
    # estimators (include/mbir_estimate.h)
    "mbir_initial_offsets": (C.c_int32, [_D, C.c_int32, C.c_int32, C.c_int32, _D, _D, _D]),
    "mbir_sigma_x_estimate": (C.c_int32, [_D, C.c_int32, C.c_int32, C.c_int32, _D, C.c_int32,
                                          C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                          C.c_double, C.c_int32, C.c_double, _D]),
    # files (include/mbir_io.h)
    "mbir_mrc_read_info": (C.c_int32, [C.c_char_p, C.POINTER(MrcInfo), _F, _F]),
    "mbir_mrc_read_tilt_series": (C.c_int32, [C.c_char_p, C.POINTER(TiltSelect),
                                              C.POINTER(TiltDims), _D, _D, _I32P]),
    "mbir_mrc_write_volume": (C.c_int32, [C.c_char_p, _F, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_int32, C.c_double]),
    "mbir_original_x_dims": (C.c_int32, [C.c_int32, C.c_double, C.c_double, C.c_int32,
                                         C.c_double, _I32P, _I32P]),
    "mbir_write_reconstructed_sinogram": (C.c_int32, [C.c_char_p, C.c_int32, C.c_int32,
                                                      C.c_int32, _D, _F, _D, _D]),
    "mbir_write_raw_volume": (C.c_int32, [C.c_char_p, _F, C.c_int32, C.c_int32, C.c_int32]),
    "mbir_read_raw_volume": (C.c_int32, [C.c_char_p, _D, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_int32]),
    "mbir_create": (C.c_int32, [C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    "mbir_destroy": (C.c_int32, [C.c_void_p]),
    "mbir_geometry_init": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_int32, _D,
                                       C.POINTER(Geometry)]),
    "mbir_job_create": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(C.c_void_p)]),
    "mbir_job_destroy": (C.c_int32, [C.c_void_p]),
    "mbir_job_sweep": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(SweepStats)]),
    "mbir_job_nuisance_update": (C.c_int32, [C.c_void_p]),
    "mbir_batch_sweep": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32,
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(SweepStats)]),
    "mbir_job_cost": (C.c_int32, [C.c_void_p, _D, _D]),
    "mbir_job_run": (C.c_int32, [C.c_void_p, C.POINTER(RunStats), _D, C.c_int32,
                                 C.POINTER(C.c_int32)]),
    "mbir_job_forward_project": (C.c_int32, [C.c_void_p, _F, _D]),
    "mbir_job_get_column": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _U, _F, C.c_int32]),
    "mbir_job_get_volume": (C.c_int32, [C.c_void_p, _F]),
    "mbir_job_set_volume": (C.c_int32, [C.c_void_p, _F]),
    "mbir_job_upsample_volume": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "mbir_job_get_error_sinogram": (C.c_int32, [C.c_void_p, _F]),
    "mbir_job_get_nuisance": (C.c_int32, [C.c_void_p, _D, _D, _D]),
    "mbir_job_get_geometry": (C.c_int32, [C.c_void_p, C.POINTER(Geometry)]),
    "mbir_job_get_detector_response": (C.c_int32, [C.c_void_p, _D]),
    "mbir_job_cancel": (C.c_int32, [C.c_void_p]),
    "mbir_job_device_bytes": (C.c_int64, [C.c_void_p]),
    "mbir_job_set_timing": (C.c_int32, [C.c_void_p, C.c_int32]),
    "mbir_problem_device_bytes": (C.c_int64, [C.POINTER(Problem)]),
    "mbir_job_get_selector": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "mbir_job_get_weights": (C.c_int32, [C.c_void_p, _F]),
    "mbir_job_set_bragg_threshold": (C.c_int32, [C.c_void_p, C.c_double]),
    "mbir_job_estimate_prior": (C.c_int32, [C.c_void_p, _D, _D]),
    "mbir_job_get_sigma_x": (C.c_int32, [C.c_void_p, _D]),
    "mbir_job_forward_project_noalloc": (C.c_int32, [C.c_void_p, _F, _D]),
    "mbir_job_get_schedule": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32,
                                          C.POINTER(C.c_int32), C.c_int64]),
    "mbir_job_set_nhicd_log": (C.c_int32, [C.c_void_p, C.c_int32]),
    "mbir_job_get_nhicd_log": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32,
                                           C.POINTER(C.c_int32), C.c_int64]),
    "mbir_server_create": (C.c_int32, [C.c_void_p, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "mbir_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(Problem), C.POINTER(C.c_uint64)]),
    "mbir_server_wait": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(JobReport)]),
    "mbir_server_wait_any": (C.c_int32, [C.c_void_p, C.c_double, C.POINTER(C.c_uint64)]),
    "mbir_server_job": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "mbir_server_release": (C.c_int32, [C.c_void_p, C.c_uint64]),
    "mbir_server_cancel": (C.c_int32, [C.c_void_p, C.c_uint64]),
    "mbir_server_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    "mbir_server_destroy": (C.c_int32, [C.c_void_p]),
}

_LIB = None


def lib():
    """Load libmbir_tem.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python openmbir-tem_amd/build.py`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _LIB = h
    return _LIB


def _chk(fn, rc):
    if rc < 0:
        raise MbirError(fn, rc)
    return rc


def _dptr(a):
    return None if a is None else a.ctypes.data_as(_D)


def _fptr(a):
    return None if a is None else a.ctypes.data_as(_F)


def batch_sweep(ctx, jobs, outer=None, inner=None):
    """One ICD sweep of every job, each launch carrying one step of every job;
    returns per-job SweepStats."""
    n = len(jobs)
    hs = (C.c_void_p * n)(*[j.h.value for j in jobs])
    o = (C.c_int32 * n)(*(outer if outer is not None else [0] * n))
    i = (C.c_int32 * n)(*(inner if inner is not None else [0] * n))
    st = (SweepStats * n)()
    _chk("mbir_batch_sweep", lib().mbir_batch_sweep(ctx.h, hs, n, o, i, st))
    return list(st)


def autotune(ctx, jobs, outer=None, inner=None, sweeps=1):
    """Time each form of the batched line kernels on these jobs and keep the faster per
    shape and launch-size bucket in ctx; returns the Tuning report."""
    n = len(jobs)
    hs = (C.c_void_p * n)(*[j.h.value for j in jobs])
    o = (C.c_int32 * n)(*(outer if outer is not None else [0] * n))
    i = (C.c_int32 * n)(*(inner if inner is not None else [0] * n))
    t = Tuning()
    _chk("mbir_autotune", lib().mbir_autotune(ctx.h, hs, n, o, i, sweeps, C.byref(t)))
    return t


def get_tuning(ctx):
    t = Tuning()
    _chk("mbir_get_tuning", lib().mbir_get_tuning(ctx.h, C.byref(t)))
    return t


def set_tuning(ctx, t):
    _chk("mbir_set_tuning", lib().mbir_set_tuning(ctx.h, C.byref(t)))


def geometry_init(N_r, N_t, angles_deg, thickness, delta_r=1.0, delta_t=1.0, delta_xz_in=1.0,
                  interpolate_factor=1.0, extend_object=0):
    ang = np.ascontiguousarray(angles_deg, np.float64)
    g = Geometry()
    _chk("mbir_geometry_init",
         lib().mbir_geometry_init(N_r, N_t, len(ang), delta_r, delta_t, thickness, delta_xz_in,
                                  interpolate_factor, extend_object, _dptr(ang), C.byref(g)))
    return g


class Context:
    def __init__(self, device=0, verbose=0, hbm_budget_bytes=0):
        self.h = C.c_void_p()
        o = Opts(device, verbose, hbm_budget_bytes)
        _chk("mbir_create", lib().mbir_create(C.byref(o), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().mbir_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_problem(geom, angles_deg, counts, *, sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3,
                 target_gain=1.0, default_offset=0.0, default_variance=1.0, stop_threshold=1e-3,
                 num_outer=1, num_inner=20, joint_estimation=0, noise_model=0, zero_skipping=1,
                 seed=12345, schedule=0, rows_per_step=1, kernel_mode=0, init_volume=None,
                 default_recon=0.0, init_gain=None, init_offset=None, init_variance=None,
                 bf_mode=0, bragg_T=3.0, bragg_delta=0.5, bf_offset=0.0, bf_nhicd=0,
                 bf_sinogram=None, estimate_prior=0, y_slabs=0):
    """mbir_problem for the arrays given; returns (Problem, arrays to keep alive)."""
    ang = np.ascontiguousarray(angles_deg, np.float64)
    cnt = np.ascontiguousarray(counts, np.float32)
    assert cnt.size == geom.N_theta * geom.N_r * geom.N_t
    keep = [ang, cnt]
    p = Problem()
    p.geom = geom
    p.angles_deg = _dptr(ang)
    p.counts = _fptr(cnt)
    if init_volume is not None:
        iv = np.ascontiguousarray(init_volume, np.float32)
        keep.append(iv)
        p.init_volume = _fptr(iv)
    p.default_recon = default_recon
    for name, arr in (("init_gain", init_gain), ("init_offset", init_offset),
                      ("init_variance", init_variance)):
        if arr is not None:
            a = np.ascontiguousarray(arr, np.float64)
            keep.append(a)
            setattr(p, name, _dptr(a))
    p.target_gain = target_gain
    p.default_offset = default_offset
    p.default_variance = default_variance
    p.sigma_x, p.mrf_q, p.mrf_c = sigma_x, mrf_q, mrf_c
    p.stop_threshold = stop_threshold
    p.num_outer, p.num_inner = num_outer, num_inner
    p.joint_estimation, p.noise_model, p.zero_skipping = joint_estimation, noise_model, zero_skipping
    p.seed = seed
    p.schedule, p.rows_per_step, p.kernel_mode = schedule, rows_per_step, kernel_mode
    p.bf_mode, p.bragg_T, p.bragg_delta, p.bf_offset = bf_mode, bragg_T, bragg_delta, bf_offset
    p.bf_nhicd = bf_nhicd
    p.estimate_prior = estimate_prior
    p.y_slabs = y_slabs
    if bf_sinogram is not None:
        b = np.ascontiguousarray(bf_sinogram, np.float32)
        assert b.size == cnt.size
        keep.append(b)
        p.bf_sinogram = _fptr(b)
    return p, keep


class Job:
    """One reconstruction job on the device (mbir_job_*)."""

    def __init__(self, ctx, geom, angles_deg, counts, **kw):
        self.ctx = ctx
        self.geom = geom
        self._owned = True
        p, keep = make_problem(geom, angles_deg, counts, **kw)
        self.h = C.c_void_p()
        _chk("mbir_job_create", lib().mbir_job_create(ctx.h, C.byref(p), C.byref(self.h)))
        del keep

    @classmethod
    def borrowed(cls, handle, geom):
        """View of a job owned elsewhere (a server's finished job): close() keeps it."""
        j = cls.__new__(cls)
        j.ctx = None
        j.geom = geom
        j._owned = False
        j.h = C.c_void_p(handle)
        return j

    @property
    def vshape(self):
        g = self.geom
        return (g.N_z, g.N_x, g.N_y)

    @property
    def sshape(self):
        g = self.geom
        return (g.N_theta, g.N_r, g.N_t)

    def close(self):
        if self.h and self._owned:
            lib().mbir_job_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sweep(self, outer=0, inner=0):
        st = SweepStats()
        rc = _chk("mbir_job_sweep", lib().mbir_job_sweep(self.h, outer, inner, C.byref(st)))
        return rc, st

    def nuisance_update(self):
        _chk("mbir_job_nuisance_update", lib().mbir_job_nuisance_update(self.h))

    def estimate_prior(self):
        """sigma_x re-estimation: (the estimate, the new sigma_x)."""
        est, sx = C.c_double(), C.c_double()
        _chk("mbir_job_estimate_prior",
             lib().mbir_job_estimate_prior(self.h, C.byref(est), C.byref(sx)))
        return est.value, sx.value

    def schedule(self):
        """The last sweep's steps: a list of (n, 3) arrays of units (line, y0, ny)."""
        n = _chk("mbir_job_get_schedule", lib().mbir_job_get_schedule(self.h, None, 0, None, 0))
        ptr = (C.c_int32 * (n + 1))()
        lib().mbir_job_get_schedule(self.h, ptr, n, None, 0)
        total = ptr[n] if n else 0
        u = np.zeros((max(total, 1), 3), np.int32)
        lib().mbir_job_get_schedule(self.h, ptr, n, u.ctypes.data_as(C.POINTER(C.c_int32)), total)
        return [u[ptr[k]:ptr[k + 1]].copy() for k in range(n)]

    def set_nhicd_log(self, on=True):
        _chk("mbir_job_set_nhicd_log", lib().mbir_job_set_nhicd_log(self.h, int(on)))

    def nhicd_lists(self):
        """The lines of every NHICD call of the last run, one array per call."""
        n = _chk("mbir_job_get_nhicd_log",
                 lib().mbir_job_get_nhicd_log(self.h, None, 0, None, 0))
        ptr = (C.c_int64 * (n + 1))()
        lib().mbir_job_get_nhicd_log(self.h, ptr, n, None, 0)
        total = ptr[n] if n else 0
        lines = np.zeros(max(total, 1), np.int32)
        lib().mbir_job_get_nhicd_log(self.h, ptr, n,
                                     lines.ctypes.data_as(C.POINTER(C.c_int32)), total)
        return [lines[ptr[c]:ptr[c + 1]].copy() for c in range(n)]

    def sigma_x(self):
        sx = C.c_double()
        _chk("mbir_job_get_sigma_x", lib().mbir_job_get_sigma_x(self.h, C.byref(sx)))
        return sx.value

    def cost(self, terms=False):
        tot = C.c_double()
        t = np.zeros(3)
        _chk("mbir_job_cost", lib().mbir_job_cost(self.h, C.byref(tot), _dptr(t)))
        return (tot.value, t) if terms else tot.value

    def run(self, cap=4096):
        """Run the job; returns the run statistics and up to cap costs (one after the
        setup, one after each sweep and nuisance step).  cap=0 records no costs, so the
        run computes none (the reference computes them only in its COST_CALCULATE
        debugging build, BrightField/BFReconstructionEngine.cpp:806-876)."""
        rs = RunStats()
        costs = np.zeros(max(cap, 0))
        n = C.c_int32()
        ptr = _dptr(costs) if cap > 0 else None
        _chk("mbir_job_run", lib().mbir_job_run(self.h, C.byref(rs), ptr, max(cap, 0), C.byref(n)))
        return rs, costs[: n.value]

    def forward_project(self, volume=None):
        out = np.zeros(self.sshape)
        v = None if volume is None else np.ascontiguousarray(volume, np.float32)
        _chk("mbir_job_forward_project", lib().mbir_job_forward_project(self.h, _fptr(v), _dptr(out)))
        return out

    def column(self, z, x):
        cap = 16 * self.geom.N_theta
        idx = np.zeros(cap, np.uint32)
        val = np.zeros(cap, np.float32)
        n = _chk("mbir_job_get_column",
                 lib().mbir_job_get_column(self.h, z, x, idx.ctypes.data_as(_U), _fptr(val), cap))
        return idx[:n].copy(), val[:n].copy()

    def volume(self):
        out = np.zeros(self.vshape, np.float32)
        _chk("mbir_job_get_volume", lib().mbir_job_get_volume(self.h, _fptr(out)))
        return out

    def set_volume(self, vol):
        v = np.ascontiguousarray(vol, np.float32)
        _chk("mbir_job_set_volume", lib().mbir_job_set_volume(self.h, _fptr(v)))

    def upsample_from(self, coarse):
        """This job's volume from a coarse job's, replicated 2 x 2 x 2 on the device."""
        _chk("mbir_job_upsample_volume", lib().mbir_job_upsample_volume(self.h, coarse.h))

    def error_sinogram(self):
        out = np.zeros(self.sshape, np.float32)
        _chk("mbir_job_get_error_sinogram", lib().mbir_job_get_error_sinogram(self.h, _fptr(out)))
        return out

    def nuisance(self):
        n = self.geom.N_theta
        I0, mu, al = np.zeros(n), np.zeros(n), np.zeros(n)
        _chk("mbir_job_get_nuisance", lib().mbir_job_get_nuisance(self.h, _dptr(I0), _dptr(mu), _dptr(al)))
        return I0, mu, al

    def detector_response(self):
        out = np.zeros((self.geom.N_theta, 64))
        _chk("mbir_job_get_detector_response", lib().mbir_job_get_detector_response(self.h, _dptr(out)))
        return out

    def selector(self):
        out = np.zeros(self.sshape, np.uint8)
        _chk("mbir_job_get_selector", lib().mbir_job_get_selector(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def weights(self):
        out = np.zeros(self.sshape, np.float32)
        _chk("mbir_job_get_weights", lib().mbir_job_get_weights(self.h, _fptr(out)))
        return out

    def set_bragg(self, T):
        _chk("mbir_job_set_bragg_threshold", lib().mbir_job_set_bragg_threshold(self.h, T))

    def cancel(self):
        _chk("mbir_job_cancel", lib().mbir_job_cancel(self.h))

    def device_bytes(self):
        return lib().mbir_job_device_bytes(self.h)

    def set_timing(self, on=True):
        _chk("mbir_job_set_timing", lib().mbir_job_set_timing(self.h, 1 if on else 0))


def problem_device_bytes(geom, angles_deg, counts, **kw):
    p, keep = make_problem(geom, angles_deg, counts, **kw)
    return lib().mbir_problem_device_bytes(C.byref(p))


class Server:
    """Request server (mbir_server_*): jobs run concurrently with continuous batching."""

    def __init__(self, ctx, max_jobs=0, setup_threads=0, hbm_budget_bytes=0, long_steps=0,
                 long_jobs=0):
        self.ctx = ctx
        self.h = C.c_void_p()
        self._geom = {}
        o = ServerOpts(max_jobs, setup_threads, hbm_budget_bytes, long_steps, long_jobs)
        _chk("mbir_server_create", lib().mbir_server_create(ctx.h, C.byref(o), C.byref(self.h)))

    def submit(self, geom, angles_deg, counts, **kw):
        p, keep = make_problem(geom, angles_deg, counts, **kw)
        t = C.c_uint64()
        _chk("mbir_server_submit", lib().mbir_server_submit(self.h, C.byref(p), C.byref(t)))
        del keep
        self._geom[t.value] = geom
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """JobReport once the job has finished, None on timeout."""
        rep = JobReport()
        rc = _chk("mbir_server_wait", lib().mbir_server_wait(self.h, ticket, timeout, C.byref(rep)))
        return None if rc == 1 else rep

    def wait_any(self, timeout=-1.0):
        t = C.c_uint64()
        rc = _chk("mbir_server_wait_any", lib().mbir_server_wait_any(self.h, timeout, C.byref(t)))
        return None if rc == 1 else t.value

    def job(self, ticket):
        """The finished job (results through the Job getters) until release()."""
        h = C.c_void_p()
        _chk("mbir_server_job", lib().mbir_server_job(self.h, ticket, C.byref(h)))
        return Job.borrowed(h.value, self._geom[ticket])

    def release(self, ticket):
        _chk("mbir_server_release", lib().mbir_server_release(self.h, ticket))
        self._geom.pop(ticket, None)

    def cancel(self, ticket):
        _chk("mbir_server_cancel", lib().mbir_server_cancel(self.h, ticket))

    def stats(self):
        st = ServerStats()
        _chk("mbir_server_get_stats", lib().mbir_server_get_stats(self.h, C.byref(st)))
        return st

    def close(self):
        if self.h:
            lib().mbir_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- files (include/mbir_io.h) ----

def mrc_read_info(path):
    """Header fields and, for an FEI-style extended header, per-section tilt and pixel size."""
    info = MrcInfo()
    _chk("mbir_mrc_read_info", lib().mbir_mrc_read_info(path.encode(), C.byref(info), None, None))
    tilts = pix = None
    if info.fei:
        tilts = np.zeros(info.nz, np.float32)
        pix = np.zeros(info.nz, np.float32)
        _chk("mbir_mrc_read_info",
             lib().mbir_mrc_read_info(path.encode(), C.byref(info), _fptr(tilts), _fptr(pix)))
    return info, tilts, pix


def mrc_read_tilt_series(path, use_subvolume=False, x=(0, 0), y=(0, 0), z=(0, 0),
                         interpolate_factor=1.0, excluded=(), tilts=None):
    """(counts [N_theta][N_r][N_t], angles (deg), good views, TiltDims) of an MRC tilt series."""
    ex = (C.c_int32 * max(1, len(excluded)))(*excluded)
    tl = None if tilts is None else np.ascontiguousarray(tilts, np.float32)
    sel = TiltSelect(1 if use_subvolume else 0, x[0], x[1], y[0], y[1], z[0], z[1],
                     interpolate_factor, ex, len(excluded), _fptr(tl),
                     0 if tl is None else len(tl))
    d = TiltDims()
    _chk("mbir_mrc_read_tilt_series",
         lib().mbir_mrc_read_tilt_series(path.encode(), C.byref(sel), C.byref(d), None, None, None))
    counts = np.zeros((d.N_theta, d.N_r, d.N_t))
    ang = np.zeros(d.N_theta)
    good = np.zeros(d.N_theta, np.int32)
    _chk("mbir_mrc_read_tilt_series",
         lib().mbir_mrc_read_tilt_series(path.encode(), C.byref(sel), C.byref(d), _dptr(counts),
                                         _dptr(ang), good.ctypes.data_as(_I32P)))
    return counts, ang, good, d


def mrc_write_volume(path, vol, x0=0, x1=None, length_x=0.0):
    v = np.ascontiguousarray(vol, np.float32)
    nz, nx, ny = v.shape
    _chk("mbir_mrc_write_volume",
         lib().mbir_mrc_write_volume(path.encode(), _fptr(v), nz, nx, ny, x0,
                                     nx if x1 is None else x1, length_x))


def original_x_dims(N_x, x0, delta_xz, N_r, delta_r):
    s, e = C.c_int32(), C.c_int32()
    _chk("mbir_original_x_dims",
         lib().mbir_original_x_dims(N_x, x0, delta_xz, N_r, delta_r, C.byref(s), C.byref(e)))
    return s.value, e.value


def write_reconstructed_sinogram(path, counts, e, I0, mu):
    c = np.ascontiguousarray(counts, np.float64)
    ef = np.ascontiguousarray(e, np.float32)
    i0 = np.ascontiguousarray(I0, np.float64)
    m = np.ascontiguousarray(mu, np.float64)
    nt, nr, ntt = c.shape
    _chk("mbir_write_reconstructed_sinogram",
         lib().mbir_write_reconstructed_sinogram(path.encode(), nt, nr, ntt, _dptr(c), _fptr(ef),
                                                 _dptr(i0), _dptr(m)))


def write_raw_volume(path, vol):
    v = np.ascontiguousarray(vol, np.float32)
    _chk("mbir_write_raw_volume", lib().mbir_write_raw_volume(path.encode(), _fptr(v), *v.shape))


def read_raw_volume(path, shape, interp=0):
    out = np.zeros(shape)
    _chk("mbir_read_raw_volume",
         lib().mbir_read_raw_volume(path.encode(), _dptr(out), *shape, interp))
    return out


# ---- estimators (include/mbir_estimate.h) ----

def initial_offsets(counts, angles_deg):
    """(offsets per view, (a, b)) of the least-squares fit mean[th] ~ a / cos + b."""
    c = np.ascontiguousarray(counts, np.float64)
    ang = np.ascontiguousarray(angles_deg, np.float64)
    out = np.zeros(c.shape[0])
    fit = np.zeros(2)
    _chk("mbir_initial_offsets",
         lib().mbir_initial_offsets(_dptr(c), *c.shape, _dptr(ang), _dptr(out), _dptr(fit)))
    return out, tuple(fit)


def sigma_x_estimate(counts, angles_deg, thickness, target_gain=1.0, x=(-1, -1), y=(-1, -1),
                     bf_offset=None):
    c = np.ascontiguousarray(counts, np.float64)
    ang = np.ascontiguousarray(angles_deg, np.float64)
    s = C.c_double()
    _chk("mbir_sigma_x_estimate",
         lib().mbir_sigma_x_estimate(_dptr(c), *c.shape, _dptr(ang), x[0], x[1], y[0], y[1],
                                     target_gain, thickness, 0 if bf_offset is None else 1,
                                     0.0 if bf_offset is None else bf_offset, C.byref(s)))
    return s.value


def nhicd_filter(mag):
    m = np.ascontiguousarray(mag, np.float32)
    out = np.zeros(m.shape)
    _chk("mbir_nhicd_filter", lib().mbir_nhicd_filter(_fptr(m), m.shape[0], m.shape[1], _dptr(out)))
    return out


def nhicd_threshold(values):
    v = np.ascontiguousarray(values, np.float64).ravel()
    t = C.c_double()
    _chk("mbir_nhicd_threshold", lib().mbir_nhicd_threshold(_dptr(v), v.size, C.byref(t)))
    return t.value
