"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure only: the checker for the HIP path, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_U8 = C.POINTER(C.c_uint8)


class Geom(C.Structure):
    _fields_ = [
        ("N_theta", C.c_int32), ("N_r", C.c_int32), ("N_t", C.c_int32),
        ("N_x", C.c_int32), ("N_y", C.c_int32), ("N_z", C.c_int32),
        ("delta_r", C.c_double), ("delta_t", C.c_double),
        ("delta_xz", C.c_double), ("delta_xy", C.c_double),
        ("R0", C.c_double), ("RMax", C.c_double), ("T0", C.c_double), ("TMax", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("LengthX", C.c_double), ("LengthY", C.c_double), ("LengthZ", C.c_double),
    ]


class QG(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("MRF_P", "SIGMA_X_P", "MRF_Q", "MRF_C", "MRF_ALPHA", "SIGMA_X_P_Q", "SIGMA_X_Q")]


class Params(C.Structure):
    _fields_ = [
        ("sigma_x", C.c_double), ("mrf_q", C.c_double), ("mrf_c", C.c_double),
        ("target_gain", C.c_double), ("stop_threshold", C.c_double),
        ("num_outer", C.c_int32), ("num_inner", C.c_int32),
        ("joint_estimation", C.c_int32), ("noise_model", C.c_int32),
        ("zero_skipping", C.c_int32), ("compute_cost", C.c_int32), ("seed", C.c_uint32),
        ("bf_mode", C.c_int32), ("bragg_T", C.c_double), ("bragg_delta", C.c_double),
        ("bf_offset", C.c_double), ("estimate_prior", C.c_int32), ("bf_nhicd", C.c_int32),
    ]


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


def load():
    if not os.path.exists(LIB_PATH):
        _build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_geometry_init": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_int32, _D,
                                        C.POINTER(Geom)]),
        "orc_voxel_profile": (None, [_D, C.c_int32, C.c_double, _D]),
        "orc_sincos": (None, [_D, C.c_int32, _D, _D]),
        "orc_beam_profile": (None, [C.c_double, _D]),
        "orc_detector_response": (None, [C.POINTER(Geom), _D, _D, C.c_double, _D]),
        "orc_init_ht": (None, [C.POINTER(Geom), C.c_double, _D]),
        "orc_vlr": (C.c_int32, [C.POINTER(Geom), _D, C.c_double, C.c_int32, _I, _D, C.c_int32]),
        "orc_acol": (C.c_int32, [C.POINTER(Geom), _D, _D, _D, C.c_double, C.c_int32, C.c_int32,
                                 _U, _D, C.c_int32]),
        "orc_offset_r": (C.c_double, [C.POINTER(Geom)]),
        "orc_offset_t": (C.c_double, [C.POINTER(Geom)]),
        "orc_qggmrf_init": (None, [C.c_double, C.c_double, C.c_double, C.POINTER(QG)]),
        "orc_qggmrf_value": (C.c_double, [C.c_double, C.POINTER(QG)]),
        "orc_qggmrf_derivative": (C.c_double, [C.c_double, C.POINTER(QG)]),
        "orc_qggmrf_second_derivative_at_zero": (C.c_double, [C.POINTER(QG)]),
        "orc_qggmrf_functional_substitution": (C.c_double, [C.c_double, C.c_double, C.c_double,
                                                             _U8, _D, _D, C.c_double, C.c_double,
                                                             C.POINTER(QG)]),
        "orc_filter": (None, [_D]),
        "orc_engine_create": (C.c_void_p, [C.POINTER(Geom), _D, _D, _D, _D, _D, _D, C.c_double,
                                           C.POINTER(Params)]),
        "orc_engine_destroy": (None, [C.c_void_p]),
        "orc_engine_sweep": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _D]),
        "orc_engine_sweep_threads": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _D]),
        "orc_engine_sweep_slab": (None, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _D]),
        "orc_engine_joint_estimation": (C.c_int32, [C.c_void_p]),
        "orc_engine_update_units": (None, [C.c_void_p, _I, C.c_int32, C.c_int32, C.c_int32, _D]),
        "orc_engine_update_weights": (None, [C.c_void_p]),
        "orc_engine_estimate_sigma_x": (C.c_double, [C.c_void_p, _D]),
        "orc_engine_get_sigma_x": (C.c_double, [C.c_void_p]),
        "orc_nhicd_filter": (None, [_D, C.c_int32, C.c_int32, _D]),
        "orc_nhicd_threshold": (C.c_double, [_D, C.c_int32]),
        "orc_engine_nhicd_log": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32, _I,
                                             C.c_int64]),
        "orc_engine_cost": (C.c_double, [C.c_void_p]),
        "orc_engine_cost_terms": (C.c_double, [C.c_void_p, _D]),
        "orc_engine_run": (C.c_int32, [C.c_void_p, _D, C.c_int32, _I, _I]),
        "orc_engine_get_x": (None, [C.c_void_p, _D]),
        "orc_engine_set_x": (None, [C.c_void_p, _D]),
        "orc_engine_set_e": (None, [C.c_void_p, _D]),
        "orc_engine_set_bf_sinogram": (None, [C.c_void_p, _D]),
        "orc_engine_get_e": (None, [C.c_void_p, _D]),
        "orc_engine_get_w": (None, [C.c_void_p, _D]),
        "orc_engine_get_yest": (None, [C.c_void_p, _D]),
        "orc_engine_get_nuisance": (None, [C.c_void_p, _D, _D, _D]),
        "orc_engine_get_H": (None, [C.c_void_p, _D]),
        "orc_engine_get_profile": (None, [C.c_void_p, _D]),
        "orc_engine_nnz": (C.c_int64, [C.c_void_p]),
        "orc_engine_get_col": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _U, _D, C.c_int32]),
        "orc_engine_project": (None, [C.c_void_p, _D, _D]),
        "orc_engine_theta": (None, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _D]),
        "orc_engine_get_sel": (None, [C.c_void_p, C.c_void_p]),
        "orc_engine_get_bragg": (C.c_double, [C.c_void_p]),
        "orc_engine_set_bragg": (None, [C.c_void_p, C.c_double]),
        "orc_engine_update_selector": (None, [C.c_void_p]),
        "orc_bf_beta": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def dp(a):
    return a.ctypes.data_as(_D)


def geometry(N_r, N_t, angles_deg, thickness, delta=1.0, extend_object=0, delta_xz=1.0,
             interpolate_factor=1.0):
    g = Geom()
    ang = np.ascontiguousarray(angles_deg, dtype=np.float64)
    rc = lib().orc_geometry_init(N_r, N_t, len(ang), delta, delta, thickness, delta_xz,
                                 interpolate_factor, extend_object, dp(ang), C.byref(g))
    assert rc == 0
    return g


def params(sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3, target_gain=1.0, stop_threshold=1e-3,
           num_outer=1, num_inner=20, joint_estimation=0, noise_model=0, zero_skipping=1,
           compute_cost=1, seed=12345, bf_mode=0, bragg_T=3.0, bragg_delta=0.5, bf_offset=0.0,
           estimate_prior=0, bf_nhicd=0):
    return Params(sigma_x, mrf_q, mrf_c, target_gain, stop_threshold, num_outer, num_inner,
                  joint_estimation, noise_model, zero_skipping, compute_cost, seed, bf_mode,
                  bragg_T, bragg_delta, bf_offset, estimate_prior, bf_nhicd)


class Engine:
    """One oracle engine: the state of HAADF_ReconstructionEngine::execute()."""

    def __init__(self, g, angles_deg, counts, prm, gains=None, offsets=None, variances=None,
                 x0=None, default_recon=0.0):
        self.g = g
        nt = g.N_theta
        self.angles = np.ascontiguousarray(angles_deg, np.float64)
        self.counts = np.ascontiguousarray(counts, np.float64).reshape(nt, g.N_r, g.N_t)
        self.gains = np.ascontiguousarray(np.ones(nt) if gains is None else gains, np.float64)
        self.offsets = np.ascontiguousarray(np.zeros(nt) if offsets is None else offsets, np.float64)
        self.variances = np.ascontiguousarray(np.ones(nt) if variances is None else variances,
                                              np.float64)
        self.prm = prm
        xin = None
        if x0 is not None:
            self._x0 = np.ascontiguousarray(x0, np.float64)
            xin = dp(self._x0)
        self.h = lib().orc_engine_create(C.byref(g), dp(self.angles), dp(self.counts),
                                         dp(self.gains), dp(self.offsets), dp(self.variances),
                                         xin, default_recon, C.byref(prm))
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_engine_destroy(self.h)
            self.h = None

    @property
    def vshape(self):
        return (self.g.N_z, self.g.N_x, self.g.N_y)

    @property
    def sshape(self):
        return (self.g.N_theta, self.g.N_r, self.g.N_t)

    def sweep(self, outer=0, inner=0, threads=0):
        st = np.zeros(4)
        if threads:
            s = lib().orc_engine_sweep_threads(self.h, outer, inner, threads, dp(st))
        else:
            s = lib().orc_engine_sweep(self.h, outer, inner, dp(st))
        return s, st

    def update_units(self, units, outer=0, inner=0):
        """Voxel updates in the order of units [(line, y0, ny), ...]."""
        u = np.ascontiguousarray(units, np.int32).reshape(-1, 3)
        st = np.zeros(4)
        lib().orc_engine_update_units(self.h, u.ctypes.data_as(_I), len(u), outer, inner, dp(st))
        return st

    def sweep_slab(self, outer, inner, y0, y1):
        """The y-slab [y0, y1) of a sweep on one thread (not counted as a sweep)."""
        st = np.zeros(4)
        lib().orc_engine_sweep_slab(self.h, outer, inner, y0, y1, dp(st))
        return st

    def joint_estimation(self):
        return lib().orc_engine_joint_estimation(self.h)

    def update_weights(self):
        lib().orc_engine_update_weights(self.h)

    def estimate_sigma_x(self):
        """estimateSigmaX + updatePriorModel: (estimate, new SigmaX)."""
        sx = C.c_double(0)
        est = lib().orc_engine_estimate_sigma_x(self.h, C.byref(sx))
        return est, sx.value

    def sigma_x(self):
        return lib().orc_engine_get_sigma_x(self.h)

    def nhicd_lists(self):
        """The lines of every NHICD call of the last run, one array per call."""
        n = lib().orc_engine_nhicd_log(self.h, (C.c_int64 * 1)(), 0, None, 0)
        ptr = (C.c_int64 * (n + 1))()
        lib().orc_engine_nhicd_log(self.h, ptr, n, None, 0)
        total = ptr[n] if n else 0
        lines = np.zeros(max(total, 1), np.int32)
        lib().orc_engine_nhicd_log(self.h, ptr, n, lines.ctypes.data_as(_I), total)
        return [lines[ptr[c]:ptr[c + 1]].copy() for c in range(n)]

    def cost(self, terms=False):
        t = np.zeros(3)
        c = lib().orc_engine_cost_terms(self.h, dp(t))
        return (c, t) if terms else c

    def run(self, cap=4096):
        costs = np.zeros(cap)
        nc = C.c_int32(0)
        ns = C.c_int32(0)
        lib().orc_engine_run(self.h, dp(costs), cap, C.byref(nc), C.byref(ns))
        return costs[: nc.value], ns.value

    def _get(self, fn, shape):
        a = np.zeros(shape)
        fn(self.h, dp(a))
        return a

    def x(self):
        return self._get(lib().orc_engine_get_x, self.vshape)

    def set_x(self, x):
        x = np.ascontiguousarray(x, np.float64)
        lib().orc_engine_set_x(self.h, dp(x))

    def set_e(self, err):
        err = np.ascontiguousarray(err, np.float64)
        assert err.size == int(np.prod(self.sshape))
        lib().orc_engine_set_e(self.h, dp(err))

    def set_bf_sinogram(self, bf_counts):
        """HAADF with a bright-field sinogram: entries scaled by counts + 32768."""
        b = np.ascontiguousarray(bf_counts, np.float64).reshape(-1)
        assert b.size == int(np.prod(self.sshape))
        lib().orc_engine_set_bf_sinogram(self.h, dp(b))

    def e(self):
        return self._get(lib().orc_engine_get_e, self.sshape)

    def w(self):
        return self._get(lib().orc_engine_get_w, self.sshape)

    def yest(self):
        return self._get(lib().orc_engine_get_yest, self.sshape)

    def H(self):
        return self._get(lib().orc_engine_get_H, (self.g.N_theta, 64))

    def profile(self):
        return self._get(lib().orc_engine_get_profile, (self.g.N_theta, 1536))

    def sel(self):
        out = np.zeros(self.sshape, np.uint8)
        lib().orc_engine_get_sel(self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def bragg(self, T=None):
        if T is not None:
            lib().orc_engine_set_bragg(self.h, float(T))
        return lib().orc_engine_get_bragg(self.h)

    def update_selector(self):
        lib().orc_engine_update_selector(self.h)

    def nuisance(self):
        nt = self.g.N_theta
        I0, mu, al = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        lib().orc_engine_get_nuisance(self.h, dp(I0), dp(mu), dp(al))
        return I0, mu, al

    def nnz(self):
        return lib().orc_engine_nnz(self.h)

    def col(self, z, x):
        cap = 9 * self.g.N_theta + 16
        idx = np.zeros(cap, np.uint32)
        val = np.zeros(cap)
        n = lib().orc_engine_get_col(self.h, z, x, idx.ctypes.data_as(_U), dp(val), cap)
        return idx[:n].copy(), val[:n].copy()

    def project(self, x):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(self.sshape)
        lib().orc_engine_project(self.h, dp(x), dp(out))
        return out

    def theta(self, z, x, y):
        t = np.zeros(2)
        lib().orc_engine_theta(self.h, z, x, y, dp(t))
        return t
