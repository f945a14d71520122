"""Seeded synthetic HAADF workloads (phantom, tilt series, nuisance, noise).

Workload names follow the sizes they describe.  The measured sinogram is made
with the engine's own forward model (I0 * A x), then per-view gain and offset
and Gaussian noise are applied.
"""
from dataclasses import dataclass

import numpy as np

from . import capi
from . import phantom as P


@dataclass
class Workload:
    name: str
    n_r: int
    n_t: int
    n_views: int
    max_deg: float
    thickness: float
    n_objects: int
    noise: float
    gain_range: tuple
    offset_range: tuple
    seed: int
    joint_estimation: int
    noise_model: int
    stop_threshold: float
    num_outer: int
    num_inner: int


HAADF_128x16x64 = Workload("haadf_128x16x64_61v", 128, 16, 61, 60.0, 32, 0, 0.02, (1.0, 1.0),
                           (0.0, 0.0), 1, 0, 0, 1e-3, 1, 20)
HAADF_512x64x256 = Workload("haadf_512x64x256_91v", 512, 64, 91, 70.0, 128, 20, 0.03, (0.8, 1.2),
                            (-0.05, 0.05), 2, 1, 1, 1e-3, 30, 10)

WORKLOADS = {w.name: w for w in (HAADF_128x16x64, HAADF_512x64x256)}


def make(ctx, wl, seed=None, **job_kw):
    """Geometry, angles, phantom, sinogram and a ready job for workload wl
    (seed overrides the workload's phantom/noise seed)."""
    if seed is not None:
        import dataclasses
        wl = dataclasses.replace(wl, seed=seed)
    ang = P.tilt_series(wl.n_views, wl.max_deg)
    geom = capi.geometry_init(wl.n_r, wl.n_t, ang, wl.thickness)
    shape = (geom.N_z, geom.N_x, geom.N_y)
    if wl.n_objects == 0:
        vol = P.two_spheres(shape, seed=wl.seed)
    else:
        vol = P.random_ellipsoids(shape, n=wl.n_objects, seed=wl.seed)
    rng = np.random.default_rng(wl.seed + 100)
    gains = rng.uniform(*wl.gain_range, wl.n_views)
    offs = rng.uniform(*wl.offset_range, wl.n_views)
    zero = np.zeros((wl.n_views, wl.n_r, wl.n_t), np.float32)
    proj_job = capi.Job(ctx, geom, ang, zero)
    proj = proj_job.forward_project(vol)
    proj_job.close()
    y = P.noisy_sinogram(proj, wl.noise, wl.seed + 1, gains=gains, offsets=offs)
    kw = dict(joint_estimation=wl.joint_estimation, noise_model=wl.noise_model,
              stop_threshold=wl.stop_threshold, num_outer=wl.num_outer, num_inner=wl.num_inner,
              sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3, seed=wl.seed)
    kw.update(job_kw)
    job = capi.Job(ctx, geom, ang, y, **kw)
    return dict(geom=geom, angles=ang, volume=vol, sinogram=y, gains=gains, offsets=offs,
                job=job, job_kw=kw)


# Many independent slab problems of uneven size: N_x (= N_r) in [128, 1024], N_y (= N_t)
# in [8, 64], N_z in [32, 256], 41-141 tilts over +-[60, 72] degrees, sphere
# phantoms, Gaussian noise sigma 0.02, gain 1, offset 0; each reconstructed until the
# stop threshold (relative update 1e-3) is met.  big_frac of the draws are replaced
# by 1024 x 128 x 512 problems (the large jobs of a served request stream).
@dataclass
class Slab:
    name: str
    n_r: int
    n_t: int
    n_views: int
    max_deg: float
    thickness: float
    seed: int


def slab_mix(n, seed=7, big_frac=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if big_frac > 0 and rng.uniform() < big_frac:
            nr, nt, nz, nv = 1024, 128, 512, int(rng.integers(41, 142))
        else:
            nr = int(rng.integers(128, 1025))
            nt = int(rng.integers(8, 65))
            nz = int(rng.integers(32, 257))
            nv = int(rng.integers(41, 142))
        deg = float(rng.uniform(60.0, 72.0))
        out.append(Slab(f"slab{i}_{nr}x{nt}x{nz}_{nv}v", nr, nt, nv, deg, nz / 2.0,
                        int(rng.integers(1, 2 ** 31))))
    return out


SLAB_KW = dict(sigma_x=0.1, mrf_q=1.2, mrf_c=1e-3, stop_threshold=1e-3, num_outer=1,
               num_inner=60, joint_estimation=0, noise_model=0)


def make_slab(ctx, s, noise=0.02):
    """Geometry, angles, phantom and noisy sinogram of slab s (projected on the GPU)."""
    ang = P.tilt_series(s.n_views, s.max_deg)
    geom = capi.geometry_init(s.n_r, s.n_t, ang, s.thickness)
    shape = (geom.N_z, geom.N_x, geom.N_y)
    vol = P.random_spheres(shape, n=12, seed=s.seed)
    zero = np.zeros((s.n_views, s.n_r, s.n_t), np.float32)
    pj = capi.Job(ctx, geom, ang, zero)
    proj = pj.forward_project(vol)
    pj.close()
    y = P.noisy_sinogram(proj, noise, s.seed + 1).astype(np.float32)
    kw = dict(SLAB_KW, seed=s.seed % 100000)
    return dict(slab=s, geom=geom, angles=ang, volume=vol, sinogram=y, job_kw=kw)


# Bright field, 512 x 64 x 256, 121 tilts over +-60 degrees: transmitted counts
# Poisson(dose_k exp(-A x)) with a mean dose of 2000 per pixel (per-view dose
# U[0.8, 1.2] x 2000, estimated as the per-view offset -log(dose)); random
# ellipsoids with attenuation U[0.2, 1] x 0.01 per voxel.
@dataclass
class BFWorkload:
    name: str
    n_r: int
    n_t: int
    n_views: int
    max_deg: float
    thickness: float
    n_objects: int
    dose: float
    dose_range: tuple
    mu_scale: float
    seed: int
    num_outer: int
    num_inner: int
    stop_threshold: float


BF_512x64x256 = BFWorkload("bf_512x64x256_121v", 512, 64, 121, 60.0, 128, 20, 2000.0, (0.8, 1.2),
                           0.01, 3, 30, 10, 1e-3)
BF_128x16x64 = BFWorkload("bf_128x16x64_61v", 128, 16, 61, 60.0, 32, 6, 2000.0, (0.8, 1.2),
                          0.01, 5, 12, 6, 1e-4)
BF_KW = dict(bf_mode=1, mrf_c=0.01, sigma_x=0.002, mrf_q=1.2, bragg_T=3.0, bragg_delta=0.5,
             joint_estimation=1, noise_model=1)


def make_bf(ctx, wl, **job_kw):
    """Geometry, angles, phantom, counts and a ready bright-field job for workload wl."""
    ang = P.tilt_series(wl.n_views, wl.max_deg)
    geom = capi.geometry_init(wl.n_r, wl.n_t, ang, wl.thickness)
    shape = (geom.N_z, geom.N_x, geom.N_y)
    vol = wl.mu_scale * P.random_ellipsoids(shape, n=wl.n_objects, seed=wl.seed)
    zero = np.zeros((wl.n_views, wl.n_r, wl.n_t), np.float32)
    pj = capi.Job(ctx, geom, ang, zero)
    proj = pj.forward_project(vol)
    pj.close()
    rng = np.random.default_rng(wl.seed + 7)
    dose = wl.dose * rng.uniform(*wl.dose_range, wl.n_views)
    counts = rng.poisson(dose[:, None, None] * np.exp(-proj)).astype(np.float32)
    counts = np.maximum(counts, 1.0)
    kw = dict(BF_KW, num_outer=wl.num_outer, num_inner=wl.num_inner,
              stop_threshold=wl.stop_threshold, seed=wl.seed,
              default_offset=-np.log(wl.dose))
    kw.update(job_kw)
    job = capi.Job(ctx, geom, ang, counts, **kw)
    return dict(geom=geom, angles=ang, volume=vol, counts=counts, dose=dose, job=job, job_kw=kw)
