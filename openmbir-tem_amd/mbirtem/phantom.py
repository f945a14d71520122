"""Seeded synthetic phantoms and tilt schedules for the HAADF workloads.

Volumes are float arrays laid out [N_z][N_x][N_y] (y fastest), the layout of the
reconstruction engine's Object (GenericFilters/InitialReconstructionInitializer.cpp:158-160).
"""
import numpy as np


def tilt_series(n_views, max_deg):
    """Evenly spaced tilts over [-max_deg, +max_deg] in degrees."""
    return np.linspace(-max_deg, max_deg, n_views).astype(np.float64)


def _grid(shape):
    nz, nx, ny = shape
    z = np.arange(nz)[:, None, None] + 0.5
    x = np.arange(nx)[None, :, None] + 0.5
    y = np.arange(ny)[None, None, :] + 0.5
    return z, x, y


def ellipsoids(shape, items, dtype=np.float32):
    """Sum of uniform ellipsoids.  items: (cz, cx, cy, rz, rx, ry, density)."""
    z, x, y = _grid(shape)
    vol = np.zeros(shape, np.float64)
    for cz, cx, cy, rz, rx, ry, dens in items:
        inside = ((z - cz) / rz) ** 2 + ((x - cx) / rx) ** 2 + ((y - cy) / ry) ** 2 <= 1.0
        vol[inside] = dens
    return vol.astype(dtype)


def two_spheres(shape, seed=1, radii=(12.0, 16.0), densities=(1.0, 0.6), dtype=np.float32):
    """Two spheres side by side in x, centred in z and y (seed jitters centres by <=2 voxels)."""
    nz, nx, ny = shape
    rng = np.random.default_rng(seed)
    j = rng.uniform(-2, 2, size=4)
    items = [
        (nz / 2 + j[0], nx * 0.3 + j[1], ny / 2, radii[0], radii[0], radii[0], densities[0]),
        (nz / 2 + j[2], nx * 0.7 + j[3], ny / 2, radii[1], radii[1], radii[1], densities[1]),
    ]
    return ellipsoids(shape, items, dtype)


def random_ellipsoids(shape, n=20, seed=2, dens=(0.2, 1.0), dtype=np.float32):
    """n random spheres/ellipsoids fully inside the volume, densities U[dens]."""
    nz, nx, ny = shape
    rng = np.random.default_rng(seed)
    items = []
    rmax = 0.22 * min(nz, nx)
    for _ in range(n):
        rz, rx = rng.uniform(0.05 * min(nz, nx), rmax, size=2)
        ry = rng.uniform(0.2 * ny, 0.5 * ny)
        if rng.uniform() < 0.5:
            rx = rz
        cz = rng.uniform(rz + 1, nz - rz - 1)
        cx = rng.uniform(rx + 1, nx - rx - 1)
        cy = rng.uniform(0, ny)
        items.append((cz, cx, cy, rz, rx, ry, rng.uniform(*dens)))
    return ellipsoids(shape, items, dtype)


def random_spheres(shape, n=12, seed=3, dens=(0.2, 1.0), dtype=np.float32):
    """n random spheres inside the volume (radius up to 0.2 of the smallest of N_z, N_x and
    twice N_y), densities U[dens]; later spheres overwrite earlier ones.  Each sphere is
    drawn inside its bounding box only, so large volumes stay cheap."""
    nz, nx, ny = shape
    rng = np.random.default_rng(seed)
    vol = np.zeros(shape, dtype)
    rmax = max(2.0, 0.2 * min(nz, nx, 2 * ny))
    for _ in range(n):
        r = rng.uniform(0.3 * rmax, rmax)
        c = [rng.uniform(r, nz - r), rng.uniform(r, nx - r), rng.uniform(0, ny)]
        lo = [max(0, int(np.floor(c[k] - r))) for k in range(3)]
        hi = [min(d, int(np.ceil(c[k] + r)) + 1) for k, d in enumerate(shape)]
        z = np.arange(lo[0], hi[0])[:, None, None] + 0.5
        x = np.arange(lo[1], hi[1])[None, :, None] + 0.5
        y = np.arange(lo[2], hi[2])[None, None, :] + 0.5
        inside = (z - c[0]) ** 2 + (x - c[1]) ** 2 + (y - c[2]) ** 2 <= r * r
        box = vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        box[inside] = rng.uniform(*dens)
    return vol


def noisy_sinogram(proj, sigma, seed, gains=None, offsets=None):
    """y = gain * proj + offset + N(0, sigma^2), proj laid out [N_theta][N_r][N_t]."""
    rng = np.random.default_rng(seed)
    y = np.array(proj, np.float64, copy=True)
    if gains is not None:
        y *= np.asarray(gains, np.float64)[:, None, None]
    if offsets is not None:
        y += np.asarray(offsets, np.float64)[:, None, None]
    if sigma > 0:
        y += rng.normal(0.0, sigma, size=y.shape)
    return y
