"""GPU: the bright-field engine adapter and its coarse-to-fine driver
(BFReconstructionEngine, BFMultiResolutionReconstruction; engine_cli --bf-multires
reads an MRC of transmitted counts) against the C-ABI's bright-field job for the
same problem, whose kernels the other bright-field tests hold to the oracle."""
import os
import subprocess

import numpy as np
import pytest

from mbirtem import capi
from problems import bf_problem

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(capi.LIB_PATH), "engine_cli")
F32 = lambda v: float(np.float32(v))  # the driver's float properties


def _series(tmp_path, counts, ang):
    from test_io import make_tilt_series
    path = str(tmp_path / "bf_counts.mrc")
    make_tilt_series(path, counts.astype(np.float32).transpose(0, 2, 1), mode=2,
                     tilts=ang.astype(np.float32), pixel=1e-9)
    return path


def _cli(mrc, levels, th, tmp, out, outer, inner, stop, seed, off, use_default):
    args = [CLI, "--bf-multires", mrc, str(levels), str(th), "0.01", "0.2", str(outer),
            str(inner), str(stop), str(seed), tmp, out, "3", "0.5", "0", repr(off),
            "1" if use_default else "0"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    return r, open(os.path.join(tmp, "levels.txt")).read()


def test_bf_engine_equals_capi_bit_for_bit(gpu_ctx, tmp_path):
    """One level: the adapter's volume, offsets, variances, selector MRC and
    ReconstructedSinogram.bin hold exactly the C-ABI job's results."""
    from test_io import parse_mrc
    g, ang, _, counts = bf_problem(bragg_frac=0.05)
    ang32 = ang.astype(np.float32)
    c32 = counts.astype(np.float32)
    mrc = _series(tmp_path, c32, ang32)
    off = F32(-np.log(2000.0))
    th = F32(g.LengthZ / 2)
    tmp, out = str(tmp_path / "t"), str(tmp_path / "bf.mrc")
    os.makedirs(tmp)
    r, lv = _cli(mrc, 1, th, tmp, out, 3, 12, 1e-4, 7, off, True)
    assert r.returncode == 0 and "error 0" in lv, r.stdout + r.stderr + lv
    d = F32(1e-9) * 1e9
    geom = capi.geometry_init(g.N_r, g.N_t, ang32.astype(np.float64), th, delta_r=d, delta_t=d)
    job = capi.Job(gpu_ctx, geom, ang32.astype(np.float64), c32, seed=7, num_outer=3,
                   num_inner=12, stop_threshold=F32(1e-4), sigma_x=F32(0.01),
                   mrf_q=F32(np.float32(0.2) + np.float32(1)), mrf_c=0.01, bf_mode=1,
                   bragg_T=3.0, bragg_delta=0.5, bf_offset=0.0, default_offset=off,
                   joint_estimation=1, noise_model=1, zero_skipping=1, bf_nhicd=1)
    job.run()
    x = job.volume()
    nz, nx, ny = x.shape
    lvl = os.path.join(tmp, "1x")
    raw = np.fromfile(os.path.join(lvl, "ReconstructedObject.bin"), "<f8").reshape(ny, nx, nz)
    np.testing.assert_array_equal(raw, x.transpose(2, 1, 0).astype(np.float64))
    I0, mu, al = job.nuisance()
    np.testing.assert_array_equal(np.fromfile(os.path.join(lvl, "FinalOffsetParameters.bin"), "<f8"), mu)
    np.testing.assert_array_equal(np.fromfile(os.path.join(lvl, "FinalVariances.bin"), "<f8"), al)
    sel = job.selector()
    h, data = parse_mrc(os.path.join(lvl, "BraggSelector.mrc"))
    nt = len(ang)
    assert (h["nx"], h["ny"], h["nz"]) == (g.N_r, g.N_t, nt)
    for k in range(nt):
        np.testing.assert_array_equal(data[k], sel[nt - 1 - k].T.astype(np.float32))
    assert sel.mean() < 1.0  # rejections took part
    y = -np.log(c32.astype(np.float64))
    want = (y - job.error_sinogram().astype(np.float64)) * I0[:, None, None] + mu[:, None, None]
    got = np.fromfile(os.path.join(lvl, "ReconstructedSinogram.bin"), "<f8").reshape(want.shape)
    np.testing.assert_allclose(got, want, rtol=1e-13)  # numpy's log and the C library's
    assert parse_mrc(out)[0]["nz"] == nz
    job.close()


def test_bf_multiresolution_driver(gpu_ctx, tmp_path):
    """Two levels from fitted initial offsets: the coarse level (two detector rows per
    slice) hands its volume and offsets to the fine one through the files, both
    write their selector, and the final volume is written."""
    from test_io import parse_mrc
    g, ang, _, counts = bf_problem(n_t=8, bragg_frac=0.05)
    mrc = _series(tmp_path, counts, ang)
    tmp, out = str(tmp_path / "t"), str(tmp_path / "bf.mrc")
    os.makedirs(tmp)
    r, lv = _cli(mrc, 2, g.LengthZ / 2, tmp, out, 2, 4, 1e-3, 5, 0.0, False)
    assert r.returncode == 0 and "error 0" in lv, r.stdout + r.stderr + lv
    lines = [l.split() for l in lv.splitlines() if l.startswith("level")]
    assert len(lines) == 2
    assert float(lines[0][3]) == 2.0 and float(lines[1][3]) == 1.0  # voxel size multiples
    for d in ("2x", "1x"):
        assert os.path.exists(os.path.join(tmp, d, "BraggSelector.mrc"))
        assert os.path.exists(os.path.join(tmp, d, "FinalOffsetParameters.bin"))
    h, data = parse_mrc(out)
    assert np.isfinite(data).all() and data.max() > 0


def test_bf_three_levels(gpu_ctx, tmp_path):
    """The reference's default of three resolution levels (BFReconstructionArgsParser.cpp:
    171): voxels of 4, 2 and 1 detector pitches, level 0 with four detector rows per
    slice; each level hands its volume and offsets to the next through the files."""
    from test_io import parse_mrc
    g, ang, _, counts = bf_problem(n_r=64, n_t=16, bragg_frac=0.05)
    mrc = _series(tmp_path, counts, ang)
    tmp, out = str(tmp_path / "t"), str(tmp_path / "bf.mrc")
    os.makedirs(tmp)
    r, lv = _cli(mrc, 3, 16.0, tmp, out, 2, 4, 1e-3, 5, 0.0, False)
    assert r.returncode == 0 and "error 0" in lv, r.stdout + r.stderr + lv
    lines = [l.split() for l in lv.splitlines() if l.startswith("level")]
    assert [float(l[3]) for l in lines] == [4.0, 2.0, 1.0]
    dims = [tuple(int(v) for v in l[15:18]) for l in lines]
    assert all(2 * a == b for a, b in zip(dims[0], dims[1]))
    assert all(2 * a == b for a, b in zip(dims[1], dims[2]))
    for d in ("4x", "2x", "1x"):
        assert os.path.exists(os.path.join(tmp, d, "BraggSelector.mrc"))
    h, data = parse_mrc(out)
    assert np.isfinite(data).all() and data.max() > 0

