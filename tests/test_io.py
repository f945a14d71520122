"""The engine's files (include/mbir_io.h), without a GPU.

The MRC checks parse the files with an independent reader written here from the
published IMOD MRC layout (1024-byte header, little-endian fields at fixed offsets,
FEI extended header of 128 bytes per section); the sub-volume and crop rules are
checked on hand-computed cases; the binary outputs against their element order.
"""
import os
import struct

import numpy as np
import pytest

from mbirtem import capi


def parse_mrc(path):
    """Header fields, FEI pixel sizes / tilts and the data of a float MRC file."""
    b = open(path, "rb").read()
    h = {
        "nx": struct.unpack_from("<i", b, 0)[0], "ny": struct.unpack_from("<i", b, 4)[0],
        "nz": struct.unpack_from("<i", b, 8)[0], "mode": struct.unpack_from("<i", b, 12)[0],
        "mxyz": struct.unpack_from("<3i", b, 28), "len": struct.unpack_from("<3f", b, 40),
        "angles": struct.unpack_from("<3f", b, 52), "map": struct.unpack_from("<3i", b, 64),
        "amin": struct.unpack_from("<f", b, 76)[0], "amax": struct.unpack_from("<f", b, 80)[0],
        "amean": struct.unpack_from("<f", b, 84)[0], "next": struct.unpack_from("<i", b, 92)[0],
        "cmap": b[208:212], "stamp": b[212:214], "nlabels": struct.unpack_from("<i", b, 220)[0],
        "label0": b[224:304].split(b"\0")[0].decode(),
    }
    nz, ny, nx = h["nz"], h["ny"], h["nx"]
    ext = b[1024:1024 + h["next"]]
    h["pixel"] = [struct.unpack_from("<f", ext, 128 * z + 44)[0] for z in range(nz)]
    data = np.frombuffer(b, "<f4", nx * ny * nz, 1024 + h["next"]).reshape(nz, ny, nx)
    return h, data


def make_tilt_series(path, data, mode=2, tilts=None, pixel=1e-9, fei=True):
    """An MRC tilt series [nz][ny][nx] with an FEI extended header (a_tilt, pixel size)."""
    nz, ny, nx = data.shape
    hdr = bytearray(1024)
    struct.pack_into("<4i", hdr, 0, nx, ny, nz, mode)
    struct.pack_into("<3i", hdr, 28, nx, ny, nz)
    struct.pack_into("<3i", hdr, 64, 1, 2, 3)
    nxt = 128 * nz if fei else 0
    struct.pack_into("<i", hdr, 92, nxt)
    hdr[208:212] = b"MAP "
    hdr[212:214] = bytes([68, 65])
    struct.pack_into("<i", hdr, 220, 1)
    hdr[224:224 + 30] = (b"Fei Company (C) Copyright 2003" if fei else b"plain tilt series".ljust(30))
    ext = bytearray(nxt)
    if fei:
        for z in range(nz):
            struct.pack_into("<f", ext, 128 * z, tilts[z])
            struct.pack_into("<f", ext, 128 * z + 44, pixel)
    dt = {1: "<i2", 2: "<f4", 6: "<u2", 4: "<f4"}[mode]  # mode 4 (complex) only to be refused
    with open(path, "wb") as f:
        f.write(bytes(hdr) + bytes(ext) + np.ascontiguousarray(data).astype(dt).tobytes())


def test_mrc_volume_header_and_layout(tmp_path):
    rng = np.random.default_rng(3)
    vol = rng.uniform(0.1, 2.0, (5, 7, 3)).astype(np.float32)  # [N_z][N_x][N_y]
    p = str(tmp_path / "v.mrc")
    capi.mrc_write_volume(p, vol, x0=1, x1=6, length_x=12.5)
    h, data = parse_mrc(p)
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (5, 3, 5, 2)
    assert h["mxyz"] == (5, 3, 5) and h["len"] == (5.0, 3.0, 5.0)
    assert h["angles"] == (90.0, 90.0, 90.0) and h["map"] == (1, 2, 3)
    assert h["next"] == 128 * 5 and h["cmap"] == b"MAP " and h["stamp"] == bytes([68, 65])
    assert "Fei Company" in h["label0"] and h["nlabels"] >= 1
    assert h["pixel"] == [12.5] * 5
    # sections from the last z down, rows y, columns x over the crop [1, 6)
    for k in range(5):
        np.testing.assert_array_equal(data[k], vol[4 - k, 1:6, :].T)
    crop = vol[:, 1:6, :].astype(np.float64)
    assert h["amin"] == np.float32(crop.min()) and h["amax"] == np.float32(crop.max())
    assert h["amean"] == pytest.approx(crop.mean(), rel=1e-6)
    info, tilts, pix = capi.mrc_read_info(p)
    assert (info.nx, info.ny, info.nz, info.fei) == (5, 3, 5, 1)
    assert info.amean == h["amean"] and list(pix) == [12.5] * 5


def test_mrc_volume_maximum_starts_at_zero(tmp_path):
    # the writer's running maximum starts from the smallest positive double rounded
    # to float, i.e. 0: an all-negative volume reports amax = 0
    vol = -np.ones((2, 3, 4), np.float32)
    p = str(tmp_path / "n.mrc")
    capi.mrc_write_volume(p, vol)
    h, _ = parse_mrc(p)
    assert h["amax"] == 0.0 and h["amin"] == -1.0 and h["amean"] == -1.0


@pytest.mark.parametrize("mode", [1, 2, 6])
def test_tilt_series_read_modes(tmp_path, mode):
    rng = np.random.default_rng(mode)
    nz, ny, nx = 6, 4, 10
    data = rng.integers(0, 3000, (nz, ny, nx)).astype(np.float64)
    tilts = np.linspace(-60, 60, nz).astype(np.float32)
    p = str(tmp_path / f"t{mode}.mrc")
    make_tilt_series(p, data, mode=mode, tilts=tilts, pixel=2.5e-9)
    counts, ang, good, d = capi.mrc_read_tilt_series(p)
    assert (d.N_theta, d.N_r, d.N_t) == (nz, nx, ny)
    # counts[theta][r = column x][t = row y]
    np.testing.assert_array_equal(counts, data.transpose(0, 2, 1))
    np.testing.assert_allclose(ang, tilts.astype(np.float64))
    assert list(good) == list(range(nz))
    assert d.delta_r == pytest.approx(2.5) and d.delta_t == pytest.approx(2.5)


def test_tilt_series_selection_rules(tmp_path):
    nz, ny, nx = 6, 20, 100  # centre column 50
    data = np.arange(nz * ny * nx, dtype=np.float64).reshape(nz, ny, nx) % 1000
    tilts = np.linspace(-50, 50, nz).astype(np.float32)
    p = str(tmp_path / "s.mrc")
    make_tilt_series(p, data, tilts=tilts)
    # x 10..79: 40 columns left of the centre, 30 right -> start moves to 50 - 30 = 20
    _, _, _, d = capi.mrc_read_tilt_series(p, True, x=(10, 79), y=(0, 8), z=(0, 5))
    assert (d.x_start, d.x_end, d.N_r) == (20, 79, 60)
    # rows 0..8 (9 rows) are already a multiple of 3
    assert (d.y_start, d.y_end, d.N_t) == (0, 8, 9)
    # x 30..89: 20 left, 40 right -> end moves to 50 + 20 - 1 = 69; rows 0..9 -> 0..11
    counts, ang, good, d = capi.mrc_read_tilt_series(p, True, x=(30, 89), y=(0, 9), z=(0, 5),
                                                     excluded=(2, 4))
    assert (d.x_start, d.x_end, d.N_r) == (30, 69, 40)
    assert (d.y_end, d.N_t) == (11, 12)
    assert list(good) == [0, 1, 3, 5] and d.N_theta == 4
    np.testing.assert_allclose(ang, tilts[[0, 1, 3, 5]])
    np.testing.assert_array_equal(counts[2], data[3, 0:12, 30:70].T)
    # explicit tilts replace the file's, one per used view
    _, ang2, _, _ = capi.mrc_read_tilt_series(p, True, x=(30, 89), y=(0, 9), z=(0, 5),
                                              excluded=(2, 4), tilts=[1, 2, 3, 4])
    np.testing.assert_array_equal(ang2, [1, 2, 3, 4])


def test_tilt_series_errors(tmp_path):
    p = str(tmp_path / "plain.mrc")
    make_tilt_series(p, np.ones((3, 2, 4)), fei=False)
    with pytest.raises(capi.MbirError) as ei:  # no tilt angles anywhere
        capi.mrc_read_tilt_series(p)
    assert ei.value.code == -6
    counts, ang, _, d = capi.mrc_read_tilt_series(p, tilts=[-1, 0, 1])
    assert counts.shape == (3, 4, 2) and d.delta_r == 1.0
    with pytest.raises(capi.MbirError):
        capi.mrc_read_tilt_series(str(tmp_path / "missing.mrc"), tilts=[0])
    q = str(tmp_path / "c.mrc")
    make_tilt_series(q, np.ones((2, 2, 2)), mode=4, tilts=[0, 1])
    with pytest.raises(capi.MbirError) as ei:
        capi.mrc_read_tilt_series(q)
    assert ei.value.code == -5


def test_original_x_dims():
    # centres x0 + (j + 0.5) delta_xz inside [-N_r delta_r / 2, N_r delta_r / 2]
    assert capi.original_x_dims(10, -5.0, 1.0, 6, 1.0) == (2, 7)
    assert capi.original_x_dims(8, -4.0, 1.0, 8, 1.0) == (0, 7)
    assert capi.original_x_dims(20, -10.0, 2.0, 4, 1.0) == (4, 5)


def test_reconstructed_sinogram_file(tmp_path):
    rng = np.random.default_rng(1)
    counts = rng.normal(size=(3, 4, 2))
    e = rng.normal(size=(3, 4, 2)).astype(np.float32)
    I0 = np.array([1.5, 0.5, 2.0])
    mu = np.array([0.1, -0.2, 0.3])
    p = str(tmp_path / "ReconstructedSinogram.bin")
    capi.write_reconstructed_sinogram(p, counts, e, I0, mu)
    got = np.fromfile(p, "<f8").reshape(3, 4, 2)
    want = (counts - e.astype(np.float64)) * I0[:, None, None] + mu[:, None, None]
    np.testing.assert_array_equal(got, want)


def test_raw_volume_layout_and_interpolated_read(tmp_path):
    rng = np.random.default_rng(2)
    vol = rng.normal(size=(4, 6, 2)).astype(np.float32)  # [N_z][N_x][N_y]
    p = str(tmp_path / "ReconstructedObject.bin")
    capi.write_raw_volume(p, vol)
    raw = np.fromfile(p, "<f8").reshape(2, 6, 4)  # y, x, z with z fastest
    np.testing.assert_array_equal(raw, vol.transpose(2, 1, 0).astype(np.float64))
    np.testing.assert_array_equal(capi.read_raw_volume(p, (4, 6, 2)), vol.astype(np.float64))
    # a half-size file read with interpolation: every voxel fills a 2 x 2 x 2 block
    up = capi.read_raw_volume(p, (8, 12, 4), interp=1)
    for z in range(8):
        for x in range(12):
            for y in range(4):
                assert up[z, x, y] == np.float64(vol[z // 2, x // 2, y // 2])
    with pytest.raises(capi.MbirError):
        capi.read_raw_volume(p, (5, 6, 2))


def test_initial_offsets_least_squares():
    # per-view means a / cos(theta) + b + noise; the fit must be the least-squares
    # solution (numpy's solver on the same design matrix)
    rng = np.random.default_rng(5)
    ang = np.linspace(-60, 60, 21)
    g = 1 / np.cos(np.radians(ang))
    means = 0.7 * g + 0.25 + rng.normal(0, 0.01, ang.size)
    counts = means[:, None, None] + rng.normal(0, 0.05, (ang.size, 6, 4))
    counts -= counts.mean(axis=(1, 2), keepdims=True) - means[:, None, None]
    off, (a, b) = capi.initial_offsets(counts, ang)
    A = np.stack([g, np.ones_like(g)], 1)
    (ra, rb), *_ = np.linalg.lstsq(A, counts.mean(axis=(1, 2)), rcond=None)
    assert a == pytest.approx(ra, rel=1e-10) and b == pytest.approx(rb, rel=1e-10)
    assert np.all(off == b)
    with pytest.raises(capi.MbirError) as ei:  # one distinct angle: no fit
        capi.initial_offsets(counts[:3], np.zeros(3))
    assert ei.value.code == -2


def test_sigma_x_estimate_hand_values():
    # view 0 (0 deg): image rows [0, 2], [0, 2] -> mean 1, mean |dev| 1
    # view 1 (60 deg): [1, 1], [1, 5] -> mean 2, mean |dev| 1.5
    # thickness 2, gain 1: (1 * 1 / 2 + 1.5 * 0.5 / 2) / 2 = 0.4375
    v0 = np.array([[0.0, 2.0], [0.0, 2.0]])
    v1 = np.array([[1.0, 1.0], [1.0, 5.0]])
    counts = np.stack([v0.T, v1.T])  # [theta][r = column][t = row]
    assert capi.sigma_x_estimate(counts, [0.0, 60.0], 2.0) == pytest.approx(0.4375, rel=1e-12)
    # gain divides, a sub-image selects: columns [1, 2) of view 0 are constant
    assert capi.sigma_x_estimate(counts, [0.0, 60.0], 2.0, target_gain=2.0) == \
        pytest.approx(0.4375 / 2, rel=1e-12)
    assert capi.sigma_x_estimate(counts[:1], [0.0], 1.0, x=(1, 2)) == 0.0
    # bright field: deviations of log(counts + offset): e^1, e^3 -> 1, 3 -> 1
    bf = np.array([[[np.e - 0.5, np.e ** 3 - 0.5]]])
    assert capi.sigma_x_estimate(bf, [0.0], 1.0, bf_offset=0.5) == pytest.approx(1.0, rel=1e-12)
