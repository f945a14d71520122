"""Pinning of the oracle's detector model (CPU): geometry, voxel profile, beam,
detector response, voxel-line response and A-matrix columns.

Each check comes from outside the restated formula: hand-derived sizes, the
exact chord-length function of a square, quadrature of the continuous
convolution, and the area / centroid that any correct projector must preserve.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate

import oracle_py as O
from mbirtem import capi
from mbirtem import phantom as P


def chord(s, phi, d=1.0):
    """Length of the line at signed offset s through the square [-d/2, d/2]^2 whose
    direction makes angle phi with the z axis (brute force: clip the line)."""
    c, sn = np.cos(phi), np.sin(phi)
    # points p = s*(c, -s) + t*(s, c) in (x, z); the detector axis is (cos, -sin)
    ux, uz = sn, c          # along the ray
    px, pz = s * c, -s * sn  # foot point on the detector axis
    lo, hi = -np.inf, np.inf
    for p, u in ((px, ux), (pz, uz)):
        if abs(u) < 1e-15:
            if abs(p) > d / 2:
                return 0.0
            continue
        a, b = (-d / 2 - p) / u, (d / 2 - p) / u
        lo, hi = max(lo, min(a, b)), min(hi, max(a, b))
    return max(0.0, hi - lo)


def test_geometry_hand_values():
    ang = P.tilt_series(61, 60.0)
    g = O.geometry(128, 16, ang, thickness=32)
    # LengthZ = round(2 * 32) = 64; N_x = 128 / 1; N_y = 16; origins are -L/2
    assert (g.N_x, g.N_y, g.N_z) == (128, 16, 64)
    assert (g.x0, g.y0, g.z0, g.R0, g.T0) == (-64.0, -8.0, -32.0, -64.0, -8.0)
    g2 = O.geometry(512, 64, P.tilt_series(91, 70.0), thickness=128)
    assert (g2.N_x, g2.N_y, g2.N_z) == (512, 64, 256)
    # rounding of LengthZ: 2 * 10.3 = 20.6 -> 21 voxels
    g3 = O.geometry(32, 4, ang, thickness=10.3)
    assert g3.N_z == 21 and g3.z0 == -10.5
    # extend_object widens x: 0.6 * N_r / cos(60) + LengthZ * tan(60) rounded
    g4 = O.geometry(128, 4, ang, thickness=32, extend_object=1)
    want = np.floor(0.6 * 128 / np.cos(np.pi / 3) + 64 * np.tan(np.pi / 3) + 0.5)
    assert g4.N_x == int(want)


def test_product_geometry_equals_oracle():
    for nr, nt, nv, deg, th, ext in ((128, 16, 61, 60.0, 32, 0), (96, 8, 45, 70.0, 21.7, 1),
                                     (64, 4, 31, 80.0, 9, 1)):
        ang = P.tilt_series(nv, deg)
        g = O.geometry(nr, nt, ang, thickness=th, extend_object=ext)
        G = capi.geometry_init(nr, nt, ang, th, extend_object=ext)
        for f, _ in O.Geom._fields_:
            assert getattr(G, f) == getattr(g, f), f


def test_voxel_profile_is_the_chord_length():
    ang_deg = np.array([-70.0, -37.0, 0.0, 12.5, 45.0, 60.0])
    a = ang_deg.copy()
    prof = np.zeros((len(a), 1536))
    O.lib().orc_voxel_profile(O.dp(a), len(a), 1.0, O.dp(prof))
    np.testing.assert_allclose(a, ang_deg * np.pi / 180)  # converted in place
    t = 2.0 * np.arange(1536) / 1536
    for i, phi in enumerate(a):
        want = np.array([chord(tt - 1.0, phi) for tt in t])
        # pointwise away from the corners (at 0 degrees the profile is a box whose edge
        # samples are a measure-zero tie), and in the mean everywhere
        c, sn = abs(np.cos(phi)), abs(np.sin(phi))
        corners = np.array([(c + sn) / 2, abs(c - sn) / 2])
        away = np.min(np.abs(np.abs(t - 1.0)[:, None] - corners[None, :]), axis=1) > 2e-3
        assert np.max(np.abs(prof[i] - want)[away]) < 2e-3, ang_deg[i]
        assert np.mean(np.abs(prof[i] - want)) < 1e-3
        # area of the unit square
        assert abs(prof[i].sum() * 2.0 / 1536 - 1.0) < 2e-3


def test_engine_profile_uses_twice_converted_angles():
    """The engine's second calculateVoxelProfile sees radians (quirk of the reference)."""
    ang = np.array([-60.0, -30.0, 0.0, 30.0, 60.0])
    g = O.geometry(16, 2, ang, thickness=4)
    eng = O.Engine(g, ang, np.zeros((5, 16, 2)), O.params())
    prof = eng.profile()
    t = 2.0 * np.arange(1536) / 1536
    for i, d in enumerate(ang):
        phi_twice = d * (np.pi / 180) ** 2
        want = np.array([chord(tt - 1.0, phi_twice) for tt in t])
        assert np.mean(np.abs(prof[i] - want)) < 1e-3
        if d != 0:
            wrong = np.array([chord(tt - 1.0, d * np.pi / 180) for tt in t])
            assert np.max(np.abs(prof[i] - wrong)) > 0.05


def test_beam_profile():
    b = np.zeros(512)
    O.lib().orc_beam_profile(2.0, O.dp(b))
    assert b.sum() * 2.0 == pytest.approx(1.0, rel=1e-12)
    n = np.arange(512)
    ham = 0.54 - 0.46 * np.cos(2 * np.pi * n / 512)
    np.testing.assert_allclose(b, ham / ham.sum() / 2.0, rtol=1e-12)


def test_detector_response_is_the_beam_convolved_chord():
    ang = np.array([0.0, 20.0, 45.0])
    g = O.geometry(16, 2, ang, thickness=4)
    eng = O.Engine(g, ang, np.zeros((3, 16, 2)), O.params())
    H = eng.H()
    offR = O.lib().orc_offset_r(C.byref(g))
    for i, d in enumerate(ang):
        phi = d * (np.pi / 180) ** 2  # the profile the engine feeds DetectorResponse

        def w(u):  # continuous Hamming density over one detector bin, / delta_t
            return (0.54 - 0.46 * np.cos(2 * np.pi * (u + 0.5))) / 0.54

        for k in (0, 10, 30, 50, 63):
            dist = k * offR
            val, _ = integrate.quad(lambda u: w(u) * chord(u - dist, phi), -0.5, 0.5, limit=200,
                                    points=[dist - 0.75, dist - 0.5, dist + 0.5, dist + 0.75])
            assert abs(H[i, k] - val) < 5e-3, (d, k, H[i, k], val)
    # support ends near (half diagonal + half bin) = 64 * OffsetR
    assert H[:, -1].max() < 2e-2


def test_voxel_line_response_is_identity_when_voxels_match_rows():
    ang = P.tilt_series(7, 60.0)
    g = O.geometry(16, 6, ang, thickness=4)
    Ht = np.zeros((7, 64))
    O.lib().orc_init_ht(C.byref(g), O.lib().orc_offset_t(C.byref(g)), O.dp(Ht))
    for y in range(g.N_y):
        t = np.zeros(8, np.int32)
        v = np.zeros(8)
        n = O.lib().orc_vlr(C.byref(g), O.dp(Ht), O.lib().orc_offset_t(C.byref(g)), y,
                            t.ctypes.data_as(C.POINTER(C.c_int32)), O.dp(v), 8)
        assert n == 1 and t[0] == y and v[0] == g.delta_t


def _overlap(d, dxy, dt):
    """Length of the intersection of a voxel slab [-dxy/2, dxy/2] with a detector row
    of width dt centred at distance d."""
    return np.maximum(0.0, np.minimum(dxy / 2, d + dt / 2) - np.maximum(-dxy / 2, d - dt / 2))


@pytest.mark.parametrize("ratio", [2, 4])
def test_voxel_line_response_of_coarse_voxels(ratio):
    """Voxels ratio rows thick (the multi-resolution driver's coarse levels,
    delta_xy = delta_xz = ratio delta_t): slice y covers rows ratio y .. ratio y + ratio - 1.
    Each value is the slab/row overlap (an interval intersection, closed form) sampled
    every OffsetT and interpolated linearly (numpy), as the response table H_t is; the
    neighbouring rows' overlap is 0 and they are left out."""
    ang = P.tilt_series(7, 60.0)
    n_t = 4 * ratio
    g = O.geometry(16 * ratio, n_t, ang, thickness=4 * ratio, delta_xz=float(ratio))
    g.delta_xy = float(ratio)  # the driver sets delta_xy with delta_xz per level
    g.N_y = int(np.floor(g.LengthY / g.delta_xy))
    g.y0 = -g.LengthY / 2
    assert g.N_y * ratio == n_t
    offT = O.lib().orc_offset_t(C.byref(g))
    assert offT == pytest.approx((ratio / 2 + 0.5) / 64)
    Ht = np.zeros((7, 64))
    O.lib().orc_init_ht(C.byref(g), offT, O.dp(Ht))
    samples = np.arange(64) * offT
    np.testing.assert_allclose(Ht[0], _overlap(samples, g.delta_xy, g.delta_t), atol=1e-15)
    for y in range(g.N_y):
        t = np.zeros(8, np.int32)
        v = np.zeros(8)
        n = O.lib().orc_vlr(C.byref(g), O.dp(Ht), offT, y,
                            t.ctypes.data_as(C.POINTER(C.c_int32)), O.dp(v), 8)
        assert n == ratio
        np.testing.assert_array_equal(t[:n], ratio * y + np.arange(ratio))
        tc = g.y0 + (y + 0.5) * g.delta_xy
        rows = g.T0 + (t[:n] + 0.5) * g.delta_t
        expect = np.interp(np.abs(rows - tc), samples, _overlap(samples, g.delta_xy, g.delta_t))
        np.testing.assert_allclose(v[:n], expect, rtol=0, atol=1e-14)
        # every row lies inside the slab (overlap delta_t, up to the interpolation's dip
        # at the trapezoid's kink), so the slice's response sums to its thickness
        assert np.all(v[:n] <= g.delta_t) and np.all(v[:n] > 0.98 * g.delta_t)
        assert v[:n].sum() == pytest.approx(g.delta_xy, rel=1e-2)


def test_columns_preserve_area_and_centroid():
    ang = P.tilt_series(31, 70.0)
    g = O.geometry(48, 2, ang, thickness=12)
    eng = O.Engine(g, ang, np.zeros((31, 48, 2)), O.params())
    rad = ang * np.pi / 180
    rc = g.R0 + (np.arange(g.N_r) + 0.5) * g.delta_r
    for z, x in ((12, 24), (3, 10), (20, 40), (0, 0)):
        idx, val = eng.col(z, x)
        assert np.all(val > 0)
        th, r = idx // g.N_r, idx % g.N_r
        assert np.all(np.bincount(th, minlength=31) <= 3)
        xc = g.x0 + (x + 0.5) * g.delta_xz
        zc = g.z0 + (z + 0.5) * g.delta_xz
        for t in range(31):
            sel = th == t
            r0 = xc * np.cos(rad[t]) - zc * np.sin(rad[t])
            if r0 < g.R0 + 2 or r0 > g.RMax - 2:
                assert r0 > g.R0 - 2 or not sel.any()  # projects off the detector
                continue  # footprint clipped at (or beyond) the detector edge
            a = val[sel]
            # unit voxel area / bin width (the response of a square convolved with the beam)
            assert a.sum() == pytest.approx(1.0, abs=1.5e-2), (z, x, t)
            # centroid at the projected voxel centre r = x cos - z sin: sampling a ~2-bin
            # wide response at bin centres moves it by < 0.15 bin; a sign or x/z swap
            # moves it by whole bins
            assert np.sum(a * rc[r[sel]]) / a.sum() == pytest.approx(r0, abs=0.15)
