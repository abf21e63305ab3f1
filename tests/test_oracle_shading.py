"""Pins of the oracle's sky, tone mapping, noise and textures against
published constants, closed forms and an independent FP64 implementation."""

import ctypes

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi
from scene_helpers import argb


def _daylight(x, y, lum=1.0):
    out = (ctypes.c_float * 3)()
    oracle_ffi.lib().oracle_daylight_xyy_to_srgb(ctypes.c_float(x), ctypes.c_float(y), ctypes.c_double(lum), out)
    return np.array(out[:])


def test_daylight_d65_is_srgb_white():
    # D65 (x, y) = (0.31271, 0.32902) is the sRGB white point: the CIE daylight
    # weights, the S0/S1/S2 integrals and the XYZ->sRGB matrix together must map
    # it to R = G = B = Y.
    rgb = _daylight(0.31271, 0.32902, 1.0)
    np.testing.assert_allclose(rgb, [1.0, 1.0, 1.0], rtol=1.5e-2)


def test_daylight_d50_matches_published_xyz():
    # D50 white XYZ = (0.9642, 1, 0.8249); through the IEC 61966-2-1 sRGB matrix
    # that is linear sRGB (1.176, 0.976, 0.722).
    rgb = _daylight(0.34567, 0.35850, 1.0)
    np.testing.assert_allclose(rgb, [1.176, 0.976, 0.722], rtol=2e-2)


def test_sky_below_horizon_and_inactive_are_constant():
    sky = dfa.sky_set((0.3, 0.8, 0.5), 2.5)
    np.testing.assert_array_equal(oracle_ffi.sky_color(sky, (0.0, -0.5, 0.866)), np.full(3, 0.01, np.float32))
    np.testing.assert_array_equal(oracle_ffi.sky_color(sky, (0.0, 1.0, 0.0), active=False), np.full(3, 0.01, np.float32))


def test_sky_symmetric_about_sun_vertical_plane():
    # Perez depends on the zenith angle and the angle to the sun only, so
    # directions mirrored across the plane of zenith and sun agree.
    sun = np.array([0.6, 0.5, 0.0])
    sky = dfa.sky_set(tuple(sun), 3.0)
    rng = np.random.default_rng(2)
    for _ in range(50):
        d = rng.normal(size=3)
        d[1] = abs(d[1]) + 0.05
        d /= np.linalg.norm(d)
        m = d * np.array([1.0, 1.0, -1.0])  # mirror z (the sun lies in the x-y plane)
        a = oracle_ffi.sky_color(sky, tuple(d.astype(np.float32)))
        b = oracle_ffi.sky_color(sky, tuple(m.astype(np.float32)))
        np.testing.assert_allclose(a, b, rtol=1e-5)


def test_sky_brightest_toward_the_sun():
    sun = np.array([0.3, 0.7, 0.2])
    sun = sun / np.linalg.norm(sun)
    sky = dfa.sky_set(tuple(sun), 2.0)
    toward = oracle_ffi.sky_color(sky, tuple(sun.astype(np.float32))).sum()
    away = oracle_ffi.sky_color(sky, tuple((np.array([-sun[0], sun[1], -sun[2]])).astype(np.float32))).sum()
    assert toward > 2.0 * away > 0.0


def _tonemap(params, rgb, prim=-1):
    f = (ctypes.c_float * 3)(*rgb)
    of = (ctypes.c_float * 3)()
    ob = (ctypes.c_uint8 * 4)()
    oracle_ffi.lib().oracle_tonemap(ctypes.byref(params), f, ctypes.c_int32(prim), of, ob)
    return np.array(of[:]), np.array(ob[:])


def _srgb(v):
    v = np.asarray(v, np.float64)
    return np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1 / 2.4) - 0.055)


def test_reinhard_and_srgb_closed_form():
    p = dfa.default_params()
    p.tone_mapper_type = 2
    p.tone_mapper_exposure = 1.0
    for x in (0.0, 0.001, 0.1, 0.5, 1.0, 3.0, 50.0):
        f, bgra = _tonemap(p, (x, x, x))
        expect = _srgb(x / (1 + x))
        # image4j's sRGB uses breakpoint 0.00304 with a matched slope: within 3e-4 of IEC sRGB
        np.testing.assert_allclose(f, [expect] * 3, atol=3e-4)
        assert bgra[3] == 255
        assert abs(int(bgra[0]) - int(np.floor(expect * 255 + 0.5))) <= 1


def test_aces_and_exposure_closed_form():
    p = dfa.default_params()
    p.tone_mapper_type = 1
    p.tone_mapper_exposure = 2.0
    for x in (0.05, 0.25, 0.5, 1.0, 4.0):
        m = 2.0 * x
        aces = min(max(m * (2.51 * m + 0.03) / (m * (2.43 * m + 0.59) + 0.14), 0.0), 1.0)
        f, _ = _tonemap(p, (x, 0.0, 0.0))
        assert f[0] == pytest.approx(_srgb(aces), abs=3e-4)


def test_reinhard_modified_closed_forms():
    p = dfa.default_params()
    p.tone_mapper_exposure = 1.0
    p.tone_mapper_type = 3
    f, _ = _tonemap(p, (0.7, 0.7, 0.7))
    assert f[0] == pytest.approx(_srgb(1 - np.exp(-0.7)), abs=3e-4)
    p.tone_mapper_type = 4
    f, _ = _tonemap(p, (0.7, 0.7, 0.7))
    assert f[0] == pytest.approx(_srgb(0.7 * (1 + 0.7 / 16) / 1.7), abs=3e-4)


def test_selection_tint_and_byte_packing():
    p = dfa.default_params()
    p.tone_mapper_type = 2
    p.selected_primitive_offset = 6
    _, bgra = _tonemap(p, (0.0, 0.0, 0.0), prim=6)
    # green += 1 before tone mapping: 1/(1+1) = 0.5 -> sRGB 0.7354 -> 188
    assert tuple(bgra) == (0, int(np.floor(_srgb(0.5) * 255 + 0.5)), 0, 255)
    _, bgra = _tonemap(p, (0.0, 0.0, 0.0), prim=3)
    assert tuple(bgra) == (0, 0, 0, 255)


# ---------------------------------------------------------------- noise

PERM = np.array([
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69, 142, 8, 99, 37, 240,
    21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252, 219, 203, 117, 35, 11, 32, 57, 177, 33, 88,
    237, 149, 56, 87, 174, 20, 125, 136, 171, 168, 68, 175, 74, 165, 71, 134, 139, 48, 27, 166, 77, 146, 158, 231, 83,
    111, 229, 122, 60, 211, 133, 230, 220, 105, 92, 41, 55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216,
    80, 73, 209, 76, 132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100, 109, 198, 173, 186,
    3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82, 85, 212, 207, 206, 59, 227, 47, 16, 58, 17,
    182, 189, 28, 42, 23, 183, 170, 213, 119, 248, 152, 2, 44, 154, 163, 70, 221, 153, 101, 155, 167, 43, 172, 9, 129,
    22, 39, 253, 19, 98, 108, 110, 79, 113, 224, 232, 178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238,
    210, 144, 12, 191, 179, 162, 241, 81, 51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31, 181, 199, 106, 157, 184,
    84, 204, 176, 115, 121, 50, 45, 127, 4, 150, 254, 138, 236, 205, 93, 222, 114, 67, 29, 24, 72, 243, 141, 128, 195,
    78, 66, 215, 61, 156, 180] * 2)
GRAD3 = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
                  [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], float)


def simplex3_fp64(x, y, z):
    """Simplex noise in 3-D after S. Gustavson, "Simplex noise demystified" (2005), in float64."""
    F3, G3 = 1.0 / 3.0, 1.0 / 6.0
    s = (x + y + z) * F3
    i, j, k = int(np.floor(x + s)), int(np.floor(y + s)), int(np.floor(z + s))
    t = (i + j + k) * G3
    x0, y0, z0 = x - (i - t), y - (j - t), z - (k - t)
    if x0 >= y0:
        if y0 >= z0:
            o1, o2 = (1, 0, 0), (1, 1, 0)
        elif x0 >= z0:
            o1, o2 = (1, 0, 0), (1, 0, 1)
        else:
            o1, o2 = (0, 0, 1), (1, 0, 1)
    else:
        if y0 < z0:
            o1, o2 = (0, 0, 1), (0, 1, 1)
        elif x0 < z0:
            o1, o2 = (0, 1, 0), (0, 1, 1)
        else:
            o1, o2 = (0, 1, 0), (1, 1, 0)
    corners = [(0, 0, 0), o1, o2, (1, 1, 1)]
    ii, jj, kk = i & 255, j & 255, k & 255
    total = 0.0
    for n, (a, b, c) in enumerate(corners):
        xc, yc, zc = x0 - a + n * G3, y0 - b + n * G3, z0 - c + n * G3
        gi = PERM[ii + a + PERM[jj + b + PERM[kk + c]]] % 12
        tt = 0.6 - xc * xc - yc * yc - zc * zc
        if tt > 0:
            total += tt ** 4 * float(GRAD3[gi] @ (xc, yc, zc))
    return 32.0 * total


def test_simplex3_matches_fp64_textbook():
    rng = np.random.default_rng(4)
    pts = rng.uniform(-40, 40, (500, 3)).astype(np.float32)
    for x, y, z in pts:
        got = oracle_ffi.lib().oracle_simplex3(x, y, z)
        assert got == pytest.approx(simplex3_fp64(float(x), float(y), float(z)), abs=2e-5)


def test_simplex_zero_on_lattice_and_bounded():
    # At a simplex vertex the own corner contributes t^4 * g.0 = 0 and every other
    # corner lies outside the kernel radius.
    # (integer points whose coordinate sum is a multiple of 3 are 3-D simplex
    # vertices: the skew s = (x+y+z)/3 and unskew t = (i+j+k)/6 cancel exactly)
    for p in [(0.0, 0.0, 0.0), (3.0, -2.0, 5.0), (17.0, 4.0, -9.0)]:
        assert oracle_ffi.lib().oracle_simplex3(*p) == 0.0
    assert oracle_ffi.lib().oracle_simplex2(0.0, 0.0) == 0.0
    rng = np.random.default_rng(9)
    v = [oracle_ffi.lib().oracle_simplex3(*map(float, q)) for q in rng.uniform(-100, 100, (3000, 3))]
    assert max(abs(np.array(v))) < 1.0
    assert np.std(v) > 0.1


def test_fbm_normalisation_range():
    # fBm rescales the octave sum by the amplitude total into [minimum, maximum]
    rng = np.random.default_rng(10)
    v = np.array([oracle_ffi.lib().oracle_fbm3(0.8, 0.5, -0.26, 0.26, 16, *map(float, q))
                  for q in rng.uniform(-10, 10, (1000, 3))])
    assert v.min() >= -0.26 and v.max() <= 0.26
    assert abs(v.mean()) < 0.02


# ---------------------------------------------------------------- textures

def _rec(uv=(0.0, 0.0), p=(0.0, 0.0, 0.0), ns=(0.0, 1.0, 0.0)):
    r = np.zeros(27, np.float32)
    r[2:5] = p
    r[8:10] = uv
    r[19:22] = ns
    return r


def test_checkerboard_tiles():
    # color0 where exactly one of frac(u*su), frac(v*sv) exceeds 0.5 (GPURendererKernel.java:1741-1752)
    tex = [3, 8, argb(1, 0, 0), argb(0, 0, 1), 1.0, 0.0, 0.5, 0.5]
    red, blue = 0xFF0000, 0x0000FF
    assert oracle_ffi.texture_color(tex, 0, _rec(uv=(0.2, 0.2))) == blue
    assert oracle_ffi.texture_color(tex, 0, _rec(uv=(1.2, 0.2))) == red
    assert oracle_ffi.texture_color(tex, 0, _rec(uv=(1.2, 1.2))) == blue
    # equal colours: 0.8 / 1.2 brightness alternation of a grey
    tex = [3, 8, argb(0.5, 0.5, 0.5), argb(0.5, 0.5, 0.5), 1.0, 0.0, 0.5, 0.5]
    c0 = oracle_ffi.texture_color(tex, 0, _rec(uv=(1.2, 0.2))) & 0xFF
    c1 = oracle_ffi.texture_color(tex, 0, _rec(uv=(0.2, 0.2))) & 0xFF
    assert c0 == int(np.float32(128 / 255 * 0.8) * 255 + 0.5)
    assert c1 == int(np.float32(128 / 255 * 1.2) * 255 + 0.5)


def test_constant_uv_and_normal_textures():
    # the constant texture hands back the packed ARGB int itself; callers mask the channels
    assert oracle_ffi.texture_color([4, 3, argb(0.2, 0.4, 0.6)], 0, _rec()) & 0xFFFFFF == (51 << 16) | (102 << 8) | 153
    assert oracle_ffi.texture_color([8, 2], 0, _rec(uv=(0.5, 1.5))) == (128 << 16) | (255 << 8)
    assert oracle_ffi.texture_color([7, 3, 0], 0, _rec(ns=(0.0, 1.0, -1.0))) == (128 << 16) | (255 << 8) | 0


def test_blend_texture_mixes_children():
    tex = [1, 5, 5, 8, 0.25, 4, 3, argb(1, 0, 0), 4, 3, argb(0, 0, 1)]
    c = oracle_ffi.texture_color(tex, 0, _rec())
    assert (c >> 16) & 0xFF == int(np.float32(0.75) * 255 + 0.5)
    assert c & 0xFF == int(np.float32(0.25) * 255 + 0.5)


def test_image_texture_bilinear():
    # 2x2 image, UV scale 1: u in [0, 1) maps to x = u * 2
    img = [argb(0, 0, 0), argb(1, 0, 0), argb(0, 1, 0), argb(1, 1, 0)]
    tex = [6, 12, 1.0, 0.0, 2.0, 2.0, 1.0, 1.0] + img
    c = oracle_ffi.texture_color(tex, 0, _rec(uv=(0.25, 0.0)))
    # x = 0.5 between texel 0 (black) and 1 (red) -> half red
    assert (c >> 16) & 0xFF == 128 and (c >> 8) & 0xFF == 0
