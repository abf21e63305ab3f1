"""Pins of the oracle's surface properties, normal mapping and remaining procedural textures against closed
forms and an independent float64 noise:

* sphere texture coordinates (GPURendererKernel.java:2452-2492): v = 1/2 + latitude / pi (1 at the north
  pole, 0 at the south pole), u an azimuthal coordinate advancing by beta / 2 pi per rotation beta about y;
* triangle attributes (:2541-2650): texture coordinates, Gouraud normals and tangents are the barycentric
  combinations of the vertex attributes (barycentric weights from a float64 Moller-Trumbore), normalised;
  flat shading gives the face normal;
* image normal maps (:2714-2846): a texel (R, G, B) decodes to (2R/255 - 1, 2G/255 - 1, 2B/255) in the
  tangent frame, so the mapped normal makes an angle with the shading normal whose cosine is
  b / |(r, g, b)|, for planes, spheres and triangles;
* noise normal maps (:2658-2712): N' = normalize(N + amount (f(p/s), f(rotated), f(rotated twice))) with f
  the 16-octave fBm in [-0.26, 0.26];
* the Bullseye texture (:1668-1712): colour A where frac(|p| / 4) > 1/2, else B, and the same-colour
  0.8 / 1.2 alternation; the fBm texture (:1776-1815): saturate(fbm(p) m + a).

Every check also runs against the oracle with one deliberate mistake switched on and must then fail."""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi
from scene_helpers import SceneBuilder, argb, default_camera
from test_oracle_shading import simplex3_fp64

R_P, R_UV, R_NG, R_NS, R_TAN = slice(2, 5), slice(8, 10), slice(5, 8), slice(19, 22), slice(24, 27)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _hit(scene, origin, target, params=None):
    o = np.asarray(origin, np.float32)
    d = _unit(np.asarray(target, np.float64) - o).astype(np.float32)
    t, rec = oracle_ffi.intersect(scene, dfa.default_params() if params is None else params, o, d)
    assert t < 1e5, (origin, target)
    return rec.astype(np.float64)


# ---------------------------------------------------------------- sphere UV

def _sphere_scene():
    b = SceneBuilder()
    b.sphere((0.0, 0.0, 0.0), 1.0, b.surface(3))
    return b.build(default_camera())


def _sphere_uv(sc, n):
    n = _unit(n)
    return _hit(sc, 3.0 * n, (0.0, 0.0, 0.0))[R_UV]


def _sphere_uv_ok():
    sc = _sphere_scene()
    known = {(0.0, 1.0, 0.0): (None, 1.0), (0.0, -1.0, 0.0): (None, 0.0), (-1.0, 0.0, 0.0): (0.5, 0.5),
             (0.0, 0.0, -1.0): (0.75, 0.5), (0.0, 0.0, 1.0): (0.25, 0.5)}
    for n, (u, v) in known.items():
        uv = _sphere_uv(sc, n)
        if abs(uv[1] - v) > 2e-6 or (u is not None and abs(uv[0] - u) > 2e-6):
            return False
    rng = np.random.default_rng(3)
    signs = set()
    for _ in range(100):
        lat = rng.uniform(-1.4, 1.4)
        lon = rng.uniform(-np.pi, np.pi)
        beta = rng.uniform(0.2, 1.0)
        n0 = np.array([np.cos(lat) * np.cos(lon), np.sin(lat), np.cos(lat) * np.sin(lon)])
        n1 = np.array([np.cos(lat) * np.cos(lon + beta), np.sin(lat), np.cos(lat) * np.sin(lon + beta)])
        a, b = _sphere_uv(sc, n0), _sphere_uv(sc, n1)
        if abs(a[1] - (0.5 + lat / np.pi)) > 3e-6:
            return False
        step = (b[0] - a[0]) % 1.0
        if abs(step - beta / (2 * np.pi)) < 3e-5:
            signs.add(1)
        elif abs(step - (1 - beta / (2 * np.pi))) < 3e-5:
            signs.add(-1)
        else:
            return False
    return len(signs) == 1


def test_sphere_texture_coordinates():
    assert _sphere_uv_ok()
    with oracle_ffi.planted("sphere_uv"):
        assert not _sphere_uv_ok()


# ---------------------------------------------------------------- triangle attributes

TRI = [np.array(p, np.float32) for p in ((0.0, 0.0, 0.0), (2.0, 0.0, 0.3), (0.4, 0.0, 1.8))]
TRI_N = [_unit(v) for v in ((0.0, 1.0, 0.0), (0.3, 1.0, 0.0), (0.0, 1.0, -0.4))]
TRI_UV = [(0.1, 0.2), (0.9, 0.3), (0.4, 0.95)]
TRI_T = [_unit(v) for v in ((1.0, 0.0, 0.0), (1.0, 0.1, 0.2), (1.0, -0.2, 0.0))]


def _triangle_scene(normal_texture=None):
    """TRI with its varying vertex frames, or, with a normal map, with the orthonormal frame N = y, T = x at
    every vertex (the map's closed form needs an orthonormal tangent frame)."""
    b = SceneBuilder()
    s = b.surface(3, normal_texture=normal_texture(b) if normal_texture else None)
    if normal_texture is None:
        b.triangle_attrs(TRI, TRI_N, TRI_UV, TRI_T, s)
    else:
        b.triangle_attrs(TRI, [(0.0, 1.0, 0.0)] * 3, TRI_UV, [(1.0, 0.0, 0.0)] * 3, s)
    return b.build(default_camera())


def _bary64(p):
    """Barycentric weights (w, u, v) of p on TRI in float64 (u of B, v of C)."""
    a, b, c = (x.astype(np.float64) for x in TRI)
    m = np.array([b - a, c - a]).T
    uv, *_ = np.linalg.lstsq(m, p - a, rcond=None)
    return 1 - uv.sum(), uv[0], uv[1]


def _triangle_ok():
    sc = _triangle_scene()
    rng = np.random.default_rng(4)
    pts = [(0.998, 0.001, 0.001), (0.001, 0.998, 0.001), (0.001, 0.001, 0.998)]
    pts += [tuple(x) for x in rng.dirichlet((1.5, 1.5, 1.5), 60)]
    for wb in pts:
        target = sum(w * x.astype(np.float64) for w, x in zip(wb, TRI))
        rec = _hit(sc, target + np.array([0.1, 2.0, 0.05]), target)
        w, u, v = _bary64(rec[R_P])
        if not np.allclose(rec[R_UV], w * np.array(TRI_UV[0]) + u * np.array(TRI_UV[1]) + v * np.array(TRI_UV[2]),
                           atol=3e-6):
            return False
        if not np.allclose(rec[R_NS], _unit(w * TRI_N[0] + u * TRI_N[1] + v * TRI_N[2]), atol=3e-6):
            return False
        if not np.allclose(rec[R_TAN], _unit(w * TRI_T[0] + u * TRI_T[1] + v * TRI_T[2]), atol=3e-6):
            return False
    p = dfa.default_params()
    p.shader_type = 1  # flat: the face normal
    face = _unit(np.cross(TRI[1] - TRI[0], TRI[2] - TRI[0]))
    rec = _hit(sc, (0.8, 2.0, 0.7), (0.8, 0.0, 0.7), p)
    return np.allclose(rec[R_NS], face, atol=2e-6) and np.allclose(rec[R_NG], face, atol=2e-6)


def test_triangle_barycentric_attributes():
    assert _triangle_ok()
    with oracle_ffi.planted("barycentric"):
        assert not _triangle_ok()


# ---------------------------------------------------------------- normal maps

TEXELS = [(255, 128, 0), (0, 128, 255), (128, 128, 255), (200, 60, 180), (30, 220, 90)]


def _image(rgb):
    def tex(b):
        c = argb(*(x / 255 for x in rgb))
        return b.texture([6, 12, 1.0, 0.0, 2.0, 2.0, 1.0, 1.0, c, c, c, c])
    return tex


def _decoded(rgb):
    r, g = ((np.float32(c) * np.float32(1 / 255) - 0.5) * 2 for c in rgb[:2])
    return np.array([r, g, np.float32(rgb[2]) * np.float32(1 / 255) * 2], np.float64)


def _normal_map_ok():
    for rgb in TEXELS:
        e = _decoded(rgb)
        cos = e[2] / np.linalg.norm(e)
        mapped = _image(rgb)
        p_off = dfa.default_params()
        p_off.normal_mapping = 0
        # plane (shape 1), sphere (2), triangle (4): the angle to the unmapped normal
        b = SceneBuilder()
        b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), b.surface(3, normal_texture=mapped(b)))
        plane = b.build(default_camera())
        b = SceneBuilder()
        b.sphere((0.0, 0.0, 0.0), 1.0, b.surface(3, normal_texture=mapped(b)))
        sphere = b.build(default_camera())
        tri = _triangle_scene(mapped)
        for sc, o, t in ((plane, (0.3, 2.0, 0.2), (0.1, 0.0, -0.4)), (sphere, (2.0, 1.5, 1.0), (0.0, 0.0, 0.0)),
                         (tri, (0.9, 2.0, 0.6), (0.8, 0.0, 0.7))):
            n0 = _hit(sc, o, t, p_off)[R_NS]
            rec = _hit(sc, o, t)
            n1 = rec[R_NS]
            if abs(np.linalg.norm(n1) - 1) > 2e-6 or abs(n1 @ n0 - cos) > 3e-5:
                return False
            if sc is tri:  # the tangent frame is the interpolated tangent: r along it
                if abs(n1 @ _unit(rec[R_TAN]) - e[0] / np.linalg.norm(e)) > 3e-5:
                    return False
    return True


def test_image_normal_map_decodes_the_tangent_frame():
    assert _normal_map_ok()
    with oracle_ffi.planted("normal_map_b"):
        assert not _normal_map_ok()


def fbm_fp64(freq, gain, lo, hi, octaves, x, y, z):
    """fBm over octaves of simplex noise at doubling frequencies, rescaled by the amplitude sum into [lo, hi];
    the octave arguments are rounded to float as the kernel's float products are."""
    amp, total, noise = 1.0, 0.0, 0.0
    f = np.float32(freq)
    for _ in range(octaves):
        noise += simplex3_fp64(float(np.float32(x) * f), float(np.float32(y) * f), float(np.float32(z) * f)) * amp
        total += amp
        amp *= gain
        f = np.float32(f * 2)
    noise /= total
    return noise * (hi - lo) / 2 + (hi + lo) / 2


def _noise_normal_ok():
    amount, scale = 0.4, 1.5
    b = SceneBuilder()
    b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), b.surface(3, noise_amount=amount, noise_scale=scale))
    sc = b.build(default_camera())
    rng = np.random.default_rng(6)
    for x, z in rng.uniform(-3, 3, (25, 2)):
        rec = _hit(sc, (x + 0.2, 2.0, z - 0.1), (x, 0.0, z))
        p = rec[R_P].astype(np.float32)
        q = p * np.float32(1 / scale)
        f = [fbm_fp64(scale, 0.5, -0.26, 0.26, 16, *args) for args in ((q[0], q[1], q[2]), (q[1], q[2], q[0]),
                                                                        (q[2], q[0], q[1]))]
        want = _unit(np.array([0.0, 1.0, 0.0]) + amount * np.array(f))
        if not np.allclose(rec[R_NS], want, atol=2e-5):
            return False
    # amount 0 or normal mapping off: the shading normal is the plane's
    p = dfa.default_params()
    p.normal_mapping = 0
    return np.allclose(_hit(sc, (0.2, 2.0, 0.1), (0.0, 0.0, 0.0), p)[R_NS], (0.0, 1.0, 0.0))


def test_noise_normal_map_is_normalised_fbm_offset():
    assert _noise_normal_ok()
    with oracle_ffi.planted("noise_normal"):
        assert not _noise_normal_ok()


# ---------------------------------------------------------------- Bullseye and fBm textures

def _rec(p):
    r = np.zeros(27, np.float32)
    r[2:5] = p
    return r


def _bytes(c):
    return np.array([(c >> 16) & 0xFF, (c >> 8) & 0xFF, c & 0xFF])


def _bullseye_ok():
    a, b = (0.9, 0.2, 0.1), (0.1, 0.3, 0.8)
    tex = [2, 4, argb(*a), argb(*b)]
    qa = np.array([int(np.float32(x) * 255 + 0.5) for x in a])
    qb = np.array([int(np.float32(x) * 255 + 0.5) for x in b])
    d = _unit([0.3, -0.5, 0.8])
    for radius, ring_a in ((1.0, False), (3.0, True), (5.0, False), (2.2, True), (6.2, True), (4.4, False)):
        got = _bytes(oracle_ffi.texture_color(tex, 0, _rec(radius * d)))
        if np.abs(got - (qa if ring_a else qb)).max() > 1:
            return False
    grey = [2, 4, argb(0.5, 0.5, 0.5), argb(0.5, 0.5, 0.5)]
    g = np.float32(128 / 255)
    return (_bytes(oracle_ffi.texture_color(grey, 0, _rec(3.0 * d)))[0] == int(np.float32(g * np.float32(0.8)) * 255 + 0.5)
            and _bytes(oracle_ffi.texture_color(grey, 0, _rec(1.0 * d)))[0] == int(np.float32(g * np.float32(1.2)) * 255 + 0.5))


def test_bullseye_rings():
    assert _bullseye_ok()
    with oracle_ffi.planted("bullseye"):
        assert not _bullseye_ok()


def _fbm_texture_ok():
    add, mul, freq, gain, octaves = (0.3, 0.2, 0.1), (0.5, 0.6, 0.7), 0.8, 0.5, 6
    tex = [5, 7, argb(*add), argb(*mul), freq, gain, octaves]
    qa = np.array([int(np.float32(x) * 255 + 0.5) / 255 for x in add])
    qm = np.array([int(np.float32(x) * 255 + 0.5) / 255 for x in mul])
    rng = np.random.default_rng(7)
    for p in rng.uniform(-5, 5, (40, 3)).astype(np.float32):
        n = fbm_fp64(freq, gain, 0.0, 1.0, octaves, *p)
        want = np.floor(np.clip(n * qm + qa, 0, 1) * 255 + 0.5)
        if np.abs(_bytes(oracle_ffi.texture_color(tex, 0, _rec(p))) - want).max() > 1:
            return False
    return True


def test_fbm_texture_closed_form():
    assert _fbm_texture_ok()
    with oracle_ffi.planted("fbm_texture"):
        assert not _fbm_texture_ok()
