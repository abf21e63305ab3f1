"""Pins of the oracle's primary rays (doCreatePrimaryRay, GPURendererKernel.java:573-724) against what the
camera models fix: the pinhole's image-centre ray and projection, the tent filter's distribution, the thin lens
converging on its focal plane from a uniformly sampled disk, the fisheye's unit-disk cut-off and its
orthographic hemisphere mapping.  Each check also runs with a deliberate mistake switched on in the oracle and
must then fail."""

import numpy as np
import pytest
from scipy import stats

import dayflower_amd as dfa
from oracle import oracle_ffi

EYE = (1.0, 2.0, 9.0)
TARGET = (0.0, 0.5, 0.0)
RC, PT = 3, 2  # renderer types: ray caster (no pixel jitter), path tracer (tent jitter)


def _cam(w, h, fov=40.0, aperture=0.0, focal=10.0, lens=None):
    c = dfa.camera_setup(w, h, fov, EYE, TARGET, aperture, focal)
    if lens is not None:
        c[18] = lens
    return c


def _frame(c):
    c = c.astype(np.float64)
    return c[1:4], c[7:10], c[10:13], c[13:16], np.tan(np.radians(c[4]) / 2), np.tan(-np.radians(c[5]) / 2)


def _lcg(seed):
    return (seed * 0x5DEECE66D + 0xB) & ((1 << 48) - 1)


def _pinhole_dir(c, x, y, jx=0.5, jy=0.5):
    """The pinhole's ray direction in float64: through the point (sx0, sy0) of the image plane one unit along W,
    with pixel centres normalised by (res - 1) as the reference does (:635-636)."""
    _, u, v, w, tx, ty = _frame(c)
    sx0 = 2.0 * (x + jx) / (c[16] - 1.0) - 1.0
    sy0 = 2.0 * (y + jy) / (c[17] - 1.0) - 1.0
    d = w + tx * sx0 * u + ty * sy0 * v
    return d / np.linalg.norm(d)


def _check_centre_ray():
    c = _cam(256, 256)
    ray, _ = oracle_ffi.primary_ray(c, RC, 256, 256, 127, 127, 1)
    # (127 + 0.5) / 255 = 1/2 exactly: the ray leaves the eye along W
    return np.array_equal(ray[:3], c[1:4]) and np.allclose(ray[3:], c[13:16], atol=1e-6)


def test_pinhole_centre_pixel_looks_along_w():
    assert _check_centre_ray()
    with oracle_ffi.planted("camera_res"):
        assert not _check_centre_ray()


def test_pinhole_projection_closed_form():
    c = _cam(16, 12, fov=55.0)
    for y in range(12):
        for x in range(16):
            ray, seed = oracle_ffi.primary_ray(c, RC, 16, 12, x, y, 5)
            assert seed == 5  # no random numbers without jitter or lens
            np.testing.assert_allclose(ray[3:], _pinhole_dir(c, x, y), atol=2e-6)
            np.testing.assert_array_equal(ray[:3], c[1:4])


def _tent_ok(n=20000):
    """The path tracer's pixel jitter recovered from its rays is the tent on [-1, 1] (cdf (t+1)^2/2 below 0),
    independently in x and y, and each ray takes exactly two random numbers."""
    w, h, x, y = 64, 48, 20, 30
    c = _cam(w, h)
    _, u, v, ww, tx, ty = _frame(c)
    seeds = np.random.default_rng(1).integers(1, 1 << 48, n)
    jit = np.zeros((n, 2))
    for i, s in enumerate(seeds):
        ray, after = oracle_ffi.primary_ray(c, PT, w, h, x, y, int(s))
        assert after == _lcg(_lcg(int(s)))
        d = ray[3:].astype(np.float64)
        sx0 = (d @ u) / (d @ ww) / tx
        sy0 = (d @ v) / (d @ ww) / ty
        jit[i] = ((sx0 + 1) / 2 * (w - 1) - x, (sy0 + 1) / 2 * (h - 1) - y)
    cdf = lambda t: np.where(t < 0, (t + 1) ** 2 / 2, 1 - (1 - t) ** 2 / 2)
    ok = all(stats.kstest(jit[:, k], cdf).pvalue > 1e-3 for k in range(2))
    return ok and abs(np.corrcoef(jit[:, 0], jit[:, 1])[0, 1]) < 0.03


def test_path_tracer_pixel_jitter_is_the_tent():
    assert _tent_ok()
    with oracle_ffi.planted("camera_box"):
        assert not _tent_ok(4000)


def _thin_lens_ok(n=4000):
    """Rays of one pixel from a thin lens of radius a leave from points spread uniformly over the lens disk
    (r^2 / a^2 and the angle uniform) in the plane through the eye normal to W, and all pass through the one
    point the pinhole ray of that pixel reaches at the focal distance along W."""
    a, f = 0.3, 8.0
    c = _cam(64, 48, aperture=a, focal=f)
    eye, u, v, w, _, _ = _frame(c)
    x, y = 41, 9
    target = eye + f * _pinhole_dir(c, x, y) / (_pinhole_dir(c, x, y) @ w)
    seeds = np.random.default_rng(2).integers(1, 1 << 48, n)
    r2, ang = np.zeros(n), np.zeros(n)
    for i, s in enumerate(seeds):
        ray, after = oracle_ffi.primary_ray(c, RC, 64, 48, x, y, int(s))
        assert after == _lcg(_lcg(int(s)))  # the two lens draws
        o, d = ray[:3].astype(np.float64), ray[3:].astype(np.float64)
        off = o - eye
        if abs(off @ w) > 1e-6 or np.linalg.norm(off) > a * (1 + 1e-6):
            return False
        # distance from the focal point to the ray's line
        if np.linalg.norm(np.cross(target - o, d)) > 2e-5 * f:
            return False
        r2[i] = (off @ off) / (a * a)
        ang[i] = np.arctan2(off @ v, off @ u)
    return stats.kstest(r2, "uniform").pvalue > 1e-3 and stats.kstest(ang, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 1e-3


def test_thin_lens_focuses_and_samples_the_disk_uniformly():
    assert _thin_lens_ok()
    with oracle_ffi.planted("lens_radius"):
        assert not _thin_lens_ok(2000)


def _fisheye_ok():
    """Fisheye (lens 2) at a 90-degree field of view: no ray outside the unit disk of the normalised image
    coordinates; inside, the direction is (sx0 U + sy0 V' + sqrt(1 - r^2) W), the orthographic hemisphere
    mapping, so the angle to W has sine r."""
    n = 33
    c = _cam(n, n, fov=90.0, lens=2.0)
    eye, u, v, w, tx, ty = _frame(c)
    assert abs(tx - 1) < 1e-6 and abs(ty + 1) < 1e-6
    inside = 0
    for y in range(n):
        for x in range(n):
            sx0 = 2.0 * (x + 0.5) / (n - 1.0) - 1.0
            sy0 = 2.0 * (y + 0.5) / (n - 1.0) - 1.0
            r2 = sx0 * sx0 + sy0 * sy0
            ray, _ = oracle_ffi.primary_ray(c, RC, n, n, x, y, 3)
            if r2 > 1.0 + 1e-6:
                if ray is not None:
                    return False
                continue
            if r2 > 1.0 - 1e-6:
                continue  # on the rim float rounding decides
            if ray is None:
                return False
            inside += 1
            d = ray[3:].astype(np.float64)
            if not (abs(d @ u - sx0) < 3e-6 and abs(d @ v + sy0) < 3e-6 and abs(d @ w - np.sqrt(1 - r2)) < 3e-6):
                return False
    return inside > 0.7 * n * n


def test_fisheye_disk_and_hemisphere_mapping():
    assert _fisheye_ok()
    with oracle_ffi.planted("fisheye_w"):
        assert not _fisheye_ok()


def test_fisheye_pixels_outside_the_disk_render_black(scenes):
    """run() gives a pixel without a ray a black sample through filmSetColor (:302-306)."""
    sc = scenes(1)
    c = _cam(24, 24, fov=90.0, lens=2.0)
    small = sc.with_camera(c)
    p = small.copy_params()
    st = oracle_ffi.OracleState(24, 24, dfa.pixel_seeds(3, 24 * 24))
    oracle_ffi.render(small, p, st, 2)
    film = st.film.reshape(24, 24, 3)
    assert np.all(film[0, 0] == 0) and np.all(film[23, 0] == 0)
    assert st.film_samples.reshape(24, 24)[0, 0] == 1
    assert np.any(film[12, 12] > 0)
    assert st.primitive_offsets.reshape(24, 24)[0, 0] == -1
