"""CPU: the oracle's FP64 adjoint of the differentiable path tracer, pinned before it checks the GPU.

The oracle differentiates its own path tracer in forward mode, in double: dual numbers through the
sky (Sky.set and the Perez evaluation) and through the sun cone sample, and a throughput tangent along
every path.  Pins:
* the dual sky's value is the forward's sky; its derivatives agree with central differences of its own
  value (double, tight) and of the forward's float sky (Richardson-extrapolated, loose);
* the dual sun cosine at the cone's centre (u = 0) is the closed form n . w(elevation);
* in differentiable mode at the texture values the film is the plain render's, bit for bit;
* the loss gradient agrees with central differences of the oracle's own loss (common random numbers,
  Russian roulette off) for albedo groups, emission, turbidity and, with the sun off, elevation;
* misuse (Russian roulette on, a sky that is not Sky.set of the parameters, a wrong gradient length)
  is refused.
The workload is the differentiation scene (347,680 triangles in 64 albedo groups, 4 area lights) at
64 x 64 with a 16 x 16 crop.
"""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi as O

AZ = 30.0
W = H = 64
CROP = (24, 8, 16, 16)


@pytest.fixture(scope="module")
def setup():
    sc = dfa.build_config(5)
    cam = sc.camera.copy()
    cam[16], cam[17] = W, H
    s = sc.with_camera(cam)
    p = s.copy_params()
    p.math_mode = dfa.MATH_EXACT
    sa, se, na, ne = dfa.diff_maps(s)
    r = np.random.default_rng(501)
    th = {"albedo": r.uniform(0.1, 0.9, (na, 3)).astype(np.float32),
          "emission": r.uniform(1.0, 20.0, ne).astype(np.float32),
          "turbidity": float(r.uniform(2.0, 6.0)), "elevation": float(r.uniform(10.0, 70.0))}
    seeds = np.random.default_rng(3).integers(0, 2 ** 48, W * H).astype(np.int64)
    target = np.zeros(W * H * 3, np.float32)
    target.reshape(H, W, 3)[:, :, :] = (0.3, 0.25, 0.2)
    return s, p, (sa, se), th, seeds, target


def _loss(setup, th, sun=True, passes=2):
    s, p, maps, _, seeds, target = setup
    sky = dfa.sky_at(th["elevation"], AZ, th["turbidity"])
    sky.is_sun_active = 1 if sun else 0
    return O.loss_gradient(s, p, sky, maps, th["albedo"], th["emission"], th["turbidity"], th["elevation"], AZ,
                           seeds, target, passes, CROP)


def _with(th, key, index, value):
    t = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in th.items()}
    if index is None:
        t[key] = value
    else:
        t[key].reshape(-1)[index] = value
    return t


def _fd(setup, th, key, index, h, sun=True, richardson=False):
    v0 = th[key] if index is None else float(th[key].reshape(-1)[index])

    def cd(step):
        return (_loss(setup, _with(th, key, index, v0 + step), sun)[0] -
                _loss(setup, _with(th, key, index, v0 - step), sun)[0]) / (2.0 * step)

    return (4.0 * cd(h) - cd(2.0 * h)) / 3.0 if richardson else cd(h)


def _directions(n, seed):
    r = np.random.default_rng(seed)
    d = r.normal(size=(n, 3))
    d[:, 1] = np.abs(d[:, 1]) + 0.05
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("T,E", [(2.2, 15.0), (4.0, 45.0), (5.8, 68.0)])
def test_dual_sky_is_the_forward_sky_and_its_derivatives(T, E):
    sky = dfa.sky_at(E, AZ, T)
    for d in _directions(40, int(T * 10 + E)):
        v, dt, de = O.sky_dual(sky, T, E, AZ, d)
        fwd = O.sky_color(sky, d).astype(np.float64)
        np.testing.assert_allclose(v, fwd, rtol=5e-5)
        h = 1e-4
        cd_t = (O.sky_dual(sky, T + h, E, AZ, d)[0] - O.sky_dual(sky, T - h, E, AZ, d)[0]) / (2 * h)
        cd_e = (O.sky_dual(sky, T, E + h, AZ, d)[0] - O.sky_dual(sky, T, E - h, AZ, d)[0]) / (2 * h)
        assert np.max(np.abs(cd_t - dt)) <= 1e-6 * np.max(np.abs(dt))
        assert np.max(np.abs(cd_e - de)) <= 1e-6 * np.max(np.abs(de))

        # the forward's own float sky, set up at parameter offsets (Richardson over h, 2h)
        def fwd_cd(f, step):
            return (f(step) - f(-step)) / (2 * step)

        ft = lambda s: O.sky_color(dfa.sky_at(E, AZ, T + s), d).astype(np.float64)
        fe = lambda s: O.sky_color(dfa.sky_at(E + s, AZ, T), d).astype(np.float64)
        rt = (4 * fwd_cd(ft, 0.05) - fwd_cd(ft, 0.1)) / 3
        re = (4 * fwd_cd(fe, 0.25) - fwd_cd(fe, 0.5)) / 3
        assert np.max(np.abs(rt - dt)) <= 2e-3 * np.max(np.abs(dt))
        assert np.max(np.abs(re - de)) <= 2e-3 * np.max(np.abs(de))


def test_dual_sky_below_the_horizon_is_constant():
    sky = dfa.sky_at(30.0, AZ, 3.0)
    assert O.sky_dual(sky, 3.0, 30.0, AZ, (0.0, -0.5, np.sqrt(0.75))) is None


def test_dual_sun_cosine_closed_form_and_differences():
    T = 3.0
    r = np.random.default_rng(9)
    for E in (12.0, 35.0, 62.0):
        sky = dfa.sky_at(E, AZ, T)
        e, a = np.radians(E), np.radians(AZ)
        w = np.array([np.cos(e) * np.sin(a), np.sin(e), np.cos(e) * np.cos(a)])
        dw = np.array([-np.sin(e) * np.sin(a), np.cos(e), -np.sin(e) * np.cos(a)]) * np.pi / 180.0
        for _ in range(10):
            p = r.uniform(-5, 5, 3).astype(np.float32)
            n = r.normal(size=3)
            n = (n / np.linalg.norm(n)).astype(np.float32)
            # u = 0: the cone's centre, the sun direction itself
            c, dc = O.sun_cosine_dual(sky, T, E, AZ, p, n, 0.0, float(r.uniform()))
            assert abs(c - np.dot(w, n)) <= 1e-12
            assert abs(dc - np.dot(dw, n)) <= 1e-12
            u, v = float(r.uniform()), float(r.uniform())
            c, dc = O.sun_cosine_dual(sky, T, E, AZ, p, n, u, v)
            assert -1.0 <= c <= 1.0
            # within the sun's cone (radius 500 at distance ~10000) of n . w
            assert abs(c - np.dot(w, n)) <= 0.051
            h = 1e-4
            cd = (O.sun_cosine_dual(sky, T, E + h, AZ, p, n, u, v)[0] -
                  O.sun_cosine_dual(sky, T, E - h, AZ, p, n, u, v)[0]) / (2 * h)
            assert abs(cd - dc) <= 1e-7


def test_differentiable_forward_at_texture_values_is_the_plain_render(setup):
    s, p, maps, _, seeds, _ = setup
    sa, se = maps
    recip = np.float32(1.0) / np.float32(255.0)
    albedo = np.zeros((int(sa.max()) + 1, 3), np.float32)
    emission = np.zeros(int(se.max()) + 1, np.float32)
    for rec in range(s.surfaces.size // 6):
        for slot, field in ((sa[rec], 1), (se[rec], 2)):
            if slot < 0:
                continue
            off = int(s.surfaces[rec * 6 + field])
            c = int(np.float32(s.textures[off + 2]))
            rgb = [np.float32((c >> sh) & 0xFF) * recip for sh in (16, 8, 0)]
            if field == 1:
                albedo[slot] = rgb
            else:
                emission[slot] = rgb[0]
    th = {"albedo": albedo, "emission": emission, "turbidity": 3.0, "elevation": 45.0}
    _, _, film = _loss(setup, th, passes=3)
    plain = s.with_camera(s.camera)
    plain.sky = dfa.sky_at(45.0, AZ, 3.0)
    st = O.OracleState(W, H, seeds)
    O.render(plain, p, st, 3, clear_first=True)
    x0, y0, cw, ch = CROP
    ref = st.film.reshape(H, W, 3)[y0:y0 + ch, x0:x0 + cw].reshape(-1)
    np.testing.assert_array_equal(film, ref)


def _close(g, fd, gmax, what, rel):
    assert abs(g - fd) <= rel * abs(fd) + 1e-6 * gmax, (what, g, fd)


def test_fp64_adjoint_matches_central_differences(setup):
    th = setup[3]
    na, ne = th["albedo"].shape[0], th["emission"].size
    loss, g, _ = _loss(setup, th)
    assert loss > 0.0
    gmax = float(np.max(np.abs(g)))
    ga = g[:3 * na]
    for i in np.argsort(-np.abs(ga))[:10]:
        _close(ga[i], _fd(setup, th, "albedo", int(i), 1e-3), gmax, "albedo %d" % i, 2e-3)
    hit = 0
    for j in range(ne):
        fd = _fd(setup, th, "emission", j, 1e-3 * float(th["emission"][j]))
        _close(g[3 * na + j], fd, gmax, "emission %d" % j, 2e-3)
        hit += fd != 0.0
    assert hit > 0
    _close(g[-2], _fd(setup, th, "turbidity", None, 0.05, richardson=True), gmax, "turbidity", 1e-2)


def test_fp64_adjoint_elevation_through_the_sky(setup):
    """Sun light off: the elevation acts through the sky radiance only, a smooth function."""
    th = setup[3]
    _, g, _ = _loss(setup, th, sun=False)
    gmax = float(np.max(np.abs(g)))
    fd = _fd(setup, th, "elevation", None, 0.05, sun=False)
    assert abs(fd) > 1e-3 * gmax
    _close(g[-1], fd, gmax, "elevation", 2e-3)
    _close(g[-2], _fd(setup, th, "turbidity", None, 0.05, sun=False, richardson=True), gmax, "turbidity", 1e-2)


def test_fp64_adjoint_refuses_misuse(setup):
    s, p, maps, th, seeds, target = setup
    sky = dfa.sky_at(th["elevation"], AZ, th["turbidity"])
    args = (maps, th["albedo"], th["emission"], th["turbidity"], th["elevation"], AZ, seeds, target, 1, CROP)
    rr = s.copy_params()
    rr.math_mode = dfa.MATH_EXACT
    rr.pt_depth_russian_roulette = 3
    with pytest.raises(RuntimeError):
        O.loss_gradient(s, rr, sky, *args)
    with pytest.raises(RuntimeError):  # a sky that is not Sky.set of these parameters
        O.loss_gradient(s, p, dfa.sky_at(th["elevation"] + 5.0, AZ, th["turbidity"]), *args)
    with pytest.raises(RuntimeError):  # crop outside the frame
        O.loss_gradient(s, p, sky, *args[:-1], (60, 60, 16, 16))
