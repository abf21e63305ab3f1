"""CPU: the oracle's FP64 adjoint of the differentiable path tracer, pinned before it checks the GPU.

The oracle differentiates its own path tracer in forward mode, in double: dual numbers through the
sky (Sky.set and the Perez evaluation) and through the sun cone sample, and a throughput tangent along
every path.  Pins:
* the dual sky's value is the forward's sky; its derivatives with respect to turbidity, sun elevation
  and sun azimuth agree with central differences of its own value (double, tight) and of the forward's
  float sky (Richardson-extrapolated, loose);
* the dual sun cosine at the cone's centre (u = 0) is the closed form n . w(elevation, azimuth), its
  two derivatives those of the closed form;
* in differentiable mode at the texture values (albedo groups, emissions and the checkerboard ground's
  two colours) the film is the plain render's, bit for bit;
* the loss gradient agrees with central differences of the oracle's own loss (common random numbers,
  Russian roulette off) for albedo groups, emission, checkerboard colours, turbidity and, with the sun
  off, elevation and azimuth;
* a planted mistake in the tile rule of a checkerboard group (the oracle's colours swapped on the
  tiles) fails the bit-exact check, so that check is not vacuous;
* misuse (Russian roulette on, a sky that is not Sky.set of the parameters, a checkerboard group on a
  surface without a checkerboard texture, a crop outside the frame) is refused.
The workload is the differentiation scene (347,680 triangles in 64 albedo groups, 4 area lights, a
checkerboard ground) at 64 x 64 with 16 x 16 crops.
"""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi as O

AZ = 30.0
W = H = 64
CROP = (24, 8, 16, 16)
GROUND = (8, 46, 16, 16)  # a crop of the checkerboard ground


@pytest.fixture(scope="module")
def setup():
    sc = dfa.build_config(5)
    cam = sc.camera.copy()
    cam[16], cam[17] = W, H
    s = sc.with_camera(cam)
    p = s.copy_params()
    p.math_mode = dfa.MATH_EXACT
    sa, se, na, ne = dfa.diff_maps(s)
    sc, nc, _ = dfa.diff_checker_map(s)
    assert nc == 1
    r = np.random.default_rng(501)
    th = {"albedo": r.uniform(0.1, 0.9, (na, 3)).astype(np.float32),
          "emission": r.uniform(1.0, 20.0, ne).astype(np.float32),
          "checker": r.uniform(0.1, 0.9, (nc, 2, 3)).astype(np.float32),
          "turbidity": float(r.uniform(2.0, 6.0)), "elevation": float(r.uniform(10.0, 70.0)),
          "azimuth": float(r.uniform(0.0, 90.0))}
    seeds = np.random.default_rng(3).integers(0, 2 ** 48, W * H).astype(np.int64)
    target = np.zeros(W * H * 3, np.float32)
    target.reshape(H, W, 3)[:, :, :] = (0.3, 0.25, 0.2)
    return s, p, (sa, se, sc), th, seeds, target


def _loss(setup, th, sun=True, passes=2, crop=CROP):
    s, p, maps, _, seeds, target = setup
    az = th.get("azimuth", AZ)
    sky = dfa.sky_at(th["elevation"], az, th["turbidity"])
    sky.is_sun_active = 1 if sun else 0
    return O.loss_gradient(s, p, sky, maps, th["albedo"], th["emission"], th["turbidity"], th["elevation"], az,
                           seeds, target, passes, crop, checker=th["checker"])


def _with(th, key, index, value):
    t = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in th.items()}
    if index is None:
        t[key] = value
    else:
        t[key].reshape(-1)[index] = value
    return t


def _fd(setup, th, key, index, h, sun=True, richardson=False, crop=CROP):
    v0 = th[key] if index is None else float(th[key].reshape(-1)[index])

    def cd(step):
        return (_loss(setup, _with(th, key, index, v0 + step), sun, crop=crop)[0] -
                _loss(setup, _with(th, key, index, v0 - step), sun, crop=crop)[0]) / (2.0 * step)

    return (4.0 * cd(h) - cd(2.0 * h)) / 3.0 if richardson else cd(h)


def _directions(n, seed):
    r = np.random.default_rng(seed)
    d = r.normal(size=(n, 3))
    d[:, 1] = np.abs(d[:, 1]) + 0.05
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("T,E,A", [(2.2, 15.0, 30.0), (4.0, 45.0, 150.0), (5.8, 68.0, 290.0)])
def test_dual_sky_is_the_forward_sky_and_its_derivatives(T, E, A):
    sky = dfa.sky_at(E, A, T)
    for d in _directions(40, int(T * 10 + E)):
        v, dt, de, da = O.sky_dual(sky, T, E, A, d)
        fwd = O.sky_color(sky, d).astype(np.float64)
        np.testing.assert_allclose(v, fwd, rtol=5e-5)
        h = 1e-4
        cd_t = (O.sky_dual(sky, T + h, E, A, d)[0] - O.sky_dual(sky, T - h, E, A, d)[0]) / (2 * h)
        cd_e = (O.sky_dual(sky, T, E + h, A, d)[0] - O.sky_dual(sky, T, E - h, A, d)[0]) / (2 * h)
        cd_a = (O.sky_dual(sky, T, E, A + h, d)[0] - O.sky_dual(sky, T, E, A - h, d)[0]) / (2 * h)
        assert np.max(np.abs(cd_t - dt)) <= 1e-6 * np.max(np.abs(dt))
        assert np.max(np.abs(cd_e - de)) <= 1e-6 * np.max(np.abs(de))
        assert np.max(np.abs(cd_a - da)) <= 1e-6 * np.max(np.abs(da)) + 1e-12

        # the forward's own float sky, set up at parameter offsets (Richardson over h, 2h)
        def fwd_cd(f, step):
            return (f(step) - f(-step)) / (2 * step)

        ft = lambda s: O.sky_color(dfa.sky_at(E, A, T + s), d).astype(np.float64)
        fe = lambda s: O.sky_color(dfa.sky_at(E + s, A, T), d).astype(np.float64)
        fa = lambda s: O.sky_color(dfa.sky_at(E, A + s, T), d).astype(np.float64)
        rt = (4 * fwd_cd(ft, 0.05) - fwd_cd(ft, 0.1)) / 3
        re = (4 * fwd_cd(fe, 0.25) - fwd_cd(fe, 0.5)) / 3
        ra = (4 * fwd_cd(fa, 1.0) - fwd_cd(fa, 2.0)) / 3  # wider steps: float noise over a smaller derivative
        assert np.max(np.abs(rt - dt)) <= 2e-3 * np.max(np.abs(dt))
        assert np.max(np.abs(re - de)) <= 2e-3 * np.max(np.abs(de))
        assert np.max(np.abs(ra - da)) <= 2e-3 * np.max(np.abs(da)) + 1e-6 * np.max(np.abs(v))


def test_dual_sky_azimuth_symmetry():
    """The sky depends on the azimuth only through the angle between the sun and the view direction: turning
    the sun and the view about the vertical by the same angle leaves the radiance and all three derivatives
    unchanged."""
    T, E = 3.5, 35.0
    for A in (10.0, 100.0):
        sky = dfa.sky_at(E, A, T)
        for d in _directions(12, int(A)):
            v0, _, de0, da0 = O.sky_dual(sky, T, E, A, d)
            c, s_ = np.cos(np.radians(20.0)), np.sin(np.radians(20.0))
            # azimuth is measured from +z towards +x: rotate the view by +20 degrees the same way
            rd = np.array([d[0] * c + d[2] * s_, d[1], -d[0] * s_ + d[2] * c], np.float32)
            v1, _, de1, da1 = O.sky_dual(dfa.sky_at(E, A + 20.0, T), T, E, A + 20.0, rd)
            np.testing.assert_allclose(v1, v0, rtol=1e-4)
            np.testing.assert_allclose(de1, de0, rtol=1e-3, atol=1e-6 * np.max(np.abs(de0)))
            np.testing.assert_allclose(da1, da0, rtol=1e-3, atol=1e-6 * np.max(np.abs(v0)))


def test_dual_sky_below_the_horizon_is_constant():
    sky = dfa.sky_at(30.0, AZ, 3.0)
    assert O.sky_dual(sky, 3.0, 30.0, AZ, (0.0, -0.5, np.sqrt(0.75))) is None


def test_dual_sun_cosine_closed_form_and_differences():
    T = 3.0
    r = np.random.default_rng(9)
    for E, A in ((12.0, AZ), (35.0, 200.0), (62.0, 75.0)):
        sky = dfa.sky_at(E, A, T)
        e, a = np.radians(E), np.radians(A)
        w = np.array([np.cos(e) * np.sin(a), np.sin(e), np.cos(e) * np.cos(a)])
        dwe = np.array([-np.sin(e) * np.sin(a), np.cos(e), -np.sin(e) * np.cos(a)]) * np.pi / 180.0
        dwa = np.array([np.cos(e) * np.cos(a), 0.0, -np.cos(e) * np.sin(a)]) * np.pi / 180.0
        for _ in range(10):
            p = r.uniform(-5, 5, 3).astype(np.float32)
            n = r.normal(size=3)
            n = (n / np.linalg.norm(n)).astype(np.float32)
            # u = 0: the cone's centre, the sun direction itself
            c, dce, dca = O.sun_cosine_dual(sky, T, E, A, p, n, 0.0, float(r.uniform()))
            assert abs(c - np.dot(w, n)) <= 1e-12
            assert abs(dce - np.dot(dwe, n)) <= 1e-12
            assert abs(dca - np.dot(dwa, n)) <= 1e-12
            u, v = float(r.uniform()), float(r.uniform())
            c, dce, dca = O.sun_cosine_dual(sky, T, E, A, p, n, u, v)
            assert -1.0 <= c <= 1.0
            # within the sun's cone (radius 500 at distance ~10000) of n . w
            assert abs(c - np.dot(w, n)) <= 0.051
            h = 1e-4
            cde = (O.sun_cosine_dual(sky, T, E + h, A, p, n, u, v)[0] -
                   O.sun_cosine_dual(sky, T, E - h, A, p, n, u, v)[0]) / (2 * h)
            cda = (O.sun_cosine_dual(sky, T, E, A + h, p, n, u, v)[0] -
                   O.sun_cosine_dual(sky, T, E, A - h, p, n, u, v)[0]) / (2 * h)
            assert abs(cde - dce) <= 1e-7
            assert abs(cda - dca) <= 1e-7


def _texture_values(s, maps):
    """The parameters equal to the scene's own textures: albedo groups and emissions from their constant
    textures, the checkerboard group's two colours from its checkerboard texture (byte / 255)."""
    sa, se, _ = maps
    recip = np.float32(1.0) / np.float32(255.0)
    albedo = np.zeros((int(sa.max()) + 1, 3), np.float32)
    emission = np.zeros(int(se.max()) + 1, np.float32)
    for rec in range(s.surfaces.size // 6):
        for slot, field in ((sa[rec], 1), (se[rec], 2)):
            if slot < 0:
                continue
            off = int(s.surfaces[rec * 6 + field])
            assert int(s.textures[off]) == 4  # constant texture
            c = int(np.float32(s.textures[off + 2]))
            rgb = [np.float32((c >> sh) & 0xFF) * recip for sh in (16, 8, 0)]
            if field == 1:
                albedo[slot] = rgb
            else:
                emission[slot] = rgb[0]
    checker = dfa.diff_checker_map(s)[2]
    assert not np.array_equal(checker[0, 0], checker[0, 1])  # two colours: no single-colour quirk
    return {"albedo": albedo, "emission": emission, "checker": checker, "turbidity": 3.0, "elevation": 45.0,
            "azimuth": AZ}


@pytest.mark.parametrize("crop", [CROP, GROUND])
def test_differentiable_forward_at_texture_values_is_the_plain_render(setup, crop):
    s, p, maps, _, seeds, _ = setup
    th = _texture_values(s, maps)
    _, _, film = _loss(setup, th, passes=3, crop=crop)
    plain = s.with_camera(s.camera)
    plain.sky = dfa.sky_at(45.0, AZ, 3.0)
    st = O.OracleState(W, H, seeds)
    O.render(plain, p, st, 3, clear_first=True)
    x0, y0, cw, ch = crop
    ref = st.film.reshape(H, W, 3)[y0:y0 + ch, x0:x0 + cw].reshape(-1)
    np.testing.assert_array_equal(film, ref)
    # a planted mistake in the checkerboard group's tile rule changes the pixels (of the sphere too, through
    # the light the ground reflects)
    with O.planted("checker_tile"):
        _, _, bad = _loss(setup, th, passes=3, crop=crop)
    assert not np.array_equal(bad, ref)


def _close(g, fd, gmax, what, rel):
    assert abs(g - fd) <= rel * abs(fd) + 1e-6 * gmax, (what, g, fd)


def test_fp64_adjoint_matches_central_differences(setup):
    th = setup[3]
    na, ne = th["albedo"].shape[0], th["emission"].size
    loss, g, _ = _loss(setup, th)
    assert loss > 0.0
    assert g.size == 3 * na + ne + 6 + 3
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
    _close(g[-3], _fd(setup, th, "turbidity", None, 0.05, richardson=True), gmax, "turbidity", 1e-2)


def test_fp64_adjoint_checker_colours(setup):
    """The ground's two checkerboard colours, on a crop of the ground: every one of the six entries."""
    th = setup[3]
    na, ne = th["albedo"].shape[0], th["emission"].size
    _, g, _ = _loss(setup, th, crop=GROUND)
    gmax = float(np.max(np.abs(g)))
    gc = g[3 * na + ne:3 * na + ne + 6]
    assert np.all(np.abs(gc) > 1e-3 * gmax)
    for i in range(6):
        _close(gc[i], _fd(setup, th, "checker", i, 1e-3, crop=GROUND), gmax, "checker %d" % i, 2e-3)


def test_fp64_adjoint_sun_angles_through_the_sky(setup):
    """Sun light off: the elevation and azimuth act through the sky radiance only, a smooth function."""
    th = setup[3]
    _, g, _ = _loss(setup, th, sun=False)
    gmax = float(np.max(np.abs(g)))
    for key, idx in (("elevation", -2), ("azimuth", -1)):
        fd = _fd(setup, th, key, None, 0.05, sun=False)
        assert abs(fd) > 1e-4 * gmax, key
        _close(g[idx], fd, gmax, key, 2e-3)
    _close(g[-3], _fd(setup, th, "turbidity", None, 0.05, sun=False, richardson=True), gmax, "turbidity", 1e-2)


def test_fp64_adjoint_refuses_misuse(setup):
    s, p, maps, th, seeds, target = setup
    sky = dfa.sky_at(th["elevation"], th["azimuth"], th["turbidity"])
    args = (th["albedo"], th["emission"], th["turbidity"], th["elevation"], th["azimuth"], seeds, target, 1, CROP)
    kw = {"checker": th["checker"]}
    O.loss_gradient(s, p, sky, maps, *args, **kw)
    rr = s.copy_params()
    rr.math_mode = dfa.MATH_EXACT
    rr.pt_depth_russian_roulette = 3
    with pytest.raises(RuntimeError):
        O.loss_gradient(s, rr, sky, maps, *args, **kw)
    with pytest.raises(RuntimeError):  # a sky that is not Sky.set of these parameters
        O.loss_gradient(s, p, dfa.sky_at(th["elevation"] + 5.0, th["azimuth"], th["turbidity"]), maps, *args, **kw)
    with pytest.raises(RuntimeError):  # nor at another azimuth
        O.loss_gradient(s, p, dfa.sky_at(th["elevation"], th["azimuth"] + 5.0, th["turbidity"]), maps, *args, **kw)
    with pytest.raises(RuntimeError):  # crop outside the frame
        O.loss_gradient(s, p, sky, maps, *args[:-1], (60, 60, 16, 16), **kw)
    sa, se, sc = maps
    bad = sc.copy()
    bad[int(np.flatnonzero(sa >= 0)[0])] = 0  # a checkerboard group on a surface with a constant texture
    with pytest.raises(RuntimeError):
        O.loss_gradient(s, p, sky, (sa, se, bad), *args, **kw)
