"""GPU: the differentiable path tracer and its path-replay adjoint (df_loss_gradient).

Bars:
* the differentiable forward at parameters equal to the texture values (albedo groups, emissions and
  the checkerboard ground's two colours) is the ordinary render, bit for bit;
* the adjoint gradient of L = sum (film - target)^2 with respect to albedo groups, emission,
  checkerboard colours, turbidity and (sun off) the sun's elevation and azimuth agrees with central
  finite differences of the loss - same seeds for every render (common random numbers), Russian
  roulette off - within 2 % (for gradient entries at least 1e-3 of the largest); the sun angles with
  the sun's direct light on are left out because moving the sun moves shadow boundaries, a visibility
  term that path replay (interior term only) does not include; at 64 x 64 x 4 samples for every kind of
  entry, and at the full workload (512 x 512 x 32 samples) for the largest entries of each kind, where the
  adjoint must also take at most 3x the forward's time;
* on a 64 x 64 crop of the full 512 x 512 frame in exact math, the film is the oracle's bit for bit and
  every gradient entry, the sun angles with the sun on included, agrees with the oracle's FP64
  forward-mode gradient (dual numbers through the sky and the sun sample on both sides) within 0.1 %
  (plus 1e-5 of the largest entry);
* the torch binding's gradients are the C ABI's; misuse is refused.
The workload is the differentiation scene (347,680 triangles in 64 albedo groups, 4 area lights, a
checkerboard ground) rendered at 64 x 64, and at 512 x 512 for the crop.
"""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi

pytestmark = pytest.mark.gpu

SEED = 77
AZIMUTH = 30.0


def _sized(scene, w, h):
    cam = scene.camera.copy()
    cam[16], cam[17] = w, h
    return scene.with_camera(cam)


@pytest.fixture(scope="module")
def diff_scene():
    return dfa.build_config(5)


def _kernel(sc, w, h):
    s = _sized(sc, w, h)
    k = dfa.HIPRendererKernel(s)
    p = s.copy_params()
    p.math_mode = dfa.MATH_FAST
    p.pt_engine = dfa.PT_MEGAKERNEL
    k.setParams(p)
    k.update(w, h)
    return s, k


def _maps(s):
    sa, se, na, ne = dfa.diff_maps(s)
    sc, nc, _ = dfa.diff_checker_map(s)
    return (sa, se, sc), na, ne, nc


def _params(n_albedo, n_emission, seed, n_checker=1):
    """The differentiation workload's parameter distribution: albedo and checkerboard colours in [0.1, 0.9],
    emission in [1, 20], turbidity in [2, 6], sun elevation in [10, 70] and azimuth in [0, 90] degrees."""
    r = np.random.default_rng(seed)
    return {"albedo": r.uniform(0.1, 0.9, (n_albedo, 3)).astype(np.float32),
            "emission": r.uniform(1.0, 20.0, n_emission).astype(np.float32),
            "turbidity": float(r.uniform(2.0, 6.0)), "elevation": float(r.uniform(10.0, 70.0)),
            "checker": r.uniform(0.1, 0.9, (n_checker, 2, 3)).astype(np.float32),
            "azimuth": float(r.uniform(0.0, 90.0))}


def _set(k, maps, th, sun=True):
    k.setDifferentiable(maps[0], maps[1], th["albedo"], th["emission"], th["turbidity"], th["elevation"],
                        th.get("azimuth", AZIMUTH), sun_active=sun, surface_checker=maps[2], checker=th["checker"])


def _loss(k, maps, th, target, passes, grad=False, sun=True, seed=SEED):
    _set(k, maps, th, sun)
    k.seed(seed)
    return k.lossGradient(target, passes, gradient=grad)


def _target(k, maps, th, w, h, passes=8, sun=True):
    _loss(k, maps, th, np.zeros(w * h * 3, np.float32), passes, sun=sun, seed=5)
    return k.getFilm()[0].copy()


def test_diff_forward_at_texture_values_is_the_plain_render(gpu, diff_scene):
    w, h, passes = 48, 40, 3
    s, k = _kernel(diff_scene, w, h)
    try:
        maps, na, ne, _ = _maps(s)
        sa, se, _ = maps
        recip = np.float32(1.0) / np.float32(255.0)
        albedo = np.zeros((na, 3), np.float32)
        emission = np.zeros(ne, np.float32)
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
                    assert rgb[0] == rgb[1] == rgb[2]
                    emission[slot] = rgb[0]
        th = {"albedo": albedo, "emission": emission, "turbidity": 3.0, "elevation": 45.0, "azimuth": AZIMUTH,
              "checker": dfa.diff_checker_map(s)[2]}
        _loss(k, maps, th, np.zeros(w * h * 3, np.float32), passes)
        film = k.getFilm()[0].copy()
        seeds = k.getSeeds()
        # the checkerboard colours swapped: the ground's pixels change, so the check above is not vacuous
        th2 = dict(th, checker=th["checker"][:, ::-1].copy())
        _loss(k, maps, th2, np.zeros(w * h * 3, np.float32), passes)
        swapped = k.getFilm()[0].copy()
    finally:
        k.dispose()
    plain = s.with_camera(s.camera)
    plain.sky = dfa.sky_at(45.0, AZIMUTH, 3.0)
    k2 = dfa.HIPRendererKernel(plain)
    try:
        p = plain.copy_params()
        p.math_mode = dfa.MATH_FAST
        p.pt_engine = dfa.PT_MEGAKERNEL
        k2.setParams(p)
        k2.update(w, h)
        k2.seed(SEED)
        k2.clear()
        k2.execute(passes)
        np.testing.assert_array_equal(film, k2.getFilm()[0])
        np.testing.assert_array_equal(seeds, k2.getSeeds())
    finally:
        k2.dispose()
    assert not np.array_equal(swapped, film)


def _fd(k, maps, th, target, passes, key, index, hstep, sun=True):
    def at(v):
        t = {kk: (vv.copy() if isinstance(vv, np.ndarray) else vv) for kk, vv in th.items()}
        if index is None:
            t[key] = v
        else:
            t[key].reshape(-1)[index] = v
        return _loss(k, maps, t, target, passes, sun=sun)[0]

    v0 = th[key] if index is None else float(th[key].reshape(-1)[index])
    lp, lm = at(v0 + hstep), at(v0 - hstep)
    return (lp - lm) / (2.0 * hstep), (lp, lm)


def _check(grad, fd, gmax, what):
    fd, losses = fd
    if abs(fd) < 1e-3 * gmax and abs(grad) < 1e-3 * gmax:
        return
    assert abs(grad - fd) <= 0.02 * abs(fd) + 1e-4 * gmax, (what, grad, fd, losses)


def test_adjoint_matches_central_differences(gpu, diff_scene):
    w = h = 64
    passes = 4
    s, k = _kernel(diff_scene, w, h)
    try:
        maps, na, ne, nc = _maps(s)
        target = _target(k, maps, _params(na, ne, 500), w, h)
        th = _params(na, ne, 501)
        loss, g, fms, ams = _loss(k, maps, th, target, passes, grad=True)
        assert loss > 0.0 and fms > 0.0 and ams > 0.0
        assert g.size == 3 * na + ne + 6 * nc + 3
        gmax = float(np.max(np.abs(g)))
        ga = g[:3 * na]
        # the albedo entries with the largest gradients, plus a few small ones
        order = np.argsort(-np.abs(ga))
        picks = list(order[:6]) + list(order[len(order) // 2:len(order) // 2 + 2])
        for i in picks:
            _check(ga[i], _fd(k, maps, th, target, passes, "albedo", int(i), 1e-3), gmax, "albedo %d" % i)
        for j in range(ne):
            _check(g[3 * na + j], _fd(k, maps, th, target, passes, "emission", j, 1e-3 * float(th["emission"][j])),
                   gmax, "emission %d" % j)
        gc = g[3 * na + ne:3 * na + ne + 6 * nc]
        assert np.count_nonzero(np.abs(gc) > 1e-3 * gmax) >= 3  # the ground is in view
        for i in range(6 * nc):
            _check(gc[i], _fd(k, maps, th, target, passes, "checker", i, 2e-3), gmax, "checker %d" % i)
        _check(g[-3], _fd(k, maps, th, target, passes, "turbidity", None, 1e-2), gmax, "turbidity")
    finally:
        k.dispose()


def test_adjoint_is_repeatable(gpu, diff_scene):
    """The same gradient twice from the same seeds: the block sums are reordered by the LDS atomics only, so
    the entries agree to double rounding (a build whose adjoint read something other than its inputs once
    gave entries 1-3 % off that also changed from run to run)."""
    w = h = 64
    s, k = _kernel(diff_scene, w, h)
    try:
        maps, na, ne, nc = _maps(s)
        target = _target(k, maps, _params(na, ne, 500), w, h)
        th = _params(na, ne, 501)
        runs = [_loss(k, maps, th, target, 4, grad=True)[1].copy() for _ in range(3)]
    finally:
        k.dispose()
    scale = float(np.max(np.abs(runs[0])))
    for g in runs[1:]:
        np.testing.assert_allclose(g, runs[0], rtol=1e-9, atol=1e-12 * scale)


def test_adjoint_sun_angles_through_the_sky(gpu, diff_scene):
    """Sun direct light off: the elevation and azimuth act through the sky model only (smooth in the
    parameters)."""
    w = h = 64
    passes = 4
    s, k = _kernel(diff_scene, w, h)
    try:
        maps, na, ne, _ = _maps(s)
        target = _target(k, maps, _params(na, ne, 500), w, h, sun=False)
        th = _params(na, ne, 502)
        loss, g, _, _ = _loss(k, maps, th, target, passes, grad=True, sun=False)
        gmax = float(np.max(np.abs(g)))
        for key, idx, step in (("elevation", -2, 0.1), ("azimuth", -1, 0.2)):
            fd = _fd(k, maps, th, target, passes, key, None, step, sun=False)
            _check(g[idx], fd, gmax, key)
            assert abs(fd[0]) > 1e-4 * gmax, key  # the check is not vacuous
        _check(g[-3], _fd(k, maps, th, target, passes, "turbidity", None, 1e-2, sun=False), gmax, "turbidity")
    finally:
        k.dispose()


def test_adjoint_at_the_full_differentiation_workload(gpu, diff_scene):
    """512 x 512 at 32 samples per pixel, as the benchmark runs it: the largest albedo, emission and
    checkerboard entries and turbidity against central differences, with the sun on; then the sun angles
    with the sun's direct light off."""
    w = h = 512
    passes = 32
    s, k = _kernel(diff_scene, w, h)
    try:
        maps, na, ne, nc = _maps(s)
        for sun in (True, False):
            target = _target(k, maps, _params(na, ne, 500), w, h, passes=passes, sun=sun)
            th = _params(na, ne, 501)
            loss, g, fms, ams = _loss(k, maps, th, target, passes, grad=True, sun=sun)
            assert ams <= 3.0 * fms, (ams, fms)
            gmax = float(np.max(np.abs(g)))
            if sun:
                ga = g[:3 * na]
                for i in np.argsort(-np.abs(ga))[:3]:
                    _check(ga[i], _fd(k, maps, th, target, passes, "albedo", int(i), 1e-3), gmax, "albedo %d" % i)
                j = int(np.argmax(np.abs(g[3 * na:3 * na + ne])))
                _check(g[3 * na + j], _fd(k, maps, th, target, passes, "emission", j, 1e-3 * float(th["emission"][j])),
                       gmax, "emission %d" % j)
                gc = g[3 * na + ne:3 * na + ne + 6 * nc]
                for i in np.argsort(-np.abs(gc))[:2]:
                    _check(gc[i], _fd(k, maps, th, target, passes, "checker", int(i), 2e-3), gmax, "checker %d" % i)
                _check(g[-3], _fd(k, maps, th, target, passes, "turbidity", None, 1e-2), gmax, "turbidity")
            else:
                for key, idx, step in (("elevation", -2, 0.1), ("azimuth", -1, 0.2)):
                    fd = _fd(k, maps, th, target, passes, key, None, step, sun=False)
                    _check(g[idx], fd, gmax, key)
                    assert abs(fd[0]) > 1e-4 * gmax, key
    finally:
        k.dispose()


def test_adjoint_matches_the_fp64_cpu_adjoint_on_a_crop(gpu, diff_scene):
    w = h = 512
    x0, y0, cw, ch = crop = (160, 300, 64, 64)  # the sphere's lower edge over the checkerboard ground
    passes = 4
    s, k = _kernel(diff_scene, w, h)
    try:
        p = s.copy_params()
        p.math_mode = dfa.MATH_EXACT
        p.pt_engine = dfa.PT_MEGAKERNEL
        k.setParams(p)
        maps, na, ne, nc = _maps(s)
        other = _target(k, maps, _params(na, ne, 500), w, h)
        th = _params(na, ne, 501)
        # the target is the render at th outside the crop, so only the crop carries loss
        _loss(k, maps, th, np.zeros(w * h * 3, np.float32), passes)
        film_th = k.getFilm()[0].copy()
        target = film_th.copy()
        view = target.reshape(h, w, 3)
        view[y0:y0 + ch, x0:x0 + cw] = other.reshape(h, w, 3)[y0:y0 + ch, x0:x0 + cw]
        _set(k, maps, th)
        k.seed(SEED)
        seeds = k.getSeeds().copy()
        loss, g, _, _ = k.lossGradient(target, passes, gradient=True)
        film = k.getFilm()[0].copy()
        np.testing.assert_array_equal(film, film_th)
    finally:
        k.dispose()
    sky = dfa.sky_at(th["elevation"], th["azimuth"], th["turbidity"])
    loss_o, g_o, film_o = oracle_ffi.loss_gradient(s, p, sky, maps, th["albedo"], th["emission"], th["turbidity"],
                                                   th["elevation"], th["azimuth"], seeds, target, passes, crop,
                                                   checker=th["checker"])
    np.testing.assert_array_equal(film.reshape(h, w, 3)[y0:y0 + ch, x0:x0 + cw].reshape(-1), film_o)
    assert abs(loss - loss_o) <= 1e-6 * loss_o  # the GPU squares float differences, the oracle double ones
    gmax = float(np.max(np.abs(g_o)))
    gc = g_o[3 * na + ne:3 * na + ne + 6 * nc]
    assert np.count_nonzero(g_o[:3 * na]) > 10 and np.count_nonzero(gc) == 6
    assert g_o[-3] != 0.0 and g_o[-2] != 0.0 and g_o[-1] != 0.0
    bad = np.abs(g - g_o) > 1e-3 * np.abs(g_o) + 1e-5 * gmax
    assert not bad.any(), [(int(i), g[i], g_o[i]) for i in np.flatnonzero(bad)]


def test_torch_autograd_binding(gpu, diff_scene):
    torch = pytest.importorskip("torch")
    from dayflower_amd.autograd import render_loss

    w = h = 32
    s, k = _kernel(diff_scene, w, h)
    try:
        maps, na, ne, nc = _maps(s)
        target = _target(k, maps, _params(na, ne, 500), w, h)
        th = _params(na, ne, 503)
        albedo = torch.tensor(th["albedo"], dtype=torch.float64, requires_grad=True)
        emission = torch.tensor(th["emission"], dtype=torch.float64, requires_grad=True)
        checker = torch.tensor(th["checker"], dtype=torch.float64, requires_grad=True)
        turb = torch.tensor(th["turbidity"], dtype=torch.float64, requires_grad=True)
        elev = torch.tensor(th["elevation"], dtype=torch.float64, requires_grad=True)
        azim = torch.tensor(th["azimuth"], dtype=torch.float64, requires_grad=True)
        loss = render_loss(k, maps, target, 2, SEED, albedo, emission, turb, elev, azim, checker=checker)
        (3.0 * loss).backward()
        _, g, _, _ = _loss(k, maps, th, target, 2, grad=True)
        np.testing.assert_allclose(albedo.grad.numpy().reshape(-1), 3.0 * g[:3 * na], rtol=1e-6)
        np.testing.assert_allclose(emission.grad.numpy(), 3.0 * g[3 * na:3 * na + ne], rtol=1e-6)
        np.testing.assert_allclose(checker.grad.numpy().reshape(-1), 3.0 * g[3 * na + ne:3 * na + ne + 6 * nc],
                                   rtol=1e-6)
        for t, i in ((turb, -3), (elev, -2), (azim, -1)):
            assert abs(t.grad.item() - 3.0 * g[i]) <= 1e-6 * abs(g[i]) + 1e-30
    finally:
        k.dispose()


def test_adjoint_argument_checks(gpu, diff_scene):
    s, k = _kernel(diff_scene, 16, 16)
    try:
        maps, na, ne, nc = _maps(s)
        sa, se, sc = maps
        th = _params(na, ne, 1)
        with pytest.raises(dfa.HIPError):  # not in differentiable mode yet
            k.lossGradient(np.zeros(16 * 16 * 3, np.float32), 1)
        _set(k, maps, th)
        with pytest.raises(dfa.HIPError):
            k.lossGradient(np.zeros(10, np.float32), 1)
        p = s.copy_params()
        p.pt_depth_russian_roulette = 3  # Russian roulette on
        p.pt_engine = dfa.PT_MEGAKERNEL
        k.setParams(p)
        with pytest.raises(dfa.HIPError):
            k.lossGradient(np.zeros(16 * 16 * 3, np.float32), 1)
        loss, g, _, _ = k.lossGradient(np.zeros(16 * 16 * 3, np.float32), 1, gradient=False)
        assert g is None and loss > 0.0
        k.setParams(s.copy_params())
        with pytest.raises(dfa.HIPError):  # surface map of the wrong length
            k.setDifferentiable(sa[:-1], se[:-1], th["albedo"], th["emission"], 3.0, 45.0, AZIMUTH)
        bad = sc.copy()
        bad[int(np.flatnonzero(sa >= 0)[0])] = 0  # a checkerboard group on a surface with an albedo group
        with pytest.raises(dfa.HIPError):
            k.setDifferentiable(sa, se, th["albedo"], th["emission"], 3.0, 45.0, AZIMUTH, surface_checker=bad,
                                checker=th["checker"])
        # the sky must stay the one df_set_diff was given
        _set(k, maps, th)
        k.lossGradient(np.zeros(16 * 16 * 3, np.float32), 1)
        k.updateSunAndSky(dfa.sky_at(th["elevation"], th["azimuth"] + 10.0, th["turbidity"]))
        with pytest.raises(dfa.HIPError):
            k.lossGradient(np.zeros(16 * 16 * 3, np.float32), 1)
    finally:
        k.dispose()
