"""ctypes wrapper of liboracle.so - TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference renderer kernel
(df_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, always as the checker or the reported CPU
baseline; the product never imports it.
"""

import ctypes
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
STAMP_PATH = os.path.join(HERE, "liboracle.stamp")
_SOURCES = [os.path.join(HERE, "df_oracle.cpp"), os.path.join(HERE, "Makefile"),
            os.path.join(os.path.dirname(HERE), "include", "dayflower_types.h")]

_c_float_p = ctypes.POINTER(ctypes.c_float)


class OracleFrame(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("film", ctypes.c_void_p),
        ("film_samples", ctypes.c_void_p),
        ("image_float", ctypes.c_void_p),
        ("image_bgra", ctypes.c_void_p),
        ("primitive_offsets", ctypes.c_void_p),
        ("seeds", ctypes.c_void_p),
    ]


class OracleCounters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("rays", "samples", "box_tests", "tri_tests", "prim_tests",
                                               "sphere_tests", "terrain_steps", "sky_evals")]


def source_hash():
    h = hashlib.sha256()
    for path in _SOURCES:
        h.update(os.path.basename(path).encode())
        h.update(open(path, "rb").read())
    return h.hexdigest()


def is_current():
    """True when liboracle.so exists and was built from the sources as they are now."""
    if not (os.path.exists(LIB_PATH) and os.path.exists(STAMP_PATH)):
        return False
    return open(STAMP_PATH).read().strip() == source_hash()


def build():
    """Compiles liboracle.so from df_oracle.cpp (make, always rebuilding) and stamps the sources' hash."""
    digest = source_hash()
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    with open(STAMP_PATH, "w") as fh:
        fh.write(digest + "\n")
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_render.restype = ctypes.c_int
        _lib.oracle_intersect.restype = ctypes.c_float
        _lib.oracle_simplex2.restype = ctypes.c_float
        _lib.oracle_simplex2.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.oracle_simplex3.restype = ctypes.c_float
        _lib.oracle_simplex3.argtypes = [ctypes.c_float] * 3
        _lib.oracle_fbm3.restype = ctypes.c_float
        _lib.oracle_fbm3.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float]
        _lib.oracle_texture.restype = ctypes.c_int32
        _lib.oracle_next_float.restype = ctypes.c_float
        _lib.oracle_next_float.argtypes = [ctypes.POINTER(ctypes.c_int64)]
    return _lib


class OracleState:
    """Film, seeds and outputs of one oracle render of width x height pixels."""

    def __init__(self, width, height, seeds):
        n = width * height
        self.width, self.height = width, height
        self.film = np.zeros(n * 3, np.float32)
        self.film_samples = np.zeros(n, np.int32)
        self.image_float = np.zeros(n * 3, np.float32)
        self.image_bgra = np.zeros(n * 4, np.uint8)
        self.primitive_offsets = np.full(n, -1, np.int32)
        self.seeds = np.ascontiguousarray(seeds, np.int64).copy()
        self.counters = {n: 0 for n, _ in OracleCounters._fields_}

    def frame(self):
        return OracleFrame(self.width, self.height, self.film.ctypes.data, self.film_samples.ctypes.data,
                           self.image_float.ctypes.data, self.image_bgra.ctypes.data,
                           self.primitive_offsets.ctypes.data, self.seeds.ctypes.data)


def render(scene, params, state, passes, clear_first=False, threads=None, pixel_begin=0, pixel_end=None, sky=None):
    """Runs `passes` passes of run() over [pixel_begin, pixel_end)."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    if pixel_end is None:
        pixel_end = state.width * state.height
    cs = scene.as_struct()
    fr = state.frame()
    cnt = OracleCounters()
    rc = lib().oracle_render(ctypes.byref(cs), ctypes.byref(scene.sky if sky is None else sky), ctypes.byref(params),
                             ctypes.byref(fr), int(passes), 1 if clear_first else 0, int(threads),
                             ctypes.c_int64(pixel_begin), ctypes.c_int64(pixel_end), ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("oracle_render failed: %d" % rc)
    for n, _ in OracleCounters._fields_:
        state.counters[n] += getattr(cnt, n)
    return state


def intersect(scene, params, origin, direction, testing=False):
    """One doIntersectPrimitives query: (t, 27-float record)."""
    cs = scene.as_struct()
    o = (ctypes.c_float * 3)(*origin)
    d = (ctypes.c_float * 3)(*direction)
    rec = (ctypes.c_float * 27)()
    t = lib().oracle_intersect(ctypes.byref(cs), ctypes.byref(scene.sky), ctypes.byref(params), o, d,
                               1 if testing else 0, rec)
    return t, np.array(rec[:], np.float32)


def sky_color(sky, direction, active=True):
    d = (ctypes.c_float * 3)(*direction)
    out = (ctypes.c_float * 3)()
    lib().oracle_sky(ctypes.byref(sky), d, 1 if active else 0, out)
    return np.array(out[:], np.float32)


def texture_color(textures, offset, record):
    t = np.ascontiguousarray(textures, np.float32)
    rec = np.ascontiguousarray(record, np.float32)
    return lib().oracle_texture(t.ctypes.data_as(_c_float_p), ctypes.c_size_t(t.size), int(offset),
                                rec.ctypes.data_as(_c_float_p))


def next_floats(seed, count):
    """`count` draws of the per-pixel LCG from `seed`; returns (draws, final seed)."""
    s = ctypes.c_int64(seed)
    out = np.array([lib().oracle_next_float(ctypes.byref(s)) for _ in range(count)], np.float32)
    return out, s.value


class OracleDiff(ctypes.Structure):
    _fields_ = [
        ("n_albedo", ctypes.c_int32),
        ("n_emission", ctypes.c_int32),
        ("surface_albedo", ctypes.c_void_p),
        ("surface_emission", ctypes.c_void_p),
        ("n_surfaces", ctypes.c_int64),
        ("albedo", ctypes.c_void_p),
        ("emission", ctypes.c_void_p),
        ("n_checker", ctypes.c_int32),
        ("surface_checker", ctypes.c_void_p),
        ("checker", ctypes.c_void_p),
        ("turbidity", ctypes.c_double),
        ("elevation_deg", ctypes.c_double),
        ("azimuth_deg", ctypes.c_double),
    ]


def loss_gradient(scene, params, sky, maps, albedo, emission, turbidity, elevation, azimuth, seeds, target, passes,
                  crop, threads=None, checker=None):
    """FP64 forward-mode gradient of L = sum over the crop (x0, y0, w, h) of (film - target)^2, the film the mean
    of `passes` path-traced samples from `seeds` (the full frame's, width x height from the scene's camera) with
    the differentiable parameters.  `maps` is (surface_albedo, surface_emission) or, with checkerboard groups,
    (surface_albedo, surface_emission, surface_checker) and `checker` their colours (groups x 2 x RGB).  `sky` is
    the sky rendered with (Sky.set of turbidity, elevation, azimuth).  Returns (loss, gradient [albedo groups x
    RGB, emission, checkerboard groups x 2 x RGB, turbidity, elevation per degree, azimuth per degree], crop
    film)."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    width, height = int(scene.camera[16]), int(scene.camera[17])
    x0, y0, cw, ch = (int(v) for v in crop)
    sa = np.ascontiguousarray(maps[0], np.int32)
    se = np.ascontiguousarray(maps[1], np.int32)
    sc = np.ascontiguousarray(maps[2], np.int32) if len(maps) > 2 else None
    al = np.ascontiguousarray(albedo, np.float32).reshape(-1)
    em = np.ascontiguousarray(emission, np.float32).reshape(-1)
    ck = np.ascontiguousarray(checker if checker is not None else np.zeros(0), np.float32).reshape(-1)
    nc = ck.size // 6 if sc is not None else 0
    d = OracleDiff(al.size // 3, em.size, sa.ctypes.data, se.ctypes.data, sa.size, al.ctypes.data, em.ctypes.data,
                   nc, sc.ctypes.data if sc is not None else None, ck.ctypes.data,
                   float(turbidity), float(elevation), float(azimuth))
    sd = np.ascontiguousarray(seeds, np.int64)
    tg = np.ascontiguousarray(target, np.float32).reshape(-1)
    if sd.size != width * height or tg.size != width * height * 3:
        raise ValueError("seeds / target do not match the camera's %d x %d" % (width, height))
    n = 3 * (al.size // 3) + em.size + 6 * nc + 3
    grad = np.zeros(n, np.float64)
    film = np.zeros(cw * ch * 3, np.float32)
    loss = ctypes.c_double()
    cs = scene.as_struct()
    rc = lib().oracle_loss_gradient(ctypes.byref(cs), ctypes.byref(sky), ctypes.byref(params), ctypes.byref(d),
                                    width, height, x0, y0, cw, ch, sd.ctypes.data_as(ctypes.c_void_p),
                                    tg.ctypes.data_as(ctypes.c_void_p), int(passes), int(threads),
                                    film.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss),
                                    grad.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(n))
    if rc != 0:
        raise RuntimeError("oracle_loss_gradient failed: %d" % rc)
    return loss.value, grad, film


def sky_dual(sky, turbidity, elevation, azimuth, direction):
    """The FP64 adjoint's sky towards `direction`: (rgb, d rgb / d turbidity, d rgb / d elevation, d rgb / d
    azimuth) or None below the horizon."""
    d = (ctypes.c_float * 3)(*direction)
    out = (ctypes.c_double * 12)()
    f = lib().oracle_sky_dual
    f.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                  ctypes.c_void_p]
    if not f(ctypes.addressof(sky), turbidity, elevation, azimuth, ctypes.addressof(d), ctypes.addressof(out)):
        return None
    v = np.array(out[:], np.float64)
    return v[0:3], v[3:6], v[6:9], v[9:12]


def sun_cosine_dual(sky, turbidity, elevation, azimuth, point, normal, u, v):
    """The FP64 adjoint's sun-sample cosine at random numbers (u, v): (cosine, d cosine / d elevation,
    d cosine / d azimuth)."""
    p = (ctypes.c_float * 3)(*point)
    n = (ctypes.c_float * 3)(*normal)
    out = (ctypes.c_double * 3)()
    f = lib().oracle_sun_cosine_dual
    f.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    f.restype = None
    f(ctypes.addressof(sky), turbidity, elevation, azimuth, ctypes.addressof(p), ctypes.addressof(n), u, v,
      ctypes.addressof(out))
    return out[0], out[1], out[2]


def primary_ray(camera, renderer_type, width, height, x, y, seed):
    """doCreatePrimaryRay for pixel (x, y): (ray [origin, direction] or None where the lens gives none,
    seed after the draws)."""
    cam = np.ascontiguousarray(camera, np.float32)
    s = ctypes.c_int64(seed)
    out = (ctypes.c_float * 6)()
    f = lib().oracle_primary_ray
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_void_p, ctypes.c_void_p]
    ok = f(cam.ctypes.data, int(renderer_type), int(width), int(height), int(x), int(y), ctypes.addressof(s),
           ctypes.addressof(out))
    return (np.array(out[:], np.float32) if ok else None), s.value


def scatter(material, point, normal, direction, albedo, seed):
    """The path tracer's BSDF sample: (next origin, next direction, throughput factor, seed after the draws)."""
    p = (ctypes.c_float * 3)(*point)
    n = (ctypes.c_float * 3)(*normal)
    d = (ctypes.c_float * 3)(*direction)
    a = (ctypes.c_float * 3)(*albedo)
    s = ctypes.c_int64(seed)
    out = (ctypes.c_float * 9)()
    f = lib().oracle_scatter
    f.restype = None
    f.argtypes = [ctypes.c_int32] + [ctypes.c_void_p] * 6
    f(int(material), ctypes.addressof(p), ctypes.addressof(n), ctypes.addressof(d), ctypes.addressof(a),
      ctypes.addressof(s), ctypes.addressof(out))
    v = np.array(out[:], np.float32)
    return v[0:3], v[3:6], v[6:9], s.value


def sun_direction(sky, point, u, v):
    """The direction doCalculateColorForSun samples from `point` with the random numbers (u, v)."""
    p = (ctypes.c_float * 3)(*point)
    out = (ctypes.c_float * 3)()
    f = lib().oracle_sun_direction
    f.restype = None
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    f(ctypes.addressof(sky), ctypes.addressof(p), float(u), float(v), ctypes.addressof(out))
    return np.array(out[:], np.float32)


def render_ray(scene, params, ray, seed, sky=None):
    """run() for one pixel whose primary ray is given: (film sample rgb, primitive offset, BGRA, seed after)."""
    r = (ctypes.c_float * 6)(*ray)
    s = ctypes.c_int64(seed)
    rgb = (ctypes.c_float * 3)()
    po = ctypes.c_int32(-1)
    bgra = (ctypes.c_uint8 * 4)()
    cs = scene.as_struct()
    f = lib().oracle_render_ray
    f.restype = ctypes.c_int
    rc = f(ctypes.byref(cs), ctypes.byref(scene.sky if sky is None else sky), ctypes.byref(params), r,
           ctypes.byref(s), rgb, ctypes.byref(po), bgra)
    if rc != 0:
        raise RuntimeError("oracle_render_ray failed: %d" % rc)
    return np.array(rgb[:], np.float32), po.value, np.array(bgra[:], np.uint8), s.value


# The deliberate mistakes of df_oracle.cpp's enum Plant, switched on one at a time by the pin tests.
PLANTS = {name: i for i, name in enumerate([
    "none", "camera_res", "camera_box", "lens_radius", "fisheye_w", "lambert_uniform", "phong_exponent",
    "fresnel_prob", "snell_ratio", "sun_cone", "ao_uniform", "phong_shader", "sphere_uv", "barycentric",
    "normal_map_b", "noise_normal", "bullseye", "perez_theta0", "perez_swap", "fbm_texture", "clearcoat_offset",
    "checker_tile"])}


class planted:
    """Context manager: the oracle with the deliberate mistake `name` switched on (TEST INFRASTRUCTURE)."""

    def __init__(self, name):
        self.id = PLANTS[name]

    def __enter__(self):
        f = lib().oracle_plant
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int]
        self.old = f(self.id)
        return self

    def __exit__(self, *exc):
        lib().oracle_plant(self.old)
        return False
