"""Compiled scenes: the reference's flat CompiledScene arrays as numpy arrays.

``CompiledScene`` holds the 14 arrays in the order of
scene/compiler/CompiledScene.java:273-295 plus the sun/sky scalars and the
render configuration a workload comes with.  Scenes come from the
host scene compiler (libdayflower_scene.so): the procedural benchmark
workloads, the uneven batch mix, or a reference ``.scene`` file.
"""

import ctypes

import numpy as np

from . import _ffi

ARRAY_NAMES = _ffi.ARRAY_NAMES
INT_ARRAYS = _ffi.INT_ARRAYS


def _check(rc, what):
    if rc != _ffi.DF_OK:
        raise RuntimeError("%s failed: %s" % (what, _ffi.ERROR_NAMES.get(rc, rc)))


def default_params():
    """Reference defaults (AbstractRendererKernel.java:171-184)."""
    p = _ffi.Params()
    _ffi.scene_lib().dfs_default_params(ctypes.byref(p))
    return p


def sky_set(direction, turbidity=2.0):
    """Sky.set(sunDirectionWorld, turbidity) (scene/Sky.java:295-333)."""
    s = _ffi.Sky()
    _check(_ffi.scene_lib().dfs_sky_set(ctypes.byref(s), float(direction[0]), float(direction[1]),
                                        float(direction[2]), float(turbidity)), "dfs_sky_set")
    return s


def sun_direction(elevation_deg, azimuth_deg):
    """World direction towards the sun: y up, azimuth measured from +z towards +x."""
    e, a = np.radians(float(elevation_deg)), np.radians(float(azimuth_deg))
    return (np.cos(e) * np.sin(a), np.sin(e), np.cos(e) * np.cos(a))


def sky_at(elevation_deg, azimuth_deg, turbidity):
    """Sky.set for a sun at (elevation, azimuth) and a turbidity."""
    return sky_set(sun_direction(elevation_deg, azimuth_deg), turbidity)


def diff_maps(scene):
    """Differentiable-mode parameter maps of a scene: albedo group k is the k-th TriangleMesh primitive's
    surface, emission j the j-th Triangle primitive's.  Returns (surface_albedo, surface_emission, n_albedo,
    n_emission), the maps holding one entry per surface record (-1: the textures)."""
    n_surf = scene.surfaces.size // 6
    sa = np.full(n_surf, -1, np.int32)
    se = np.full(n_surf, -1, np.int32)
    na = ne = 0
    for st, _, so in scene.primitives.reshape(-1, 3):
        if st == 5 and sa[so // 6] < 0:
            sa[so // 6] = na
            na += 1
        elif st == 4 and se[so // 6] < 0:
            se[so // 6] = ne
            ne += 1
    return sa, se, na, ne


def diff_checker_map(scene):
    """Differentiable-mode checkerboard groups of a scene: group g is the g-th surface record, in record order,
    whose albedo texture is a CheckerboardTexture and that no TriangleMesh primitive uses (those are albedo
    groups).  Returns (surface_checker, n_checker, colours): one entry per surface record (-1: none) and the
    groups' texture colours (n_checker x 2 x RGB, colour0 then colour1, each channel byte / 255 as the
    reference's texture lookup decodes it, CheckerboardTexture.java / GPURendererKernel.java:1714-1767)."""
    n_surf = scene.surfaces.size // 6
    sa = diff_maps(scene)[0]
    sc = np.full(n_surf, -1, np.int32)
    cols = []
    recip = np.float32(1.0) / np.float32(255.0)
    for rec in range(n_surf):
        off = int(scene.surfaces[rec * 6 + 1])
        if sa[rec] >= 0 or int(scene.textures[off]) != 3:
            continue
        sc[rec] = len(cols)
        pair = []
        for k in (2, 3):
            c = int(np.float32(scene.textures[off + k]))
            pair.append([np.float32((c >> sh) & 0xFF) * recip for sh in (16, 8, 0)])
        cols.append(pair)
    return sc, len(cols), np.array(cols, np.float32).reshape(len(cols), 2, 3)


def camera_setup(width, height, fov_x, eye, target, aperture=0.0, focal=30.0):
    """Camera array for a look-at (Camera.java:743-748, 877-946)."""
    cam = np.zeros(19, np.float32)
    _check(_ffi.scene_lib().dfs_camera_setup(
        cam.ctypes.data_as(_ffi.c_float_p), float(width), float(height), float(fov_x),
        float(eye[0]), float(eye[1]), float(eye[2]), float(target[0]), float(target[1]), float(target[2]),
        float(aperture), float(focal)), "dfs_camera_setup")
    return cam


class CompiledScene:
    """The flat arrays of one compiled scene plus its sky and render setup."""

    def __init__(self, name, arrays, sky, width, height, spp, params):
        self.name = name
        self.arrays = arrays
        self.sky = sky
        self.width = width
        self.height = height
        self.spp = spp
        self.params = params

    def __getattr__(self, item):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and item in arrays:
            return arrays[item]
        raise AttributeError(item)

    @property
    def primitive_count(self):
        return len(self.arrays["primitives"]) // 3

    def as_struct(self):
        """A df_compiled_scene view of the arrays (valid while self lives)."""
        cs = _ffi.CompiledScene()
        for n in ARRAY_NAMES:
            a = self.arrays[n]
            ptr_t = _ffi.c_int32_p if n in INT_ARRAYS else _ffi.c_float_p
            setattr(cs, n, a.ctypes.data_as(ptr_t))
            setattr(cs, n + "_len", a.size)
        return cs

    def copy_params(self):
        """An independent copy of the workload's renderer parameters."""
        p = _ffi.Params()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self.params), ctypes.sizeof(_ffi.Params))
        return p

    def validate(self):
        """df_validate_scene: (status, message) of the host-side range checks."""
        buf = ctypes.create_string_buffer(512)
        rc = _ffi.hip_lib().df_validate_scene(ctypes.byref(self.as_struct()), buf, 512)
        return rc, buf.value.decode()

    def with_camera(self, camera):
        arrays = dict(self.arrays)
        arrays["camera"] = np.ascontiguousarray(camera, np.float32)
        return CompiledScene(self.name, arrays, self.sky, int(camera[16]), int(camera[17]), self.spp,
                             self.copy_params())

    def write(self, path):
        """CompiledScene.write (CompiledScene.java:273-320)."""
        lib = _ffi.scene_lib()
        h = ctypes.c_void_p()
        cs = self.as_struct()
        _check(lib.dfs_scene_from_arrays(ctypes.byref(cs), self.name.encode(), ctypes.byref(h)), "dfs_scene_from_arrays")
        try:
            _check(lib.dfs_write_scene(h, str(path).encode()), "dfs_write_scene")
        finally:
            lib.dfs_free(h)


def _from_handle(h):
    lib = _ffi.scene_lib()
    cs = _ffi.CompiledScene()
    _check(lib.dfs_get_compiled(h, ctypes.byref(cs)), "dfs_get_compiled")
    arrays = {}
    for n in ARRAY_NAMES:
        length = getattr(cs, n + "_len")
        ptr = getattr(cs, n)
        dtype = np.int32 if n in INT_ARRAYS else np.float32
        arrays[n] = np.ctypeslib.as_array(ptr, shape=(length,)).astype(dtype, copy=True) if length else np.zeros(0, dtype)
    sky = _ffi.Sky()
    _check(lib.dfs_get_sky(h, ctypes.byref(sky)), "dfs_get_sky")
    w, hh, spp = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    params = _ffi.Params()
    _check(lib.dfs_get_render_config(h, ctypes.byref(w), ctypes.byref(hh), ctypes.byref(spp), ctypes.byref(params)),
           "dfs_get_render_config")
    buf = ctypes.create_string_buffer(256)
    lib.dfs_get_name(h, buf, 256)
    return CompiledScene(buf.value.decode(), arrays, sky, w.value, hh.value, spp.value, params)


def _build(fn, *args):
    lib = _ffi.scene_lib()
    h = ctypes.c_void_p()
    _check(fn(*args, ctypes.byref(h)), fn.__name__)
    try:
        return _from_handle(h)
    finally:
        lib.dfs_free(h)


def build_config(config_id, seed=0):
    """Benchmark workload 1, 2 or 3 (see include/dayflower_scene.h)."""
    return _build(_ffi.scene_lib().dfs_build_config, int(config_id), ctypes.c_uint64(seed))


def build_batch_scene(seed):
    """One scene of the uneven batch mix (seeds 100..355)."""
    return _build(_ffi.scene_lib().dfs_build_batch_scene, ctypes.c_uint64(seed))


def read_scene(path):
    """CompiledScene.read (CompiledScene.java:323-377)."""
    return _build(_ffi.scene_lib().dfs_read_scene, str(path).encode())


def stats(scene):
    lib = _ffi.scene_lib()
    h = ctypes.c_void_p()
    cs = scene.as_struct()
    _check(lib.dfs_scene_from_arrays(ctypes.byref(cs), scene.name.encode(), ctypes.byref(h)), "dfs_scene_from_arrays")
    try:
        st = _ffi.SceneStats()
        _check(lib.dfs_get_stats(h, ctypes.byref(st)), "dfs_get_stats")
        return {n: getattr(st, n) for n, _ in _ffi.SceneStats._fields_}
    finally:
        lib.dfs_free(h)


def pixel_seeds(master, n):
    """The per-pixel seeds df_seed(master) produces (SplitMix64 of
    master + (i + 1) * 0x9E3779B97F4A7C15), computed on the host."""
    golden = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(master) + (np.arange(n, dtype=np.uint64) + np.uint64(1)) * golden
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z.view(np.int64)
