"""Python mirror of the reference renderer-kernel interface over the C ABI.

``HIPRendererKernel`` has the method names, argument meaning and error
behaviour of the reference's AbstractRendererKernel
(kernel/AbstractRendererKernel.java:36-727) and GPURendererKernel
(kernel/GPURendererKernel.java:166-569), plus the Aparapi Kernel calls the
application makes (execute(range), get(...), dispose()).  Every call goes to
libdayflower_hip.so; there is no CPU path.
"""

import ctypes

import numpy as np

from . import _ffi
from .scene import CompiledScene

BOOLEAN_FALSE = 0
BOOLEAN_TRUE = 1

RENDERER_TYPE_AMBIENT_OCCLUSION = 1
RENDERER_TYPE_PATH_TRACER = 2
RENDERER_TYPE_RAY_CASTER = 3
RENDERER_TYPE_RAY_MARCHER = 4
RENDERER_TYPE_RAY_TRACER = 5
RENDERER_TYPE_SURFACE_NORMALS = 6

SHADER_TYPE_FLAT = 1
SHADER_TYPE_GOURAUD = 2

TONE_MAPPER_TYPE_FILMIC_CURVE_ACES_MODIFIED = 1
TONE_MAPPER_TYPE_REINHARD = 2
TONE_MAPPER_TYPE_REINHARD_MODIFIED_1 = 3
TONE_MAPPER_TYPE_REINHARD_MODIFIED_2 = 4

MATH_FAST = 0
MATH_EXACT = 1
PT_AUTO = 0
PT_MEGAKERNEL = 1
PT_WAVEFRONT = 2
TILE_ROWS = 0
TILE_SQUARE = 1
TAIL_BOUNCE_DEFAULT = 4  # DF_TAIL_BOUNCE_DEFAULT


class HIPError(RuntimeError):
    """A failed C-ABI call; ``code`` is the DF_E* status."""

    def __init__(self, code, message):
        super().__init__("%s: %s" % (_ffi.ERROR_NAMES.get(code, code), message))
        self.code = code


def device_count():
    n = ctypes.c_int32()
    _ffi.hip_lib().df_device_count(ctypes.byref(n))
    return n.value


ROOFS = {"valu_fp32": 1, "l2_read": 2, "mall_read": 3, "hbm_read": 4}


def measure_roof(kind, device=0):
    """df_measure_roof: the sustained rate of one microbenchmark on `device` (FLOP/s for "valu_fp32", bytes/s
    for "l2_read", "mall_read", "hbm_read") and the work of one timed launch."""
    rate, work = ctypes.c_double(), ctypes.c_double()
    rc = _ffi.hip_lib().df_measure_roof(int(device), ROOFS[kind], ctypes.byref(rate), ctypes.byref(work))
    if rc != _ffi.DF_OK:
        raise HIPError(rc, "df_measure_roof(%s) failed" % kind)
    return rate.value, work.value


def check_reciprocal(device=0):
    """df_check_reciprocal: (values tested, mismatches, smallest mismatching bit pattern) of the triangle test's
    reciprocal against the correctly rounded division over 2^-14 < |x| < 2^126."""
    tested, bad, first = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
    rc = _ffi.hip_lib().df_check_reciprocal(int(device), ctypes.byref(tested), ctypes.byref(bad), ctypes.byref(first))
    if rc != _ffi.DF_OK:
        raise HIPError(rc, "df_check_reciprocal failed")
    return tested.value, bad.value, first.value


def check_terrain_march(rays, math_mode=None, device=0):
    """df_check_terrain_march: (t_skip, t_plain, evals) of the kernels' terrain march and of the reference's march
    over every position for rays of shape (n, 6) (object-space origin, direction)."""
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
    n = rays.shape[0]
    t_skip = np.empty(n, np.float32)
    t_plain = np.empty(n, np.float32)
    evals = np.empty(n, np.uint32)
    mode = MATH_FAST if math_mode is None else math_mode
    rc = _ffi.hip_lib().df_check_terrain_march(
        int(device), int(mode), rays.ctypes.data_as(_ffi.c_float_p), n, t_skip.ctypes.data_as(_ffi.c_float_p),
        t_plain.ctypes.data_as(_ffi.c_float_p), evals.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    if rc != _ffi.DF_OK:
        raise HIPError(rc, "df_check_terrain_march failed")
    return t_skip, t_plain, evals


class HIPRendererKernel:
    """Renderer bound to one compiled scene on one HIP device."""

    def __init__(self, scene, device=0):
        if not isinstance(scene, CompiledScene):
            raise TypeError("scene must be a CompiledScene")
        self._lib = _ffi.hip_lib()
        self._ctx = ctypes.c_void_p()
        rc = self._lib.df_create(ctypes.byref(self._ctx), int(device))
        if rc != _ffi.DF_OK:
            self._ctx = None
            raise HIPError(rc, "df_create(device=%d)" % device)
        self._scene = scene
        self._image = None
        self.width = 0
        self.height = 0
        self._changed = False
        self._check(self._lib.df_load_scene(self._ctx, ctypes.byref(scene.as_struct())), "df_load_scene")
        self._check(self._lib.df_upload_sky(self._ctx, ctypes.byref(scene.sky)), "df_upload_sky")
        self._params = _ffi.Params()
        self._check(self._lib.df_get_params(self._ctx, ctypes.byref(self._params)), "df_get_params")

    # ------------------------------------------------------------ plumbing
    def _check(self, rc, what):
        if rc != _ffi.DF_OK:
            msg = self._lib.df_last_error(self._ctx) if self._ctx else b""
            raise HIPError(rc, "%s: %s" % (what, (msg or b"").decode()))

    def _push_params(self):
        self._check(self._lib.df_set_params(self._ctx, ctypes.byref(self._params)), "df_set_params")
        self._changed = True

    @classmethod
    def _borrowed(cls, lib, ctx, scene, width, height):
        """A view of a renderer handle owned elsewhere (a batch job); dispose() leaves it alone."""
        self = cls.__new__(cls)
        self._lib = lib
        self._ctx = ctx
        self._owned = False
        self._scene = scene
        self._image = None
        self.width, self.height = int(width), int(height)
        self._changed = False
        self._params = _ffi.Params()
        self._check(self._lib.df_get_params(self._ctx, ctypes.byref(self._params)), "df_get_params")
        return self

    def dispose(self):
        """Kernel.dispose(): releases the device resources."""
        if self._ctx and getattr(self, "_owned", True):
            self._lib.df_destroy(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.dispose()

    # ------------------------------------------------------------ lifecycle
    def update(self, resolutionX, resolutionY, imageDataByte=None, localSize=256):
        """update(W, H, byte[], localSize) (GPURendererKernel.java:411-449).

        Raises ValueError when ``W * H * 4 != len(imageDataByte)``
        (AbstractImageKernel.java:146-148)."""
        n = int(resolutionX) * int(resolutionY)
        if imageDataByte is None:
            imageDataByte = np.zeros(n * 4, np.uint8)
        if n * 4 != len(imageDataByte):
            raise ValueError("resolutionX * resolutionY * 4 != imageDataByte.length: resolutionX=%d, resolutionY=%d, "
                             "imageDataByte.length=%d" % (resolutionX, resolutionY, len(imageDataByte)))
        self._check(self._lib.df_update(self._ctx, int(resolutionX), int(resolutionY), int(localSize)), "df_update")
        self.width, self.height = int(resolutionX), int(resolutionY)
        self._image = imageDataByte

    def seed(self, master):
        """Deterministic per-pixel seeds (replaces ThreadLocalRandom)."""
        self._check(self._lib.df_seed(self._ctx, ctypes.c_uint64(master)), "df_seed")

    def setSeeds(self, seeds):
        s = np.ascontiguousarray(seeds, np.int64)
        self._check(self._lib.df_set_seeds(self._ctx, s.ctypes.data_as(_ffi.c_int64_p), s.size), "df_set_seeds")

    def getSeeds(self):
        s = np.zeros(self.width * self.height, np.int64)
        self._check(self._lib.df_get_seeds(self._ctx, s.ctypes.data_as(_ffi.c_int64_p), s.size), "df_get_seeds")
        return s

    def updateCamera(self, camera=None):
        """updateCamera() (GPURendererKernel.java:455-461)."""
        cam = np.ascontiguousarray(self._scene.camera if camera is None else camera, np.float32)
        self._check(self._lib.df_upload_camera(self._ctx, cam.ctypes.data_as(_ffi.c_float_p), cam.size), "df_upload_camera")

    def updatePrimitives(self, objectToWorld, worldToObject):
        """updatePrimitives() (GPURendererKernel.java:467-521)."""
        o = np.ascontiguousarray(objectToWorld, np.float32)
        w = np.ascontiguousarray(worldToObject, np.float32)
        self._check(self._lib.df_upload_transforms(self._ctx, o.ctypes.data_as(_ffi.c_float_p),
                                                   w.ctypes.data_as(_ffi.c_float_p), o.size // 16),
                    "df_upload_transforms")
        self._changed = True

    def updateSurfaces(self, surfaces):
        s = np.ascontiguousarray(surfaces, np.float32)
        self._check(self._lib.df_upload_surfaces(self._ctx, s.ctypes.data_as(_ffi.c_float_p), s.size),
                    "df_upload_surfaces")
        self._changed = True

    def updateSunAndSky(self, sky=None):
        """updateSunAndSky() (GPURendererKernel.java:527-569)."""
        self._check(self._lib.df_upload_sky(self._ctx, ctypes.byref(self._scene.sky if sky is None else sky)),
                    "df_upload_sky")
        self._changed = True

    # ------------------------------------------------------------ passes
    def clear(self):
        """clear(): the next pass overwrites the film (AbstractImageKernel.java:119-121)."""
        self._check(self._lib.df_clear(self._ctx), "df_clear")

    def clearFilmFlags(self):
        """clearFilmFlags(): folded into execute (each execute ends a pass)."""

    def execute(self, passes=1):
        """execute(range) + clearFilmFlags(), ``passes`` times (asynchronous)."""
        self._check(self._lib.df_execute(self._ctx, int(passes)), "df_execute")

    def schedule(self):
        """df_get_schedule: {"engine", "wave_tile", "shade_binned", "bvh_pairs", "tail_bounce"} of the next passes
        (engine PT_AUTO while the autotuner is still timing its candidates)."""
        sc = _ffi.Schedule()
        self._check(self._lib.df_get_schedule(self._ctx, ctypes.byref(sc)), "df_get_schedule")
        return {"engine": sc.engine, "wave_tile": sc.wave_tile, "shade_binned": sc.shade_binned,
                "bvh_pairs": sc.bvh_pairs, "tail_bounce": sc.tail_bounce}

    def setSchedule(self, engine=None, wave_tile=0, shade_binned=0, bvh_pairs=0, tail_bounce=TAIL_BOUNCE_DEFAULT):
        """df_set_schedule: fix the path tracer's schedule (engine None: back to the parameters' pt_engine).
        tail_bounce: the first bounce whose trace launch runs the kernel built for short queues."""
        if engine is None:
            self._check(self._lib.df_set_schedule(self._ctx, None), "df_set_schedule")
        else:
            sc = _ffi.Schedule(int(engine), int(wave_tile), int(shade_binned), int(bvh_pairs), int(tail_bounce))
            self._check(self._lib.df_set_schedule(self._ctx, ctypes.byref(sc)), "df_set_schedule")

    def bvhPairStack(self):
        """df_get_bvh_pair_stack: the per-lane stack depth of the scene's two-child walk (0: not available)."""
        d = ctypes.c_int32()
        self._check(self._lib.df_get_bvh_pair_stack(self._ctx, ctypes.byref(d)), "df_get_bvh_pair_stack")
        return d.value

    def setGraphs(self, enabled):
        """df_set_graphs: replay the frame loop's launches from captured HIP graphs (default off)."""
        self._check(self._lib.df_set_graphs(self._ctx, 1 if enabled else 0), "df_set_graphs")

    def synchronize(self):
        self._check(self._lib.df_synchronize(self._ctx), "df_synchronize")

    def lastExecuteMillis(self):
        ms = ctypes.c_float()
        self._check(self._lib.df_last_execute_ms(self._ctx, ctypes.byref(ms)), "df_last_execute_ms")
        return ms.value

    # ------------------------------------------------------------ readback
    def get(self):
        """get(imageDataByte): copies the BGRA image into the array given to update()."""
        buf = np.zeros(self.width * self.height * 4, np.uint8)
        self._check(self._lib.df_get_image_bgra(self._ctx, buf.ctypes.data_as(_ffi.c_uint8_p), buf.size),
                    "df_get_image_bgra")
        if isinstance(self._image, np.ndarray):
            self._image[:] = buf
        else:
            self._image[:] = buf.tobytes()
        return self._image

    def getImageBGRA(self):
        buf = np.zeros(self.width * self.height * 4, np.uint8)
        self._check(self._lib.df_get_image_bgra(self._ctx, buf.ctypes.data_as(_ffi.c_uint8_p), buf.size),
                    "df_get_image_bgra")
        return buf

    def getImageFloat(self):
        buf = np.zeros(self.width * self.height * 3, np.float32)
        self._check(self._lib.df_get_image_float(self._ctx, buf.ctypes.data_as(_ffi.c_float_p), buf.size),
                    "df_get_image_float")
        return buf

    def getPrimitiveOffsets(self):
        """get(primitiveOffsets): the ray-cast buffer."""
        buf = np.zeros(self.width * self.height, np.int32)
        self._check(self._lib.df_get_primitive_offsets(self._ctx, buf.ctypes.data_as(_ffi.c_int32_p), buf.size),
                    "df_get_primitive_offsets")
        return buf

    def getFilm(self):
        rgb = np.zeros(self.width * self.height * 3, np.float32)
        n = np.zeros(self.width * self.height, np.int32)
        self._check(self._lib.df_get_film(self._ctx, rgb.ctypes.data_as(_ffi.c_float_p), rgb.size,
                                          n.ctypes.data_as(_ffi.c_int32_p), n.size), "df_get_film")
        return rgb, n

    def setFilm(self, rgb, samples):
        rgb = np.ascontiguousarray(rgb, np.float32)
        n = np.ascontiguousarray(samples, np.int32)
        self._check(self._lib.df_set_film(self._ctx, rgb.ctypes.data_as(_ffi.c_float_p), rgb.size,
                                          n.ctypes.data_as(_ffi.c_int32_p), n.size), "df_set_film")

    def counters(self):
        c = _ffi.Counters()
        self._check(self._lib.df_get_counters(self._ctx, ctypes.byref(c)), "df_get_counters")
        return {n: getattr(c, n) for n in _ffi.COUNTER_NAMES}

    def resetCounters(self):
        self._check(self._lib.df_reset_counters(self._ctx), "df_reset_counters")

    def setDifferentiable(self, surface_albedo, surface_emission, albedo, emission, turbidity, elevation_deg,
                          azimuth_deg, sky_active=True, sun_active=True, surface_checker=None, checker=None):
        """Differentiable mode: the surfaces mapped by surface_albedo / surface_emission (one entry per surface
        record, -1 = texture) take the continuous albedo (groups x RGB) and emission (white radiance) parameters,
        those mapped by surface_checker the two colours (groups x 2 x RGB) of their checkerboard albedo texture;
        the sky is Sky.set of (turbidity, sun elevation / azimuth in degrees), differentiated with respect to all
        three.  Uploads that sky (with the sky / sun switches given) first."""
        from .scene import sky_at

        sc = None if surface_checker is None else np.ascontiguousarray(surface_checker, np.int32)
        ck = np.ascontiguousarray(checker if checker is not None else np.zeros(0), np.float32).reshape(-1)
        self._diff_keep = [np.ascontiguousarray(surface_albedo, np.int32), np.ascontiguousarray(surface_emission, np.int32),
                           np.ascontiguousarray(albedo, np.float32).reshape(-1), np.ascontiguousarray(emission, np.float32),
                           sc, ck]
        sa, se, al, em = self._diff_keep[:4]
        d = _ffi.Diff()
        d.n_albedo = al.size // 3
        d.n_emission = em.size
        d.surface_albedo = sa.ctypes.data_as(_ffi.c_int32_p)
        d.surface_emission = se.ctypes.data_as(_ffi.c_int32_p)
        d.n_surfaces = sa.size
        d.albedo = al.ctypes.data_as(_ffi.c_float_p)
        d.emission = em.ctypes.data_as(_ffi.c_float_p)
        d.n_checker = ck.size // 6 if sc is not None else 0
        if sc is not None:
            d.surface_checker = sc.ctypes.data_as(_ffi.c_int32_p)
            d.checker = ck.ctypes.data_as(_ffi.c_float_p)
        d.turbidity = float(turbidity)
        d.elevation_deg = float(elevation_deg)
        d.azimuth_deg = float(azimuth_deg)
        sky = sky_at(elevation_deg, azimuth_deg, turbidity)
        sky.is_sky_active = 1 if sky_active else 0
        sky.is_sun_active = 1 if sun_active else 0
        self._check(self._lib.df_upload_sky(self._ctx, ctypes.byref(sky)), "df_upload_sky")
        self._check(self._lib.df_set_diff(self._ctx, ctypes.byref(d)), "df_set_diff")
        self._n_grad = 3 * d.n_albedo + d.n_emission + 6 * d.n_checker + 3

    def clearDifferentiable(self):
        self._check(self._lib.df_clear_diff(self._ctx), "df_clear_diff")

    def lossGradient(self, target, passes, gradient=True):
        """Forward render of `passes` samples into a cleared film, L = sum (film - target)^2, and with gradient
        dL/d[albedo groups x RGB, emission, checkerboard groups x 2 x RGB, turbidity, elevation, azimuth] by
        path replay.  Returns (loss, grad or None, forward_ms, adjoint_ms)."""
        t = np.ascontiguousarray(target, np.float32).reshape(-1)
        loss = ctypes.c_double()
        fms, ams = ctypes.c_float(), ctypes.c_float()
        grad = None
        gp, ng = None, 0
        if gradient:
            ng = getattr(self, "_n_grad", 3)
            grad = np.zeros(ng, np.float64)
            gp = grad.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._check(self._lib.df_loss_gradient(self._ctx, t.ctypes.data_as(_ffi.c_float_p), t.size, int(passes),
                                               ctypes.byref(loss), gp, ng, ctypes.byref(fms), ctypes.byref(ams)),
                    "df_loss_gradient")
        return loss.value, grad, fms.value, (ams.value if gradient else None)

    def setUnitCounters(self, enabled):
        """df_set_unit_counters: with False the wavefront's trace kernels skip the unit counters (box, triangle,
        primitive and sphere tests, terrain heights), about 3 % of their time; rays and samples are still counted
        and the results are the same bits."""
        self._check(self._lib.df_set_unit_counters(self._ctx, 1 if enabled else 0), "df_set_unit_counters")

    def setStageTiming(self, enabled):
        """Brackets every kernel launch with HIP events (see stageStats)."""
        self._check(self._lib.df_set_stage_timing(self._ctx, 1 if enabled else 0), "df_set_stage_timing")

    def stageStats(self):
        """{stage: {"counters": {...}, "ms": summed kernel time, "launches": n}} since resetCounters()."""
        out = {}
        for i, name in enumerate(_ffi.STAGE_NAMES):
            c = _ffi.Counters()
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            self._check(self._lib.df_get_stage_stats(self._ctx, i, ctypes.byref(c), ctypes.byref(ms), ctypes.byref(n)),
                        "df_get_stage_stats")
            out[name] = {"counters": {k: getattr(c, k) for k in _ffi.COUNTER_NAMES}, "ms": ms.value,
                         "launches": n.value}
        return out

    # ------------------------------------------------------------ setters (AbstractRendererKernel.java:392-575)
    # Invalid enum values are ignored, as in the reference.
    def setRendererType(self, rendererType):
        if rendererType in (1, 2, 3, 4, 5, 6) and rendererType != self._params.renderer_type:
            self._params.renderer_type = rendererType
            self._push_params()

    def getRendererType(self):
        return self._params.renderer_type

    def setToneMapperType(self, toneMapperType):
        if toneMapperType in (1, 2, 3, 4) and toneMapperType != self._params.tone_mapper_type:
            self._params.tone_mapper_type = toneMapperType
            self._push_params()

    def getToneMapperType(self):
        return self._params.tone_mapper_type

    def setToneMapperExposure(self, exposure):
        self._params.tone_mapper_exposure = float(exposure)
        self._push_params()

    def setRendererPTRayDepthMaximum(self, depth):
        self._params.pt_depth_maximum = int(depth)
        self._push_params()

    def getRendererPTRayDepthMaximum(self):
        return self._params.pt_depth_maximum

    def setRendererPTRayDepthRussianRoulette(self, depth):
        self._params.pt_depth_russian_roulette = int(depth)
        self._push_params()

    def setRendererAOMaximumDistance(self, distance):
        self._params.ao_maximum_distance = float(distance)
        self._push_params()

    def setRendererNormalMapping(self, value):
        if value in (BOOLEAN_FALSE, BOOLEAN_TRUE):
            self._params.normal_mapping = value
            self._push_params()

    def setRendererWireframes(self, value):
        if value in (BOOLEAN_FALSE, BOOLEAN_TRUE):
            self._params.wireframes = value
            self._push_params()

    def setShaderType(self, shaderType):
        if shaderType in (SHADER_TYPE_FLAT, SHADER_TYPE_GOURAUD):
            self._params.shader_type = shaderType
            self._push_params()

    def setGlobalAmplitude(self, v):
        self._params.global_amplitude = float(v)
        self._push_params()

    def setGlobalFrequency(self, v):
        self._params.global_frequency = float(v)
        self._push_params()

    def setGlobalGain(self, v):
        self._params.global_gain = float(v)
        self._push_params()

    def setGlobalLacunarity(self, v):
        self._params.global_lacunarity = float(v)
        self._push_params()

    def setGlobalOctaves(self, v):
        self._params.global_octaves = int(v)
        self._push_params()

    def setMathMode(self, mode):
        """MATH_FAST (device float library) or MATH_EXACT (double-rounded transcendentals)."""
        self._params.math_mode = int(mode)
        self._push_params()

    def setPathTracerEngine(self, engine):
        """PT_AUTO, PT_MEGAKERNEL or PT_WAVEFRONT: how the path tracer is scheduled (same results)."""
        self._params.pt_engine = int(engine)
        self._push_params()

    def engine(self):
        """PT_MEGAKERNEL or PT_WAVEFRONT, or PT_AUTO while the autotuner is still timing both."""
        e = ctypes.c_int32()
        self._check(self._lib.df_get_engine(self._ctx, ctypes.byref(e)), "df_get_engine")
        return e.value

    def setParams(self, params):
        ctypes.memmove(ctypes.byref(self._params), ctypes.byref(params), ctypes.sizeof(_ffi.Params))
        self._push_params()

    def getParams(self):
        p = _ffi.Params()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self._params), ctypes.sizeof(_ffi.Params))
        return p

    def toggleSky(self):
        self._check(self._lib.df_toggle_sky(self._ctx), "df_toggle_sky")
        self._changed = True

    def toggleSun(self):
        self._check(self._lib.df_toggle_sun(self._ctx), "df_toggle_sun")
        self._changed = True

    def togglePrimitiveSelection(self, x, y):
        self._check(self._lib.df_toggle_primitive_selection(self._ctx, int(x), int(y)), "df_toggle_primitive_selection")
        self._check(self._lib.df_get_params(self._ctx, ctypes.byref(self._params)), "df_get_params")

    def togglePrimitiveMaterial(self):
        self._check(self._lib.df_toggle_primitive_material(self._ctx), "df_toggle_primitive_material")
        self._changed = True

    def getSelectedPrimitiveIndex(self):
        o = self._params.selected_primitive_offset
        return -1 if o == -1 else o // 3

    def isChanged(self):
        return self._changed

    def setChanged(self, changed):
        self._changed = bool(changed)


class HIPBatchRenderer:
    """Many independent scenes rendered by one persistent launch (df_batch_*).

    ``add`` registers a scene with its own parameters, resolution, passes per
    execute and seed; ``execute`` renders every scene's passes; ``job(i)``
    returns a HIPRendererKernel view of scene i for readbacks (or to render
    it alone).  Each scene's result equals rendering it alone."""

    def __init__(self, device=0):
        self._lib = _ffi.hip_lib()
        self._b = ctypes.c_void_p()
        rc = self._lib.df_batch_create(ctypes.byref(self._b), int(device))
        if rc != _ffi.DF_OK:
            self._b = None
            raise HIPError(rc, "df_batch_create(device=%d)" % device)
        self._jobs = []

    def _check(self, rc, what):
        if rc != _ffi.DF_OK:
            msg = self._lib.df_batch_last_error(self._b) if self._b else b""
            raise HIPError(rc, "%s: %s" % (what, (msg or b"").decode()))

    def add(self, scene, params=None, width=None, height=None, passes=1, seed=0):
        if not isinstance(scene, CompiledScene):
            raise TypeError("scene must be a CompiledScene")
        p = scene.copy_params() if params is None else params
        w = scene.width if width is None else int(width)
        h = scene.height if height is None else int(height)
        idx = ctypes.c_int32()
        self._check(self._lib.df_batch_add(self._b, ctypes.byref(scene.as_struct()), ctypes.byref(scene.sky),
                                           ctypes.byref(p), w, h, int(passes), ctypes.c_uint64(seed),
                                           ctypes.byref(idx)), "df_batch_add")
        self._jobs.append((scene, w, h))
        return idx.value

    def __len__(self):
        return len(self._jobs)

    def execute(self):
        self._check(self._lib.df_batch_execute(self._b), "df_batch_execute")

    def setGraphs(self, enabled):
        """df_batch_set_graphs: replay the batch's launches from a captured HIP graph (default off)."""
        self._check(self._lib.df_batch_set_graphs(self._b, 1 if enabled else 0), "df_batch_set_graphs")

    def executeNoAlloc(self):
        """df_batch_execute_noalloc: the same renders with no device allocation (work tables in the library's
        static per-GPU scratch array; at most 1024 scenes and 262144 tiles per call)."""
        self._check(self._lib.df_batch_execute_noalloc(self._b), "df_batch_execute_noalloc")

    def lastExecuteMillis(self):
        ms = ctypes.c_float()
        self._check(self._lib.df_batch_last_execute_ms(self._b, ctypes.byref(ms)), "df_batch_last_execute_ms")
        return ms.value

    def job(self, i):
        ctx = ctypes.c_void_p()
        self._check(self._lib.df_batch_job(self._b, int(i), ctypes.byref(ctx)), "df_batch_job")
        scene, w, h = self._jobs[i]
        return HIPRendererKernel._borrowed(self._lib, ctx, scene, w, h)

    def dispose(self):
        if self._b:
            self._lib.df_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass
