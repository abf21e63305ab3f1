"""Scene construction and asset ingest: Wavefront OBJ models and image textures.

``SceneBuilder`` drives the scene library's builder (include/dayflower_scene.h): textures, surfaces and
primitives (shape + surface + Transform) as the reference constructs them in Java (scene/Scene.java,
scene/Primitive.java, scene/Surface.java, scene/texture, scene/shape), OBJ models through
ObjectLoader.load (scene/wavefront/ObjectLoader.java:118-212, restated in C++), then SceneCompiler.compile.

``image_rgb(path)`` decodes an image into what ImageTexture.load keeps (ImageTexture.java:470-548): the
pixels of a TYPE_INT_RGB BufferedImage, 0x00RRGGBB per pixel, rows from the top.  The reference reads the file
with ImageIO and draws it onto a black TYPE_INT_RGB image; here Pillow decodes and the draw is restated:
opaque pixels are copied, grey is replicated into R, G and B, and a pixel with alpha a below 255 becomes
round(a * c / 255) per channel (Java2D's source-over onto black, whose 8-bit multiply table holds exactly
these values).  Decoding a PNG is exact; a JPEG decodes with libjpeg's integer IDCT and
fancy upsampling, the defaults of both ImageIO's and Pillow's decoders.  Images with 16-bit channels, CMYK or
floating-point pixels are refused (ImageIO converts them through colour spaces this does not restate).

``material_showcase_scene`` is Scenes.newMaterialShowcaseScene (scene/loader/Scenes.java:318-408) given the
two images it loads.
"""

import ctypes

import numpy as np

from . import _ffi
from .scene import _check, _from_handle

CLEAR_COAT, GLASS, LAMBERTIAN, PHONG, REFLECTION = 1, 2, 3, 4, 5

BLACK = (0.0, 0.0, 0.0)
GRAY = (0.5, 0.5, 0.5)
RED = (1.0, 0.0, 0.0)
WHITE = (1.0, 1.0, 1.0)


def transform(position=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0, 1.0), scale=(1.0, 1.0, 1.0)):
    """Transform(position, rotation quaternion (x, y, z, w), scale) (scene/Transform.java)."""
    t = _ffi.Transform()
    t.position[:] = [float(v) for v in position]
    t.rotation[:] = [float(v) for v in rotation]
    t.scale[:] = [float(v) for v in scale]
    return t


def _mul8(a, c):
    return (a.astype(np.int64) * c.astype(np.int64) + 127) // 255


def image_rgb(path):
    """(width, height, int32 pixels) of ImageTexture.load's TYPE_INT_RGB raster for the image at `path`."""
    from PIL import Image

    with Image.open(path) as im:
        im.load()
        mode = im.mode
        if mode == "P":
            im = im.convert("RGBA")
        elif mode == "1":
            im = im.convert("L")
        elif mode not in ("RGB", "RGBA", "L", "LA"):
            raise ValueError("%s: %s images are not supported (16-bit, CMYK and float pixels go through colour "
                             "spaces ImageIO handles differently)" % (path, mode))
        a = np.asarray(im)
        w, h = im.size
    if a.ndim == 2:
        rgb = np.repeat(a[:, :, None], 3, axis=2).astype(np.int64)
        alpha = None
    elif a.shape[2] == 2:  # LA
        rgb = np.repeat(a[:, :, :1], 3, axis=2).astype(np.int64)
        alpha = a[:, :, 1].astype(np.int64)
    else:
        rgb = a[:, :, :3].astype(np.int64)
        alpha = a[:, :, 3].astype(np.int64) if a.shape[2] == 4 else None
    if alpha is not None:
        al = np.repeat(alpha[:, :, None], 3, axis=2)
        rgb = np.where(al == 255, rgb, _mul8(al, rgb))
    packed = (rgb[:, :, 0] << 16) | (rgb[:, :, 1] << 8) | rgb[:, :, 2]
    return w, h, packed.astype(np.int32).reshape(-1)


class SceneBuilder:
    """new Scene(name) with Camera() and Sky() defaults; add textures, surfaces and primitives, then compile()."""

    def __init__(self, name="Scene"):
        self._lib = _ffi.scene_lib()
        self._h = ctypes.c_void_p()
        _check(self._lib.dfs_builder_create(str(name).encode(), ctypes.byref(self._h)), "dfs_builder_create")
        self._mappers = []

    def close(self):
        if self._h:
            self._lib.dfs_builder_free(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- plumbing
    def _call(self, name, *args):
        rc = getattr(self._lib, name)(self._h, *args)
        if rc != _ffi.DF_OK:
            buf = ctypes.create_string_buffer(512)
            self._lib.dfs_builder_last_error(self._h, buf, 512)
            raise RuntimeError("%s failed (%s): %s" % (name, _ffi.ERROR_NAMES.get(rc, rc), buf.value.decode()))

    def _index(self, name, *args):
        out = ctypes.c_int32()
        self._call(name, *args, ctypes.byref(out))
        return out.value

    @staticmethod
    def _f3(v):
        return (ctypes.c_float * 3)(*[float(x) for x in v])

    @staticmethod
    def _t(t):
        return ctypes.byref(t) if t is not None else None

    # ---- textures
    def constant(self, rgb):
        """ConstantTexture(color)."""
        return self._index("dfs_builder_constant_texture", self._f3(rgb))

    def checkerboard(self, color0=GRAY, color1=WHITE, scale_u=5.0, scale_v=5.0, degrees=0.0):
        """CheckerboardTexture(color0, color1, scaleU, scaleV, degrees); defaults as CheckerboardTexture(c0, c1)."""
        return self._index("dfs_builder_checkerboard_texture", self._f3(color0), self._f3(color1), float(scale_u),
                           float(scale_v), float(degrees))

    def fbm(self, addend, multiplier, frequency, gain, octaves):
        """FractionalBrownianMotionTexture(addend, multiplier, frequency, gain, octaves)."""
        return self._index("dfs_builder_fbm_texture", self._f3(addend), self._f3(multiplier), float(frequency),
                           float(gain), int(octaves))

    def blend(self, texture_a, texture_b, factor):
        """BlendTexture(textureA, textureB, factor)."""
        return self._index("dfs_builder_blend_texture", int(texture_a), int(texture_b), float(factor))

    def bullseye(self, color_a, color_b):
        return self._index("dfs_builder_bullseye_texture", self._f3(color_a), self._f3(color_b))

    def uv(self):
        return self._index("dfs_builder_uv_texture")

    def surface_normal(self, tangent_space=False):
        return self._index("dfs_builder_surface_normal_texture", 1 if tangent_space else 0)

    def image(self, source, degrees=0.0, scale_u=1.0, scale_v=1.0):
        """ImageTexture.load(file, degrees, scaleU, scaleV) for a path, or ImageTexture(degrees, width, height,
        scaleU, scaleV, data) for a (width, height, pixels) triple as image_rgb returns."""
        w, h, data = image_rgb(source) if isinstance(source, (str, bytes)) or hasattr(source, "__fspath__") \
            else source
        d = np.ascontiguousarray(data, np.int32).reshape(-1)
        if d.size != int(w) * int(h):
            raise ValueError("image data holds %d pixels, not %d x %d" % (d.size, w, h))
        return self._index("dfs_builder_image_texture", d.ctypes.data_as(_ffi.c_int32_p), int(w), int(h),
                           float(degrees), float(scale_u), float(scale_v))

    # ---- surfaces and primitives
    def surface(self, material, albedo, emission, normal, noise_amount=0.0, noise_scale=0.0):
        """Surface(material, textureAlbedo, textureEmission, textureNormal, noiseAmount, noiseScale)."""
        return self._index("dfs_builder_surface", int(material), int(albedo), int(emission), int(normal),
                           float(noise_amount), float(noise_scale))

    def plane(self, a, b, c, surface, transform=None):
        self._call("dfs_builder_plane", self._f3(a), self._f3(b), self._f3(c), int(surface), self._t(transform))

    def sphere(self, center, radius, surface, transform=None):
        self._call("dfs_builder_sphere", self._f3(center), float(radius), int(surface), self._t(transform))

    def terrain(self, frequency, gain, minimum, maximum, octaves, surface, transform=None):
        self._call("dfs_builder_terrain", float(frequency), float(gain), float(minimum), float(maximum), int(octaves),
                   int(surface), self._t(transform))

    def triangle(self, vertices, surface, transform=None):
        """Triangle(a, b, c): vertices (3, 11) = position, normal, uv, tangent per vertex."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
        if v.size != 33:
            raise ValueError("a triangle takes 3 x 11 floats")
        self._call("dfs_builder_triangle", v.ctypes.data_as(_ffi.c_float_p), int(surface), self._t(transform))

    def mesh(self, vertices, surface, transform=None):
        """TriangleMesh(triangles): vertices (n, 3, 11)."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
        if v.size == 0 or v.size % 33:
            raise ValueError("a mesh takes n x 3 x 11 floats")
        self._call("dfs_builder_mesh", v.ctypes.data_as(_ffi.c_float_p), ctypes.c_int64(v.size // 33), int(surface),
                   self._t(transform))

    def load_obj(self, path, surface=None, transform=None):
        """ObjectLoader.load(file, surfaceMapper, transform): `surface` is a surface index for every mesh, a
        callable (group, material) -> surface index, a dict keyed by material name (missing keys: its None
        entry, as surfaces.getOrDefault(materialName, default)), or None for the reference's default
        Surface(Lambertian, GRAY, BLACK, BLACK).  Returns the number of primitives added."""
        mapper = None
        fixed = -1
        if callable(surface):
            fn = surface
        elif isinstance(surface, dict):
            table = dict(surface)
            fn = lambda g, m: table.get(m, table.get(None, -1))  # noqa: E731
        else:
            fn = None
            fixed = -1 if surface is None else int(surface)
        if fn is not None:
            def thunk(g, m, _user):
                try:
                    return int(fn(g.decode(), m.decode()))
                except Exception:
                    return -1
            mapper = _ffi.SurfaceMapper(thunk)
            self._mappers.append(mapper)
        else:
            mapper = _ffi.SurfaceMapper()
        n = ctypes.c_int32()
        self._call("dfs_builder_load_obj", str(path).encode(), mapper, None, fixed, self._t(transform),
                   ctypes.byref(n))
        return n.value

    # ---- camera, sky, compile
    def set_camera(self, camera):
        cam = np.ascontiguousarray(camera, np.float32)
        if cam.size != 19:
            raise ValueError("a camera is 19 floats")
        self._call("dfs_builder_set_camera", cam.ctypes.data_as(_ffi.c_float_p))

    def set_sky(self, sky):
        self._call("dfs_builder_set_sky", ctypes.byref(sky))

    def compile(self):
        """SceneCompiler.compile: a CompiledScene."""
        h = ctypes.c_void_p()
        self._call("dfs_builder_compile", ctypes.byref(h))
        try:
            return _from_handle(h)
        finally:
            self._lib.dfs_free(h)


def default_camera(eye=None, width=None, height=None):
    """Camera() defaults, then setEye / setResolution as given."""
    lib = _ffi.scene_lib()
    cam = np.zeros(19, np.float32)
    p = cam.ctypes.data_as(_ffi.c_float_p)
    _check(lib.dfs_camera_default(p), "dfs_camera_default")
    if eye is not None:
        _check(lib.dfs_camera_set_eye(p, float(eye[0]), float(eye[1]), float(eye[2])), "dfs_camera_set_eye")
    if width is not None:
        _check(lib.dfs_camera_set_resolution(p, float(width), float(height)), "dfs_camera_set_resolution")
    return cam


def material_showcase_scene(texture_2, bricks2_normal, width=None, height=None):
    """Scenes.newMaterialShowcaseScene (scene/loader/Scenes.java:318-408).  `texture_2` and `bricks2_normal`
    stand for Texture_2.png and bricks2_normal.jpg: paths, or (width, height, pixels) triples."""
    with SceneBuilder("Material_Showcase_Scene") as b:
        black = lambda: b.constant(BLACK)  # noqa: E731
        ground = b.surface(LAMBERTIAN, b.constant(GRAY), black(), black())
        b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), ground)
        spheres = [
            (CLEAR_COAT, b.constant(GRAY), 0.0, 0.0, None, 40.0),
            (GLASS, b.constant(GRAY), 0.0, 0.0, None, 80.0),
            (LAMBERTIAN, b.constant(GRAY), 0.0, 0.0, None, 120.0),
            (REFLECTION, b.constant(GRAY), 0.0, 0.0, None, 160.0),
            (PHONG, b.constant(GRAY), 0.0, 0.0, None, 200.0),
            (LAMBERTIAN, b.checkerboard(GRAY, WHITE), 0.0, 0.0, None, 280.0),
            (LAMBERTIAN, b.constant(RED), 0.0, 0.0, None, 320.0),
            (LAMBERTIAN, b.fbm((0.5, 0.05, 0.05), WHITE, 0.8, 0.5, 16), 0.0, 0.0, None, 360.0),
            (LAMBERTIAN, b.image(texture_2, 0.0, 1.0, 1.0), 0.0, 0.0, None, 400.0),
            (LAMBERTIAN, b.surface_normal(), 0.0, 0.0, None, 440.0),
            (LAMBERTIAN, b.blend(b.fbm((0.05, 0.05, 0.05), WHITE, 0.5, 0.5, 16), b.checkerboard(GRAY, RED), 0.5),
             0.0, 0.0, None, 480.0),
            (LAMBERTIAN, b.uv(), 0.0, 0.0, None, 520.0),
            (LAMBERTIAN, b.bullseye((1.0, 0.1, 0.1), (0.1, 1.0, 0.1)), 0.0, 0.0, None, 560.0),
            (REFLECTION, b.constant(GRAY), 0.0, 0.0, "bricks", 640.0),
            (GLASS, b.constant(WHITE), 1.0, 16.0, None, 680.0),
        ]
        for material, albedo, noise_amount, noise_scale, normal, x in spheres:
            emission = black()
            nrm = b.image(bricks2_normal, 0.0, 4.0, 4.0) if normal == "bricks" else black()
            s = b.surface(material, albedo, emission, nrm, noise_amount, noise_scale)
            b.sphere((0.0, 0.0, 0.0), 16.5, s, transform((x, 16.5, 20.0)))
        b.set_camera(default_camera(eye=(295.0, 42.0, 332.6), width=width, height=height))
        return b.compile()
