"""Hand-built compiled scenes for unit tests (flat arrays written directly
in the reference's layout, scene/compiler/SceneCompiler.java:634-1053)."""

import numpy as np

import dayflower_amd as dfa
from dayflower_amd.scene import CompiledScene


def argb(r, g, b):
    """image4j Color.pack() of an opaque colour, as the signed int stored in a float."""
    def q(c):
        c = min(max(c, 0.0), 1.0)
        return int(np.float32(np.float32(c) * np.float32(255.0) + np.float32(0.5)))
    v = (0xFF << 24) | (q(r) << 16) | (q(g) << 8) | q(b)
    return float(v - (1 << 32))


class SceneBuilder:
    """Minimal builder of the 14 arrays; every shape gets its own surface."""

    def __init__(self):
        self.p2 = [0.0, 0.0]
        self.p3 = []
        self.v3 = []
        self.planes = []
        self.spheres = []
        self.terrains = []
        self.triangles = []
        self.bvh = []
        self.textures = []
        self.surfaces = []
        self.primitives = []
        self.o2w = []
        self.w2o = []

    def point(self, p):
        self.p3 += [float(x) for x in p]
        return len(self.p3) - 3

    def vector(self, v):
        self.v3 += [float(x) for x in v]
        return len(self.v3) - 3

    def constant_texture(self, rgb):
        off = len(self.textures)
        self.textures += [4.0, 3.0, argb(*rgb)]
        return off

    def texture(self, record):
        off = len(self.textures)
        self.textures += [float(x) for x in record]
        return off

    def surface(self, material, albedo=(0.8, 0.8, 0.8), emission=(0.0, 0.0, 0.0), albedo_texture=None,
                noise_amount=0.0, noise_scale=0.0, normal_texture=None):
        a = self.constant_texture(albedo) if albedo_texture is None else albedo_texture
        e = self.constant_texture(emission)
        n = self.constant_texture((0.0, 0.0, 0.0)) if normal_texture is None else normal_texture
        off = len(self.surfaces)
        self.surfaces += [float(material), float(a), float(e), float(n), float(noise_amount), float(noise_scale)]
        return off

    def _prim(self, shape_type, shape_offset, surface, translation=(0.0, 0.0, 0.0)):
        tx, ty, tz = translation
        self.primitives += [shape_type, shape_offset, surface]
        self.o2w += [1, 0, 0, tx, 0, 1, 0, ty, 0, 0, 1, tz, 0, 0, 0, 1]
        self.w2o += [1, 0, 0, -tx, 0, 1, 0, -ty, 0, 0, 1, -tz, 0, 0, 0, 1]
        return len(self.primitives) - 3

    def sphere(self, center, radius, surface, translation=(0.0, 0.0, 0.0)):
        off = len(self.spheres)
        self.spheres += [float(self.point(center)), float(radius)]
        return self._prim(2, off, surface, translation)

    def plane(self, a, b, c, surface):
        a, b, c = (np.array(x, np.float32) for x in (a, b, c))
        n = np.cross(b - a, c - a).astype(np.float32)
        n = (n / np.float32(np.sqrt(np.float32(np.dot(n, n))))).astype(np.float32)
        off = len(self.planes)
        self.planes += [self.point(a), self.point(b), self.point(c), self.vector(n)]
        return self._prim(1, off, surface)

    def terrain(self, surface):
        off = len(self.terrains)
        self.terrains += [0.25, 0.5, -2.0, 2.0, 8.0]
        return self._prim(3, off, surface)

    def _triangle(self, a, b, c, normal=None):
        a, b, c = (np.array(x, np.float32) for x in (a, b, c))
        if normal is None:
            n = np.cross(b - a, c - a)
            n = n / np.linalg.norm(n)
        else:
            n = np.array(normal, np.float32)
        off = len(self.triangles)
        ni = self.vector(n)
        ti = self.vector((1.0, 0.0, 0.0))
        self.triangles += [self.point(a), self.point(b), self.point(c), ni, ni, ni, 0, 0, 0, ti, ti, ti]
        return off

    def triangle(self, a, b, c, surface):
        return self._prim(4, self._triangle(a, b, c), surface)

    def uv(self, uv):
        self.p2 += [float(x) for x in uv]
        return len(self.p2) - 2

    def triangle_attrs(self, abc, normals, uvs, tangents, surface):
        """A Triangle primitive with its own normal, texture coordinate and tangent at each vertex."""
        pts = [self.point(np.array(p, np.float32)) for p in abc]
        ns = [self.vector(n) for n in normals]
        us = [self.uv(u) for u in uvs]
        ts = [self.vector(t) for t in tangents]
        off = len(self.triangles)
        self.triangles += pts + ns + us + ts
        return self._prim(4, off, surface)

    def mesh_single_leaf(self, tris, surface):
        """A TriangleMesh whose BVH is one leaf holding every triangle."""
        offs = [self._triangle(*t) for t in tris]
        pts = np.array([p for t in tris for p in t], np.float32)
        base = len(self.bvh)
        self.bvh += [2, -1, self.point(pts.min(axis=0)), self.point(pts.max(axis=0)), len(offs)] + offs
        return self._prim(5, base, surface)

    def build(self, camera, sky=None, params=None, name="test"):
        def f(x):
            return np.array(x if x else [0.0], np.float32)

        def i(x):
            return np.array(x if x else [0], np.int32)

        arrays = {
            "point2fs": f(self.p2), "point3fs": f(self.p3), "vector3fs": f(self.v3), "planes": i(self.planes),
            "spheres": f(self.spheres), "terrains": f(self.terrains), "triangles": i(self.triangles),
            "bvh": i(self.bvh), "textures": f(self.textures), "surfaces": f(self.surfaces),
            "primitives": np.array(self.primitives, np.int32), "object_to_world": f(self.o2w),
            "world_to_object": f(self.w2o), "camera": np.ascontiguousarray(camera, np.float32),
        }
        if sky is None:
            sky = dfa.sky_set((1.0, 1.0, 1.0), 2.0)
        if params is None:
            params = dfa.default_params()
        return CompiledScene(name, arrays, sky, int(camera[16]), int(camera[17]), 1, params)


def default_camera(w=32, h=32, eye=(0.0, 2.0, 10.0), target=(0.0, 0.5, 0.0), fov=40.0, aperture=0.0, focal=10.0):
    return dfa.camera_setup(w, h, fov, eye, target, aperture, focal)
