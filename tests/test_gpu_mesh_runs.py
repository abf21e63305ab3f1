"""GPU: runs of consecutive meshes, walked as one loop of node visits per lane, give the oracle's bits.

The trace, megakernel and adjoint kernels walk a run of consecutive mesh primitives as one per-lane loop of node
visits, a run being cut where the primitive loop's 64-primitive blocks end.  The scene here holds a run of 70 small
meshes (cut at the block end after 63 of them), four more runs of one or two large meshes separated by spheres and
a triangle, and a plane, so runs start and end at every kind of boundary.  Bars: exact math, both path-tracer
schedules, film / samples / float image / BGRA / primitive offsets / seeds bit for bit against the oracle.
"""

import numpy as np
import pytest

import dayflower_amd as dfa
from dayflower_amd import assets
from oracle import oracle_ffi

pytestmark = pytest.mark.gpu

W, H = 96, 64


def grid_mesh(nx, nz, size, height, phase):
    """A wavy nx x nz grid of quads (2 nx nz triangles), (n, 3, 11) vertices: position, normal, uv, tangent."""
    def vertex(i, k):
        x = (i / nx - 0.5) * size
        z = (k / nz - 0.5) * size
        y = height * np.sin(3.0 * x + phase) * np.cos(2.0 * z)
        return [x, y, z, 0.0, 1.0, 0.0, i / nx, k / nz, 1.0, 0.0, 0.0]
    tris = []
    for i in range(nx):
        for k in range(nz):
            a, b, c, d = vertex(i, k), vertex(i + 1, k), vertex(i + 1, k + 1), vertex(i, k + 1)
            tris.append([a, b, c])
            tris.append([a, c, d])
    return np.asarray(tris, np.float32)


def mesh_run_scene():
    with dfa.SceneBuilder("Mesh_Runs") as b:
        black = b.constant((0.0, 0.0, 0.0))
        ground = b.surface(assets.LAMBERTIAN, b.checkerboard((0.3, 0.3, 0.3), (0.9, 0.9, 0.9), 0.5, 0.5), black,
                           black)
        b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), ground)
        warm = b.surface(assets.LAMBERTIAN, b.constant((0.8, 0.5, 0.3)), black, black)
        cool = b.surface(assets.LAMBERTIAN, b.constant((0.3, 0.5, 0.8)), black, black)
        glass = b.surface(assets.GLASS, b.constant((1.0, 1.0, 1.0)), black, black)
        mirror = b.surface(assets.REFLECTION, b.constant((0.9, 0.9, 0.9)), black, black)
        small = grid_mesh(5, 10, 0.7, 0.1, 0.0)  # 100 triangles
        for j in range(70):  # primitives 1..70
            x = -4.0 + (j % 10) * 0.9
            z = -5.0 + (j // 10) * 0.9
            b.mesh(small, warm if j % 2 else cool, assets.transform((x, 0.4 + 0.05 * (j % 3), z)))
        b.sphere((0.0, 0.0, 0.0), 0.6, glass, assets.transform((-1.0, 0.6, 1.0)))  # 71
        b.mesh(grid_mesh(50, 60, 3.0, 0.3, 0.5), cool, assets.transform((2.5, 1.5, -1.0)))  # 72: 6000
        b.sphere((0.0, 0.0, 0.0), 0.5, mirror, assets.transform((1.0, 0.5, 1.5)))  # 73
        b.mesh(grid_mesh(50, 30, 2.0, 0.2, 1.0), warm, assets.transform((-2.5, 1.2, -2.0)))  # 74: 3000
        b.mesh(grid_mesh(50, 30, 2.0, 0.2, 2.0), cool, assets.transform((-2.0, 2.0, -3.0)))  # 75: 3000
        b.triangle(grid_mesh(1, 1, 3.0, 0.0, 0.0)[0], mirror, assets.transform((0.0, 3.0, -5.0)))  # 76
        b.mesh(grid_mesh(50, 50, 2.5, 0.25, 1.5), warm, assets.transform((0.5, 2.4, -4.0)))  # 77: 5000
        b.sphere((0.0, 0.0, 0.0), 0.4, glass, assets.transform((0.0, 0.4, 2.5)))  # 78
        b.mesh(grid_mesh(45, 50, 2.0, 0.2, 0.7), cool, assets.transform((3.0, 0.8, 1.0)))  # 79: 4500
        b.set_camera(dfa.camera_setup(W, H, 55.0, (0.0, 5.0, 8.0), (0.0, 0.5, -1.5)))
        b.set_sky(dfa.sky_at(45.0, 30.0, 2.5))
        return b.compile()


@pytest.fixture(scope="module")
def scene():
    return mesh_run_scene()


@pytest.mark.parametrize("engine", [dfa.PT_MEGAKERNEL, dfa.PT_WAVEFRONT])
def test_mesh_runs_bit_exact(gpu, scene, engine):
    sc = scene
    p = sc.copy_params()
    p.renderer_type = 2
    p.math_mode = dfa.MATH_EXACT
    p.pt_engine = engine
    passes = 2
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.updateCamera()
        k.seed(11)
        seeds0 = k.getSeeds()
        k.execute(passes)
        k.synchronize()
        film, samples = k.getFilm()
        prims, bgra, imgf, seeds = k.getPrimitiveOffsets(), k.getImageBGRA(), k.getImageFloat(), k.getSeeds()
    finally:
        k.dispose()
    st = oracle_ffi.OracleState(sc.width, sc.height, seeds0)
    oracle_ffi.render(sc, p, st, passes)
    np.testing.assert_array_equal(prims, st.primitive_offsets)
    np.testing.assert_array_equal(samples, st.film_samples)
    np.testing.assert_array_equal(film, st.film)
    np.testing.assert_array_equal(imgf, st.image_float)
    np.testing.assert_array_equal(bgra, st.image_bgra)
    np.testing.assert_array_equal(seeds, st.seeds)
    assert np.count_nonzero(prims >= 0) > 0.5 * prims.size
