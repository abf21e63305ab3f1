"""GPU: scenes built from assets render as the oracle renders them.

Bars (exact math, every renderer mode, bit for bit against the oracle: film, samples, BGRA, float image,
primitive offsets, seeds):
* the material showcase scene of Scenes.java:318-408 (plane + 15 spheres placed by transforms, every texture
  kind, an image albedo and an image normal map) with the two images it loads, the reference's own
  Texture_2.png and bricks2_normal.jpg (tests/golden);
* the icosphere OBJ model the reference's own BVH test loads, under a rotation + non-uniform scale + translation,
  with an image albedo and an image normal map (synthetic images), over a checkerboard plane;
* the reference's smoothMonkey0.obj (Suzanne with its own smooth normals) with Texture_2.png as albedo and
  bricks2_normal.jpg as normal map.
"""

import os

import numpy as np
import pytest

import dayflower_amd as dfa
from dayflower_amd import assets
from oracle import oracle_ffi

from test_assets import _synthetic_textures

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
W, H = 160, 90


def _obj_scene():
    t2, nrm = _synthetic_textures()
    with dfa.SceneBuilder("Icosphere_Scene") as b:
        black = b.constant((0.0, 0.0, 0.0))
        s = b.surface(assets.LAMBERTIAN, b.image(t2, 0.0, 1.0, 1.0), black, b.image(nrm, 30.0, 2.0, 2.0))
        b.load_obj(os.path.join(GOLDEN, "icosphere.obj"), s,
                   assets.transform((0.0, 1.2, 0.0), (0.2, 0.3, 0.1, 0.927), (1.0, 1.2, 0.8)))
        ground = b.surface(assets.LAMBERTIAN, b.checkerboard((0.3, 0.3, 0.3), (0.9, 0.9, 0.9), 0.5, 0.5), black,
                           black)
        b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), ground)
        b.set_camera(dfa.camera_setup(W, H, 50.0, (2.5, 2.0, 4.0), (0.0, 1.0, 0.0)))
        b.set_sky(dfa.sky_at(40.0, 30.0, 2.5))
        return b.compile()


def _monkey_scene():
    with dfa.SceneBuilder("Monkey_Scene") as b:
        black = b.constant((0.0, 0.0, 0.0))
        s = b.surface(assets.LAMBERTIAN, b.image(os.path.join(GOLDEN, "Texture_2.png"), 0.0, 1.0, 1.0), black,
                      b.image(os.path.join(GOLDEN, "bricks2_normal.jpg"), 0.0, 2.0, 2.0))
        b.load_obj(os.path.join(GOLDEN, "smoothMonkey0.obj"), s,
                   assets.transform((0.0, 1.1, 0.0), (0.0, 1.0, 0.0, 0.6), (1.0, 1.0, 1.0)))
        ground = b.surface(assets.LAMBERTIAN, b.checkerboard((0.3, 0.3, 0.3), (0.9, 0.9, 0.9), 0.5, 0.5), black,
                           black)
        b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), ground)
        b.set_camera(dfa.camera_setup(W, H, 45.0, (1.5, 2.2, 3.5), (0.0, 1.0, 0.0)))
        b.set_sky(dfa.sky_at(35.0, 60.0, 2.5))
        return b.compile()


@pytest.fixture(scope="module")
def asset_scenes():
    t2, nrm = os.path.join(GOLDEN, "Texture_2.png"), os.path.join(GOLDEN, "bricks2_normal.jpg")
    return {"showcase": assets.material_showcase_scene(t2, nrm, width=W, height=H), "obj": _obj_scene(),
            "monkey": _monkey_scene()}


@pytest.mark.parametrize("renderer", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("name", ["showcase", "obj", "monkey"])
def test_asset_scenes_render_bit_exact(gpu, asset_scenes, name, renderer):
    sc = asset_scenes[name]
    p = sc.copy_params()
    p.renderer_type = renderer
    p.math_mode = dfa.MATH_EXACT
    passes = 2
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.updateCamera()
        k.seed(7)
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
    if renderer != 4:  # the ray marcher draws its own height field, not the primitives
        assert np.count_nonzero(prims >= 0) > 0.2 * prims.size
