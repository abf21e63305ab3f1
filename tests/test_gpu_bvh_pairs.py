"""The two-child walk (schedule bvh_pairs: the bounce traces test both children of a tree node from one 64-byte
record and keep the second child's decision key on a per-lane stack in LDS) visits the boxes and triangles the
reference's skip-pointer walk visits, in its order and against the same closest hit
(GPURendererKernel.java:1097-1194): exact math gives the oracle's bits on the mesh workloads and on a scene of
mesh runs cut at every kind of boundary; fast math gives the node walk's bits at full size; both walks count the
same triangle and primitive tests (and the same box tests but for the second boxes of tree nodes whose first
child's subtree ends an any-hit query)."""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi

pytestmark = pytest.mark.gpu

UNITS = ("box_tests", "tri_tests", "prim_tests", "sphere_tests", "terrain_steps")


def _small(scene, w, h):
    cam = scene.camera.copy()
    cam[16], cam[17] = w, h
    return scene.with_camera(cam)


def _render(sc, p, pairs, passes, seed=5):
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.updateCamera()
        k.setSchedule(dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0, 1 if pairs else 0)
        assert k.schedule()["bvh_pairs"] == (1 if pairs else 0)
        depth = k.bvhPairStack()
        k.seed(seed)
        seeds0 = k.getSeeds()
        k.execute(passes)
        film, n = k.getFilm()
        return {"film": film, "n": n, "bgra": k.getImageBGRA(), "seeds": k.getSeeds(), "imgf": k.getImageFloat(),
                "prims": k.getPrimitiveOffsets(), "seeds0": seeds0, "counters": k.counters(), "depth": depth}
    finally:
        k.dispose()


def _mesh_runs():
    from test_gpu_mesh_runs import mesh_run_scene
    return mesh_run_scene()


@pytest.mark.parametrize("case", ["w2", "w3", "runs"])
def test_pair_walk_is_bit_exact(scenes, gpu, case):
    sc = {"w2": lambda: _small(scenes(2), 128, 96), "w3": lambda: _small(scenes(3), 160, 90),
          "runs": _mesh_runs}[case]()
    p = sc.copy_params()
    p.renderer_type = 2
    p.math_mode = dfa.MATH_EXACT
    g = _render(sc, p, True, 2)
    assert 1 <= g["depth"] <= 16  # the scene is walked over the two-child records
    st = oracle_ffi.OracleState(sc.width, sc.height, g["seeds0"])
    oracle_ffi.render(sc, p, st, 2)
    for key, ref in (("film", st.film), ("n", st.film_samples), ("bgra", st.image_bgra), ("seeds", st.seeds),
                     ("prims", st.primitive_offsets), ("imgf", st.image_float)):
        np.testing.assert_array_equal(g[key], ref, err_msg=key)
    nodes = _render(sc, p, False, 2)
    for name in ("tri_tests", "prim_tests", "sphere_tests", "terrain_steps", "rays", "samples", "sky_evals"):
        assert g["counters"][name] == nodes["counters"][name], name
    # a tree node's two boxes are tested together, so an any-hit query that ends in the first child's subtree
    # has tested the second child's box too (the node walk never reaches it)
    assert nodes["counters"]["box_tests"] <= g["counters"]["box_tests"] <= 1.02 * nodes["counters"]["box_tests"]


@pytest.mark.parametrize("cid", [2, 3])
def test_pair_walk_equals_node_walk_at_full_size(scenes, gpu, cid):
    sc = scenes(cid)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_FAST
    a = _render(sc, p, False, 1)
    b = _render(sc, p, True, 1)
    assert b["depth"] >= 1
    for key in ("film", "n", "bgra", "seeds", "prims"):
        np.testing.assert_array_equal(b[key], a[key], err_msg=key)
    assert a["counters"]["box_tests"] <= b["counters"]["box_tests"] <= 1.02 * a["counters"]["box_tests"]
