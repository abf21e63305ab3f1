"""Frame-loop graphs (df_set_graphs, df_batch_set_graphs): a frame replayed from captured HIP graphs gives the
bits of the same frame launched eagerly, after every replay (each checked with the program's own readbacks and
synchronisation between replays), across changes that must re-capture (camera, sky, parameters, transforms,
clear), for both path-tracer schedules and for batches; and in exact math the graph path equals the oracle."""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi

pytestmark = pytest.mark.gpu


def _small(scene, w, h):
    cam = scene.camera.copy()
    cam[16], cam[17] = w, h
    return scene.with_camera(cam)


def _pair(sc, p):
    """Two renderers of one scene, one replaying graphs, one launching eagerly."""
    out = []
    for graphs in (True, False):
        k = dfa.HIPRendererKernel(sc)
        k.setGraphs(graphs)
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.updateCamera()
        k.seed(42)
        out.append(k)
    return out


def _state(k):
    film, n = k.getFilm()
    return film, n, k.getImageBGRA(), k.getSeeds(), k.getPrimitiveOffsets()


def _equal(a, b):
    for x, y in zip(_state(a), _state(b)):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("engine", [dfa.PT_MEGAKERNEL, dfa.PT_WAVEFRONT])
@pytest.mark.parametrize("cid,w,h", [(1, 96, 96), (2, 128, 96), (3, 160, 90)])
def test_graph_replays_equal_eager_launches(scenes, gpu, cid, w, h, engine):
    sc = _small(scenes(cid), w, h)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_FAST
    p.pt_engine = engine
    g, e = _pair(sc, p)
    try:
        for passes in (1, 1, 3, 1, 2):  # repeated replays of the same graphs, and other (clear, last) variants
            for k in (g, e):
                k.execute(passes)
            _equal(g, e)
        for k in (g, e):  # a cleared pass: the clear-flag variant
            k.clear()
            k.execute(2)
        _equal(g, e)
    finally:
        g.dispose()
        e.dispose()


@pytest.mark.parametrize("engine", [dfa.PT_MEGAKERNEL, dfa.PT_WAVEFRONT])
def test_graphs_follow_every_change(scenes, gpu, engine):
    """Camera, sky, parameters and transforms change the captured kernel arguments: each change is re-captured."""
    sc = _small(scenes(3), 128, 72)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_FAST
    p.pt_engine = engine
    g, e = _pair(sc, p)
    try:
        cam = dfa.camera_setup(128, 72, 45.0, (2.0, 4.0, 12.0), (0.0, 1.0, 0.0), 0.0, 10.0)
        o2w = sc.object_to_world.reshape(-1, 4, 4).copy()
        w2o = sc.world_to_object.reshape(-1, 4, 4).copy()
        o2w[:, 0, 3] += 0.3
        w2o[:, 0, 3] -= 0.3  # a translation along x composed with every primitive's transform
        p2 = sc.copy_params()
        p2.math_mode = dfa.MATH_FAST
        p2.pt_engine = engine
        p2.pt_depth_maximum = 5
        p2.tone_mapper_type = 1
        steps = [lambda k: k.updateCamera(cam), lambda k: k.updateSunAndSky(dfa.sky_at(30.0, 200.0, 4.0)),
                 lambda k: k.setParams(p2), lambda k: k.updatePrimitives(o2w.reshape(-1), w2o.reshape(-1)),
                 lambda k: k.toggleSun()]
        for change in steps:
            for k in (g, e):
                k.execute(1)
                change(k)
                k.execute(2)
            _equal(g, e)
    finally:
        g.dispose()
        e.dispose()


def test_graph_path_equals_the_oracle_in_exact_math(scenes, gpu):
    sc = _small(scenes(3), 96, 54)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_EXACT
    p.pt_engine = dfa.PT_WAVEFRONT
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setGraphs(True)
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.seed(5)
        st = oracle_ffi.OracleState(sc.width, sc.height, k.getSeeds())
        for passes in (1, 2, 1):
            k.execute(passes)
            oracle_ffi.render(sc, p, st, passes)
            np.testing.assert_array_equal(k.getFilm()[0], st.film)
            np.testing.assert_array_equal(k.getSeeds(), st.seeds)
        np.testing.assert_array_equal(k.getImageBGRA(), st.image_bgra)
    finally:
        k.dispose()


def test_batch_graph_replays_equal_eager_batches(gpu):
    out = []
    for graphs in (True, False):
        b = dfa.HIPBatchRenderer()
        try:
            b.setGraphs(graphs)
            for i, sd in enumerate((100, 101, 102, 103)):
                sc = dfa.build_batch_scene(sd)
                p = sc.copy_params()
                p.renderer_type = (2, 1, 2, 3)[i]
                b.add(sc, params=p, width=(64, 96, 48, 80)[i], height=(64, 40, 48, 72)[i], passes=2, seed=sd)
            films = []
            for _ in range(3):  # the second and third executes replay the first's graph
                b.execute()
                films.append([b.job(i).getFilm()[0].copy() for i in range(4)])
            out.append(films)
        finally:
            b.dispose()
    for fa, fb in zip(*out):
        for x, y in zip(fa, fb):
            np.testing.assert_array_equal(x, y)
