"""GPU: batches of independent scenes (df_batch_*) in one persistent launch.

Bar: every job of a batch equals the same scene rendered alone with
df_execute(passes) - film, sample counts, seeds, BGRA, primitive offsets and
work counters bit for bit - whatever else shares the launch; in exact math
mode, also the oracle.
"""

import ctypes

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi

pytestmark = pytest.mark.gpu


def _sized(scene, w, h):
    cam = scene.camera.copy()
    cam[16], cam[17] = w, h
    return scene.with_camera(cam)


def _alone(sc, p, passes, seed):
    """The scene rendered by itself (megakernel schedule, same seeds)."""
    q = dfa.default_params()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
    q.pt_engine = dfa.PT_MEGAKERNEL
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(q)
        k.update(sc.width, sc.height)
        k.seed(seed)
        k.execute(passes)
        k.synchronize()
        return _outputs(k)
    finally:
        k.dispose()


def _outputs(k):
    film, n = k.getFilm()
    return {"film": film, "n": n, "seeds": k.getSeeds(), "bgra": k.getImageBGRA(), "prims": k.getPrimitiveOffsets(),
            "counters": k.counters()}


def _assert_same(a, b):
    for key in ("film", "n", "seeds", "bgra", "prims"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert a["counters"] == b["counters"]


# (batch seed, width, height, passes, renderer, math): sizes that are not
# multiples of the 16x16 tile, several renderers and both math modes
JOBS = [(100, 40, 24, 3, 2, dfa.MATH_EXACT), (101, 64, 64, 2, 2, dfa.MATH_FAST), (102, 17, 33, 4, 2, dfa.MATH_EXACT),
        (103, 48, 32, 2, 3, dfa.MATH_FAST), (104, 32, 32, 2, 1, dfa.MATH_EXACT), (105, 80, 16, 1, 2, dfa.MATH_FAST),
        (106, 16, 16, 5, 6, dfa.MATH_FAST)]


def _job_scene(seed, w, h):
    return _sized(dfa.build_batch_scene(seed), w, h)


def test_batch_jobs_equal_each_scene_alone(gpu):
    b = dfa.HIPBatchRenderer()
    try:
        specs = []
        for seed, w, h, passes, renderer, math in JOBS:
            sc = _job_scene(seed, w, h)
            p = sc.copy_params()
            p.renderer_type = renderer
            p.math_mode = math
            b.add(sc, p, passes=passes, seed=1000 + seed)
            specs.append((sc, p, passes, 1000 + seed))
        b.execute()
        assert b.lastExecuteMillis() > 0.0
        for i, (sc, p, passes, seed) in enumerate(specs):
            got = _outputs(b.job(i))
            _assert_same(got, _alone(sc, p, passes, seed))
    finally:
        b.dispose()


def test_batch_second_execute_continues_each_film(gpu):
    """Two executes of a batch equal one render of 2 x passes alone."""
    b = dfa.HIPBatchRenderer()
    try:
        sc0, sc1 = _job_scene(110, 32, 24), _job_scene(111, 24, 40)
        p0, p1 = sc0.copy_params(), sc1.copy_params()
        p0.math_mode = p1.math_mode = dfa.MATH_EXACT
        b.add(sc0, p0, passes=2, seed=7)
        b.add(sc1, p1, passes=3, seed=8)
        b.execute()
        b.execute()
        _assert_same(_outputs(b.job(0)), _alone(sc0, p0, 4, 7))
        _assert_same(_outputs(b.job(1)), _alone(sc1, p1, 6, 8))
    finally:
        b.dispose()


def test_batch_job_matches_oracle(gpu):
    sc = _job_scene(120, 40, 32)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_EXACT
    b = dfa.HIPBatchRenderer()
    try:
        b.add(_job_scene(121, 64, 48), passes=2, seed=3)  # shares the launch
        i = b.add(sc, p, passes=2, seed=99)
        k = b.job(i)
        seeds0 = k.getSeeds()
        b.execute()
        got = _outputs(b.job(i))
    finally:
        b.dispose()
    st = oracle_ffi.OracleState(sc.width, sc.height, seeds0)
    oracle_ffi.render(sc, p, st, 2)
    np.testing.assert_array_equal(got["film"], st.film)
    np.testing.assert_array_equal(got["n"], st.film_samples)
    np.testing.assert_array_equal(got["bgra"], st.image_bgra)
    np.testing.assert_array_equal(got["prims"], st.primitive_offsets)
    np.testing.assert_array_equal(got["seeds"], st.seeds)


def test_batch_errors(gpu):
    b = dfa.HIPBatchRenderer()
    try:
        b.execute()  # an empty batch is a no-op
        with pytest.raises(dfa.HIPError) as e:
            b.job(0)
        assert e.value.code == -1
        sc = _job_scene(130, 16, 16)
        with pytest.raises(dfa.HIPError):
            b.add(sc, passes=-1)
        bad = sc.copy_params()
        bad.renderer_type = 9
        with pytest.raises(dfa.HIPError):
            b.add(sc, bad, passes=1)
        assert len(b) == 0
    finally:
        b.dispose()
