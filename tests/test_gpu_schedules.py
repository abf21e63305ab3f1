"""Path-tracer schedules (df_set_schedule, df_get_schedule): every schedule the autotuner chooses among - the
megakernel with strip or square waves, the wavefront with one shade queue or with material-binned shading -
gives the oracle's bits in exact math and the same bits as each other in fast math; the autotuner settles on one
of them; a batch renders a job whose own schedule is the wavefront on that schedule, equal to the job alone."""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi

pytestmark = pytest.mark.gpu

SCHEDULES = [(dfa.PT_MEGAKERNEL, dfa.TILE_ROWS, 0), (dfa.PT_MEGAKERNEL, dfa.TILE_SQUARE, 0),
             (dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0), (dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 1)]


def _small(scene, w, h):
    cam = scene.camera.copy()
    cam[16], cam[17] = w, h
    return scene.with_camera(cam)


def _render(sc, p, schedule, passes, seed=42, size=None):
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(*(size or (sc.width, sc.height)))
        k.seed(seed)
        seeds0 = k.getSeeds()
        if schedule is not None:
            k.setSchedule(*schedule)
            got = k.schedule()
            assert (got["engine"], got["wave_tile"], got["shade_binned"], got["bvh_pairs"],
                    got["tail_bounce"])[:len(schedule)] == tuple(schedule)
        k.execute(passes)
        film, n = k.getFilm()
        return {"film": film, "n": n, "bgra": k.getImageBGRA(), "seeds": k.getSeeds(),
                "prims": k.getPrimitiveOffsets(), "seeds0": seeds0, "counters": k.counters()}
    finally:
        k.dispose()


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("cid,w,h", [(1, 96, 96), (2, 128, 96), (3, 160, 90)])
def test_every_schedule_is_bit_exact(scenes, gpu, cid, w, h, schedule):
    sc = _small(scenes(cid), w, h)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_EXACT
    g = _render(sc, p, schedule, 2)
    st = oracle_ffi.OracleState(sc.width, sc.height, g["seeds0"])
    oracle_ffi.render(sc, p, st, 2)
    np.testing.assert_array_equal(g["film"], st.film)
    np.testing.assert_array_equal(g["n"], st.film_samples)
    np.testing.assert_array_equal(g["bgra"], st.image_bgra)
    np.testing.assert_array_equal(g["seeds"], st.seeds)
    np.testing.assert_array_equal(g["prims"], st.primitive_offsets)
    for name in ("rays", "samples", "sky_evals"):
        assert g["counters"][name] == st.counters[name], name


@pytest.mark.parametrize("cid", [2, 3])
def test_schedules_agree_in_fast_math_at_full_size(scenes, gpu, cid):
    sc = scenes(cid)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_FAST
    ref = _render(sc, p, SCHEDULES[0], 1)
    for schedule in SCHEDULES[1:]:
        g = _render(sc, p, schedule, 1)
        for key in ("film", "n", "bgra", "seeds", "prims"):
            np.testing.assert_array_equal(g[key], ref[key], err_msg="%s %s" % (schedule, key))


@pytest.mark.parametrize("tail", [0, 2, 99])
def test_tail_kernel_gives_the_oracles_bits(scenes, gpu, tail):
    """tail_bounce picks which bounces' trace launches run the kernel built for short queues: 0 = all of them,
    2 = from the third bounce on, 99 = none (the escape launch always does)."""
    sc = _small(scenes(3), 160, 90)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_EXACT
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.seed(5)
        seeds0 = k.getSeeds()
        k.setSchedule(dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0, 0, tail)
        assert k.schedule()["tail_bounce"] == tail
        k.execute(2)
        film, n = k.getFilm()
        seeds = k.getSeeds()
    finally:
        k.dispose()
    st = oracle_ffi.OracleState(sc.width, sc.height, seeds0)
    oracle_ffi.render(sc, p, st, 2)
    np.testing.assert_array_equal(film, st.film)
    np.testing.assert_array_equal(n, st.film_samples)
    np.testing.assert_array_equal(seeds, st.seeds)


def test_tail_kernel_agrees_in_fast_math_at_full_size(scenes, gpu):
    sc = scenes(3)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_FAST
    ref = _render(sc, p, (dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0, 0, 99), 1)
    for tail in (0, dfa.TAIL_BOUNCE_DEFAULT):
        g = _render(sc, p, (dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0, 0, tail), 1)
        for key in ("film", "n", "bgra", "seeds", "prims"):
            np.testing.assert_array_equal(g[key], ref[key], err_msg="tail %d %s" % (tail, key))


def test_schedule_rejects_a_negative_tail_bounce(scenes, gpu):
    sc = _small(scenes(1), 32, 32)
    k = dfa.HIPRendererKernel(sc)
    try:
        with pytest.raises(dfa.HIPError):
            k.setSchedule(dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0, 0, -1)
    finally:
        k.dispose()


def test_autotuner_settles_on_a_candidate(scenes, gpu):
    sc = _small(scenes(3), 160, 90)
    p = sc.copy_params()
    p.math_mode = dfa.MATH_FAST
    p.pt_engine = dfa.PT_AUTO
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.seed(3)
        passes = 0
        while k.engine() == dfa.PT_AUTO:
            k.execute(1)
            passes += 1
            assert passes <= 8  # four candidates, two timed passes each
        got = k.schedule()
        assert (got["engine"], got["wave_tile"], got["shade_binned"]) in SCHEDULES
        k.execute(2)
        film = k.getFilm()[0]
    finally:
        k.dispose()
    ref = _render(sc, p, SCHEDULES[0], passes + 2, seed=3)
    np.testing.assert_array_equal(film, ref["film"])


def test_batch_renders_a_wavefront_job_on_its_own_schedule(scenes, gpu):
    big = _small(scenes(2), 192, 144)
    small = dfa.build_batch_scene(101)
    pb = big.copy_params()
    ps = small.copy_params()
    b = dfa.HIPBatchRenderer()
    try:
        b.add(big, params=pb, passes=2, seed=9)
        b.add(small, params=ps, width=64, height=48, passes=3, seed=10)
        b.job(0).setSchedule(dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 1)
        b.execute()
        films = [b.job(i).getFilm()[0].copy() for i in range(2)]
    finally:
        b.dispose()
    alone_big = _render(big, pb, SCHEDULES[0], 2, seed=9)
    np.testing.assert_array_equal(films[0], alone_big["film"])
    # the batch renders the scene's own camera array on a 64 x 48 frame
    alone_small = _render(small, ps, SCHEDULES[0], 3, seed=10, size=(64, 48))
    np.testing.assert_array_equal(films[1], alone_small["film"])
