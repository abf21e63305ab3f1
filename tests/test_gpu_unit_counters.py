"""Unit counters (df_set_unit_counters): with them off the wavefront's trace kernels give the same bits, rays,
samples and sky evaluations and leave the unit counters at zero; counting the same passes again from the same
seeds gives the counts of a run that counted throughout, within the oracle's bound."""

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


def _run(k, seeds, passes, units):
    k.setSeeds(seeds)
    k.clear()
    k.setUnitCounters(units)
    k.resetCounters()
    k.execute(passes)
    film, n = k.getFilm()
    return {"film": film, "n": n, "bgra": k.getImageBGRA(), "seeds": k.getSeeds(), "counters": k.counters()}


@pytest.mark.parametrize("math", [dfa.MATH_EXACT, dfa.MATH_FAST])
@pytest.mark.parametrize("cid,w,h", [(2, 128, 96), (3, 160, 90)])
def test_unit_counters_off_same_bits(scenes, gpu, cid, w, h, math):
    sc = _small(scenes(cid), w, h)
    p = sc.copy_params()
    p.math_mode = math
    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(p)
        k.update(sc.width, sc.height)
        k.setSchedule(dfa.PT_WAVEFRONT, dfa.TILE_ROWS, 0)
        k.seed(7)
        seeds0 = k.getSeeds()
        on = _run(k, seeds0, 2, True)
        off = _run(k, seeds0, 2, False)
        again = _run(k, seeds0, 2, True)
    finally:
        k.dispose()
    for key in ("film", "n", "bgra", "seeds"):
        np.testing.assert_array_equal(off[key], on[key], err_msg=key)
    for name in ("rays", "samples", "sky_evals"):
        assert off["counters"][name] == on["counters"][name], name
    for name in UNITS:
        assert off["counters"][name] == 0, name
        assert again["counters"][name] == on["counters"][name], name
    assert on["counters"]["box_tests"] > 0
    if math == dfa.MATH_EXACT:
        st = oracle_ffi.OracleState(sc.width, sc.height, seeds0)
        oracle_ffi.render(sc, p, st, 2)
        np.testing.assert_array_equal(off["film"], st.film)
        for name in UNITS:  # never more work than the oracle's
            assert on["counters"][name] <= st.counters[name], name
