"""df_batch_execute_noalloc through the C ABI: the allocation-free batch execution renders every job exactly as
df_batch_execute does (film, sample counts, BGRA, seeds bit for bit), on two calls in a row whose second has
fewer and smaller scenes (the shared scratch array still holds the first call's larger tables), and rejects a
batch beyond its tile limit with DF_EINVAL before rendering anything."""

import ctypes

import numpy as np
import pytest

import dayflower_amd as dfa
from dayflower_amd import _ffi

pytestmark = pytest.mark.gpu


def _mix(seeds, sizes, passes):
    """Batch scenes of the mix with a spread of renderers and math modes (several launch groups)."""
    jobs = []
    for k, (sd, (w, h), n) in enumerate(zip(seeds, sizes, passes)):
        sc = dfa.build_batch_scene(sd)
        p = sc.copy_params()
        p.renderer_type = (2, 2, 1, 3, 2, 6)[k % 6]
        p.math_mode = dfa.MATH_EXACT if k % 3 == 1 else dfa.MATH_FAST
        jobs.append((sc, p, w, h, n, 1000 + sd))
    return jobs


def _render(jobs, noalloc):
    b = dfa.HIPBatchRenderer()
    try:
        for sc, p, w, h, n, seed in jobs:
            b.add(sc, params=p, width=w, height=h, passes=n, seed=seed)
        lib = _ffi.hip_lib()
        rc = (lib.df_batch_execute_noalloc if noalloc else lib.df_batch_execute)(b._b)
        assert rc == _ffi.DF_OK, lib.df_batch_last_error(b._b)
        out = []
        for i in range(len(jobs)):
            j = b.job(i)
            film, samples = j.getFilm()
            out.append((film, samples, j.getImageBGRA(), j.getSeeds()))
        return out
    finally:
        b.dispose()


def _same(a, b):
    for (fa, na, ia, sa), (fb, nb, ib, sb) in zip(a, b):
        np.testing.assert_array_equal(fa, fb)
        np.testing.assert_array_equal(na, nb)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(sa, sb)


def test_noalloc_batch_equals_batch_execute_on_two_calls(gpu):
    first = _mix([100, 101, 102, 103, 104, 105], [(96, 80), (130, 64), (64, 64), (200, 120), (48, 40), (80, 80)],
                 [3, 2, 2, 1, 4, 2])
    _same(_render(first, noalloc=True), _render(first, noalloc=False))
    # a smaller second call: fewer jobs and tiles than the scratch array still holds
    second = _mix([110, 111], [(40, 24), (33, 17)], [2, 3])
    _same(_render(second, noalloc=True), _render(second, noalloc=False))


def test_noalloc_batch_rejects_more_tiles_than_its_maximum(gpu):
    sc = dfa.build_batch_scene(100)
    b = dfa.HIPBatchRenderer()
    try:
        b.add(sc, width=8192, height=8200, passes=1, seed=1)  # 512 x 513 = 262656 tiles > 262144
        lib = _ffi.hip_lib()
        rc = lib.df_batch_execute_noalloc(b._b)
        assert rc == _ffi.DF_EINVAL
        assert b"262144" in lib.df_batch_last_error(b._b)
        ms = ctypes.c_float()
        assert lib.df_batch_last_execute_ms(b._b, ctypes.byref(ms)) == _ffi.DF_ESTATE  # nothing was launched
    finally:
        b.dispose()
