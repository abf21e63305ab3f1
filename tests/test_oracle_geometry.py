"""Pins of the oracle's ray queries (doIntersectPrimitives,
GPURendererKernel.java:1048-1411) against closed forms, an FP64 brute force
and the BVH-vs-linear-scan invariant."""

import numpy as np
import pytest

import dayflower_amd as dfa
from oracle import oracle_ffi
from scene_helpers import SceneBuilder, default_camera

INF = 100000.0


def _params():
    return dfa.default_params()


def test_sphere_closed_form_and_transform():
    b = SceneBuilder()
    s = b.surface(3)
    b.sphere((0.0, 0.0, 0.0), 1.0, s, translation=(2.0, 0.0, 0.0))
    sc = b.build(default_camera())
    # ray along x from the origin: enters the unit sphere centred at x=2 at t=1
    t, rec = oracle_ffi.intersect(sc, _params(), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert t == pytest.approx(1.0, abs=1e-6)
    assert int(rec[1]) == 0
    np.testing.assert_allclose(rec[2:5], (1.0, 0.0, 0.0), atol=1e-6)   # world hit point
    np.testing.assert_allclose(rec[19:22], (-1.0, 0.0, 0.0), atol=1e-6)  # shading normal
    # from inside: the far root
    t, _ = oracle_ffi.intersect(sc, _params(), (2.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert t == pytest.approx(1.0, abs=1e-6)
    # miss
    t, rec = oracle_ffi.intersect(sc, _params(), (0.0, 5.0, 0.0), (1.0, 0.0, 0.0))
    assert t == INF and int(rec[1]) == -1


def test_plane_closed_form_and_uv():
    b = SceneBuilder()
    s = b.surface(3)
    b.plane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), s)
    sc = b.build(default_camera())
    t, rec = oracle_ffi.intersect(sc, _params(), (0.3, 5.0, -0.7), (0.0, -1.0, 0.0))
    assert t == 5.0
    # dominant-axis projection of the plane y=0: u = z, v = x (GPURendererKernel.java:2361-2396)
    assert rec[8] == pytest.approx(-0.7, abs=1e-6) and rec[9] == pytest.approx(0.3, abs=1e-6)
    # parallel ray: no hit
    t, _ = oracle_ffi.intersect(sc, _params(), (0.0, 1.0, 0.0), (1.0, 0.0, 0.0))
    assert t == INF
    # t below the 0.001 epsilon: no hit
    t, _ = oracle_ffi.intersect(sc, _params(), (0.0, 0.0005, 0.0), (0.0, -1.0, 0.0))
    assert t == INF


def _mt64(o, d, a, b, c):
    """Moller-Trumbore in float64 with the reference's epsilons."""
    e0, e1 = b - a, c - a
    v0 = np.cross(d, e1)
    det = np.einsum("ij,ij->i", e0, v0)
    ok = np.abs(det) > 1e-4
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    v1 = o - a
    u = np.einsum("ij,ij->i", v1, v0) * inv
    v2 = np.cross(v1, e0)
    v = (v2 @ d) * inv
    t = np.einsum("ij,ij->i", e1, v2) * inv
    hit = ok & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > 0.001)
    return np.where(hit, t, np.inf), u, v


def test_triangle_matches_fp64_brute_force():
    rng = np.random.default_rng(3)
    tris = [tuple(rng.uniform(-2, 2, 3) for _ in range(3)) for _ in range(40)]
    b = SceneBuilder()
    s = b.surface(3)
    for tr in tris:
        b.triangle(*tr, s)
    sc = b.build(default_camera())
    A = np.array([np.array(t[0], np.float32) for t in tris], np.float64)
    B = np.array([np.array(t[1], np.float32) for t in tris], np.float64)
    C = np.array([np.array(t[2], np.float32) for t in tris], np.float64)
    checked = 0
    for _ in range(400):
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        tgt = rng.uniform(-1, 1, 3)
        d = (tgt - o)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t32, rec = oracle_ffi.intersect(sc, _params(), o, d)
        t64, u, v = _mt64(o.astype(np.float64), d.astype(np.float64), A, B, C)
        order = np.argsort(t64)
        best = t64[order[0]]
        if not np.isfinite(best):
            assert t32 == INF
            continue
        # skip near-ties and grazing hits, where fp32 and fp64 may legitimately disagree
        second = t64[order[1]]
        k = order[0]
        if (np.isfinite(second) and abs(second - best) <= 1e-5 * best) or min(u[k], v[k], 1 - u[k] - v[k]) < 1e-4:
            continue
        assert t32 == pytest.approx(best, rel=1e-4)
        assert int(rec[1]) == 3 * k   # each triangle is its own primitive
        checked += 1
    assert checked > 50


def test_mesh_bvh_equals_linear_scan(scenes):
    """Closest hit through the skip-pointer BVH == FP64 closest hit over every
    triangle of the ~100k-triangle workload (plus the ground plane)."""
    sc = scenes(2)
    p3 = sc.point3fs.reshape(-1, 3).astype(np.float64)
    tri = sc.triangles.reshape(-1, 12)
    A, B, C = p3[tri[:, 0] // 3], p3[tri[:, 1] // 3], p3[tri[:, 2] // 3]
    prim = sc.primitives.reshape(-1, 3)
    # primitive of each triangle: meshes own the triangles listed in their BVH leaves
    owner = np.full(len(tri), -1)
    bvh = sc.bvh
    for pi, (st, so, _) in enumerate(prim):
        if st != 5:
            continue
        rel, todo = 0, 1
        while todo:
            n = bvh[so + rel: so + rel + 5]
            if n[0] == 1:
                todo += 1
                rel += 5
            else:
                owner[bvh[so + rel + 5: so + rel + 5 + n[4]] // 12] = pi
                todo -= 1
                rel += 5 + n[4]
    assert np.all(owner >= 0)
    rng = np.random.default_rng(11)
    cam = sc.camera
    eye = cam[1:4].astype(np.float64)
    checked = 0
    for _ in range(120):
        tgt = rng.uniform((-3, 0.5, -3), (3, 6, 3))
        d = tgt - eye
        d = (d / np.linalg.norm(d)).astype(np.float32)
        o = eye.astype(np.float32)
        t32, rec = oracle_ffi.intersect(sc, _params(), o, d)
        t64, u, v = _mt64(o.astype(np.float64), d.astype(np.float64), A, B, C)
        # ground plane y = 0
        tp = -o[1] / d[1] if d[1] < 0 else np.inf
        order = np.argsort(t64)
        k = order[0]
        best = min(t64[k], tp)
        if not np.isfinite(best):
            assert t32 == INF
            continue
        cands = np.sort(np.append(t64[order[:2]], tp))
        if np.isfinite(cands[1]) and cands[1] - cands[0] <= 1e-5 * cands[0]:
            continue
        if t64[k] < tp and min(u[k], v[k], 1 - u[k] - v[k]) < 1e-4:
            continue
        assert t32 == pytest.approx(best, rel=1e-4)
        expect_prim = 3 * owner[k] if t64[k] < tp else 0
        assert int(rec[1]) == expect_prim
        checked += 1
    assert checked > 60


def test_any_hit_agrees_with_closest_hit_on_hit_or_miss(scenes):
    sc = scenes(1)
    rng = np.random.default_rng(5)
    for _ in range(200):
        o = rng.uniform((-4, 0.1, -4), (4, 3, 4)).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        tc, _ = oracle_ffi.intersect(sc, _params(), o, d, testing=False)
        ta, _ = oracle_ffi.intersect(sc, _params(), o, d, testing=True)
        assert (tc < INF) == (ta < INF)
        assert ta >= tc


def test_terrain_march_against_dense_fp64_march():
    """The sin(x)sin(z) height field at scale 10 (GPURendererKernel.java:1470-1506)
    compared with a dense FP64 search for the first crossing."""
    b = SceneBuilder()
    s = b.surface(3)
    b.terrain(s)
    sc = b.build(default_camera())
    rng = np.random.default_rng(8)
    checked = 0
    for _ in range(60):
        o = np.array([rng.uniform(-5, 5), rng.uniform(2, 6), rng.uniform(-5, 5)], np.float32)
        d = np.array([rng.uniform(-1, 1), rng.uniform(-1, -0.2), rng.uniform(-1, 1)])
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t32, _ = oracle_ffi.intersect(sc, _params(), o, d)
        ts = np.linspace(0.001, 10.0, 400001)
        p = o.astype(np.float64)[None, :] / 10.0 + ts[:, None] * d.astype(np.float64)[None, :]
        below = p[:, 1] < np.sin(p[:, 0]) * np.sin(p[:, 2])
        if not below.any():
            continue
        t_true = ts[np.argmax(below)]
        if t_true < 0.05:
            continue
        # the march uses a 1% geometric step and a linear refinement
        assert t32 == pytest.approx(t_true, rel=0.011, abs=2e-3)
        checked += 1
    assert checked > 20


def test_terrain_normal_is_the_central_difference_gradient():
    """The terrain's normal (GPURendererKernel.java:2495-2539) is (h(x-e) - h(x+e), 2e, h(z-e) - h(z+e))
    normalised, e = 0.02, at the marched point: within O(e^2) of the analytic normal
    (-cos x sin z, 1, -sin x cos z) / |.| of h = sin x sin z."""
    b = SceneBuilder()
    b.terrain(b.surface(3))
    sc = b.build(default_camera())
    rng = np.random.default_rng(12)
    checked = 0
    for _ in range(60):
        o = np.array([rng.uniform(-5, 5), rng.uniform(2, 6), rng.uniform(-5, 5)], np.float32)
        d = np.array([rng.uniform(-1, 1), rng.uniform(-1, -0.2), rng.uniform(-1, 1)])
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t, rec = oracle_ffi.intersect(sc, _params(), o, d)
        if t >= INF:
            continue
        x, z = float(rec[2]), float(rec[4])
        n = np.array([-np.cos(x) * np.sin(z), 1.0, -np.sin(x) * np.cos(z)])
        n /= np.linalg.norm(n)
        np.testing.assert_allclose(rec[19:22], n, atol=2e-4)
        np.testing.assert_allclose(rec[5:8], rec[19:22], atol=0)
        assert rec[8] == rec[2] and rec[9] == rec[4]  # uv = (x, z) of the marched point
        checked += 1
    assert checked > 20
