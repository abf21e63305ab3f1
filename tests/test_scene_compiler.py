"""The host scene compiler: BVH invariants (BoundingVolumeHierarchy.java:805-986,
SceneCompiler.java:892-995), .scene file format (CompiledScene.java:273-377)
and the offset validation that guards every kernel launch."""

import struct

import numpy as np
import pytest

import dayflower_amd as dfa
from dayflower_amd import scene as dscene


def _parse_segment(bvh, base):
    """Pre-order nodes of one mesh segment: list of (rel, type, next, min, max, left_or_count, tris, depth)."""
    nodes = []
    rel, stack = 0, [0]  # depth of nodes still to read
    while stack:
        depth = stack.pop()
        n = bvh[base + rel: base + rel + 5]
        if n[0] == 1:
            nodes.append((rel, 1, n[1], n[2], n[3], n[4], [], depth))
            stack += [depth + 1, depth + 1]
            rel += 5
        else:
            tris = list(bvh[base + rel + 5: base + rel + 5 + n[4]])
            nodes.append((rel, 2, n[1], n[2], n[3], n[4], tris, depth))
            rel += 5 + n[4]
    return nodes


@pytest.mark.parametrize("cid", [2])
def test_bvh_invariants(scenes, cid):
    sc = scenes(cid)
    p3 = sc.point3fs.reshape(-1, 3)
    tri = sc.triangles.reshape(-1, 12)
    seen = np.zeros(len(tri), np.int32)
    for st, so, _ in sc.primitives.reshape(-1, 3):
        if st != 5:
            continue
        nodes = _parse_segment(sc.bvh, so)
        by_rel = {n[0]: n for n in nodes}
        for i, (rel, typ, nxt, mn, mx, x, tris, depth) in enumerate(nodes):
            lo, hi = p3[mn // 3], p3[mx // 3]
            assert np.all(lo <= hi)
            # skip pointer: the first later node (pre-order) at depth <= this depth
            later = [m for m in nodes[i + 1:] if m[7] <= depth]
            assert nxt == (later[0][0] if later else -1)
            if typ == 1:
                left = by_rel[x]
                assert x == rel + 5 and left[7] == depth + 1
                right = by_rel[left[2]]
                for child in (left, right):
                    assert np.all(p3[child[3] // 3] >= lo) and np.all(p3[child[4] // 3] <= hi)
            else:
                # leaves hold fewer than 4 triangles unless no split was found
                assert 1 <= x
                for t in tris:
                    seen[t // 12] += 1
                    pts = p3[tri[t // 12, :3] // 3]
                    assert np.all(pts >= lo) and np.all(pts <= hi)
        leaves = [n for n in nodes if n[1] == 2]
        assert sum(1 for n in leaves if n[5] >= 4) <= len(leaves) * 0.01
    assert np.all(seen == 1)  # every triangle in exactly one leaf


def test_workload_statistics(scenes):
    s1 = dscene.stats(scenes(1))
    assert s1["spheres"] == 3 and s1["planes"] == 1 and s1["triangles"] == 0
    s2 = dscene.stats(scenes(2))
    assert s2["triangles"] == 81920 + 20000
    s3 = dscene.stats(scenes(3))
    assert s3["spheres"] == 50 and s3["terrains"] == 1 and s3["triangles"] == 81920 + 20000 + 1


def test_scene_file_round_trip(tmp_path, scenes):
    sc = scenes(1)
    path = tmp_path / "c1.scene"
    sc.write(path)
    raw = path.read_bytes()
    # DataOutputStream.writeUTF: big-endian length then the name bytes
    (n,) = struct.unpack(">H", raw[:2])
    assert raw[2:2 + n].decode() == sc.name
    # first array: point2Fs length, big-endian
    (l0,) = struct.unpack(">i", raw[2 + n:6 + n])
    assert l0 == len(sc.point2fs)
    back = dfa.read_scene(path)
    for name in dscene.ARRAY_NAMES:
        np.testing.assert_array_equal(back.arrays[name], sc.arrays[name], err_msg=name)


def test_read_scene_rejects_truncated_file(tmp_path, scenes):
    path = tmp_path / "bad.scene"
    scenes(1).write(path)
    data = path.read_bytes()
    path.write_bytes(data[: len(data) // 2])
    with pytest.raises(RuntimeError):
        dfa.read_scene(path)


def test_validation_accepts_workloads(scenes):
    for cid in (1, 2, 3):
        rc, msg = scenes(cid).validate()
        assert rc == 0, msg
    for seed in (100, 101, 355):
        rc, msg = dfa.build_batch_scene(seed).validate()
        assert rc == 0, msg


def _corrupt(sc, name, index, value):
    arrays = {k: v.copy() for k, v in sc.arrays.items()}
    arrays[name][index] = value
    return dscene.CompiledScene(sc.name, arrays, sc.sky, sc.width, sc.height, sc.spp, sc.params)


@pytest.mark.parametrize("name,index,value", [
    ("primitives", 2, 10_000),        # surface offset past the surfaces array
    ("primitives", 1, 9_999),         # plane offset past the planes array
    ("surfaces", 1, 1e6),             # albedo texture offset past the textures
    ("spheres", 0, 3e6),              # sphere position offset past point3Fs
])
def test_validation_rejects_out_of_range_offsets(scenes, name, index, value):
    rc, msg = _corrupt(scenes(1), name, index, value).validate()
    assert rc == -1 and msg != "ok"


def test_validation_rejects_bad_bvh(scenes):
    sc = scenes(2)
    mesh_base = int(sc.primitives.reshape(-1, 3)[1, 1])
    # a skip pointer that points backwards would loop forever
    rc, msg = _corrupt(sc, "bvh", mesh_base + 5 + 1, 0).validate()
    assert rc == -1 and "BVH" in msg
    # a triangle offset past the triangle array
    bvh = sc.bvh
    rel = 0
    while bvh[mesh_base + rel] != 2:
        rel += 5
    rc, msg = _corrupt(sc, "bvh", mesh_base + rel + 5, 10 ** 8).validate()
    assert rc == -1


def test_batch_mix_is_uneven_and_deterministic():
    a = [dfa.build_batch_scene(s) for s in (100, 150, 200, 250, 300, 355)]
    b = dfa.build_batch_scene(150)
    np.testing.assert_array_equal(a[1].triangles, b.triangles)
    sizes = {x.width for x in a}
    assert len(sizes) > 1
    tri_counts = [dscene.stats(x)["triangles"] for x in a]
    assert max(tri_counts) > 10 * (min(tri_counts) + 1)


def test_camera_setup_orthonormal_and_looks_at_target():
    cam = dfa.camera_setup(320, 240, 40.0, (1.0, 2.0, 9.0), (0.0, 0.5, 0.0), 0.05, 8.0)
    u, v, w = cam[7:10], cam[10:13], cam[13:16]
    for a in (u, v, w):
        assert np.dot(a, a) == pytest.approx(1.0, abs=1e-6)
    assert abs(np.dot(u, v)) < 1e-6 and abs(np.dot(u, w)) < 1e-6 and abs(np.dot(v, w)) < 1e-6
    d = np.array([0.0, 0.5, 0.0]) - np.array([1.0, 2.0, 9.0])
    # the image-centre ray is W (GPURendererKernel.java:659-674)
    assert np.dot(w, d / np.linalg.norm(d)) == pytest.approx(1.0, abs=1e-6)
    # fovY from fovX and the aspect ratio (Camera.java:743-748)
    fy = np.degrees(2 * np.arctan(np.tan(np.radians(40.0) / 2) * 240 / 320))
    assert cam[5] == pytest.approx(fy, abs=1e-4)
    assert cam[0] == pytest.approx(0.05) and cam[6] == pytest.approx(8.0)


def test_sky_set_matches_preetham_zenith():
    # Preetham, Shirley, Smits 1999, appendix: zenith luminance and chromaticity.
    sun = np.array([np.cos(np.radians(45)) * 0.5, np.sin(np.radians(45)), np.cos(np.radians(45)) * 0.866])
    for t in (2.0, 3.0, 6.0):
        s = dfa.sky_set(tuple(sun), t)
        th = float(s.theta)
        chi = (4.0 / 9.0 - t / 120.0) * (np.float32(np.pi) - 2.0 * th)
        assert s.zenith_relative_luminance == pytest.approx(((4.0453 * t - 4.9710) * np.tan(chi) - 0.2155 * t + 2.4192) * 1000.0)
        assert th == pytest.approx(np.arccos(sun[1] / np.linalg.norm(sun)), abs=1e-6)
        assert s.sun_origin[1] == pytest.approx(10000.0 * sun[1] / np.linalg.norm(sun), rel=1e-6)
