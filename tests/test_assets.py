"""CPU: scene construction and asset ingest (Wavefront OBJ models, image textures, transforms).

Pins:
* ObjectLoader behaviour (scene/wavefront/ObjectLoader.java:118-637) on the icosphere model the reference's own
  BoundingVolumeHierarchyTest loads (tests/golden/icosphere.obj) and on hand-written files: fan triangulation,
  shared vertices, normals computed as the normalized sum of unit face normals, the tangent formula whose z
  repeats y, v flipped to 1 - v, mesh splits where the group changes, the per-file scale of 100, and the
  file-wide texture-coordinate flag;
* transforms against Rodrigues' rotation formula and the inverse they must be;
* image decoding against the pixels written (PNG is lossless) and the source-over-black rule in closed form;
* the material showcase scene of Scenes.java:318-408 compiles with its 16 primitives and both image textures.
"""

import os

import numpy as np
import pytest

import dayflower_amd as dfa
from dayflower_amd import assets

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _tris(sc):
    """Per triangle: positions (3, 3), normals (3, 3), uvs (3, 2), tangents (3, 3) from the compiled arrays."""
    t = sc.triangles.reshape(-1, 12)
    p3, v3, p2 = sc.point3fs.reshape(-1, 3), sc.vector3fs.reshape(-1, 3), sc.point2fs.reshape(-1, 2)
    return p3[t[:, 0:3] // 3], v3[t[:, 3:6] // 3], p2[t[:, 6:9] // 2], v3[t[:, 9:12] // 3]


def _load(path, surface=None, transform=None):
    with dfa.SceneBuilder("obj") as b:
        n = b.load_obj(path, surface, transform)
        return n, b.compile()


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def _obj_tokens(path):
    v, vt, vn, f = [], [], [], []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            v.append([float(x) for x in t[1:4]])
        elif t[0] == "vt":
            vt.append([float(x) for x in t[1:3]])
        elif t[0] == "vn":
            vn.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            f.append([[int(x) - 1 for x in c.split("/")] for c in t[1:]])
    return (np.array(v, np.float32), np.array(vt, np.float32), np.array(vn, np.float32), f)


def test_icosphere_loads_as_one_mesh_with_its_attributes():
    path = os.path.join(GOLDEN, "icosphere.obj")
    n, sc = _load(path)
    assert n == 1 and sc.primitive_count == 1 and sc.primitives[0] == 5
    v, vt, vn, faces = _obj_tokens(path)
    assert len(faces) == 80 and all(len(c) == 3 for c in faces)
    pos, nrm, uv, tan = _tris(sc)
    assert pos.shape[0] == 80
    for i, face in enumerate(faces):
        for k, (vi, ti, ni) in enumerate(face):
            np.testing.assert_array_equal(pos[i, k], v[vi])
            np.testing.assert_array_equal(nrm[i, k], vn[ni])  # given normals are kept as written
            assert uv[i, k, 0] == vt[ti, 0]
            assert uv[i, k, 1] == np.float32(np.float32(1.0) - vt[ti, 1])  # flipTextureCoordinatesForY
    # tangents: unit length, and the reference's z component repeats the y terms before normalization
    lens = np.linalg.norm(tan.astype(np.float64), axis=2)
    np.testing.assert_allclose(lens, 1.0, atol=1e-6)
    np.testing.assert_array_equal(tan[..., 1], tan[..., 2])
    # BVH over the 80 triangles, every triangle in one leaf
    stats = dfa.scene.stats(sc)
    assert stats["triangles"] == 80 and stats["bvh_nodes"] > 1


def test_quad_fan_default_surface_and_computed_normals(tmp_path):
    path = _write(tmp_path, "quad.obj", "v 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nv 0 1 0\n"
                                        "f 1 4 3 2\nf 1 2 5\n")
    n, sc = _load(path)
    assert n == 1
    pos, nrm, uv, tan = _tris(sc)
    assert pos.shape[0] == 3
    # fan: (1, 4, 3), (1, 3, 2), then (1, 2, 5)
    np.testing.assert_array_equal(pos[0], [[0, 0, 0], [0, 0, 1], [1, 0, 1]])
    np.testing.assert_array_equal(pos[1], [[0, 0, 0], [1, 0, 1], [1, 0, 0]])
    # computed normals: normalized sums of the unit face normals at each position
    up = np.array([0, 1, 0], np.float64)
    side = np.array([0, 0, 1], np.float64)  # (1,0,0)x(0,1,0) for face (1, 2, 5)
    expect = {(0, 0, 0): up + up + side, (1, 0, 0): up + side, (0, 0, 1): up + up, (1, 0, 1): up + up,
              (0, 1, 0): side}
    for i in range(3):
        for k in range(3):
            e = expect[tuple(int(x) for x in pos[i, k])]
            np.testing.assert_allclose(nrm[i, k], e / np.linalg.norm(e), atol=1e-7)
    # no texture coordinates anywhere: (0, 0) flipped to (0, 1); the tangent sums are zero -> NaN
    assert np.all(uv[..., 0] == 0.0) and np.all(uv[..., 1] == 1.0)
    assert np.all(np.isnan(tan))
    # the default surface: Lambertian with a GRAY constant albedo
    so = int(sc.primitives[2])
    assert int(sc.surfaces[so]) == 3
    off = int(sc.surfaces[so + 1])
    assert int(sc.textures[off]) == 4 and int(np.float32(sc.textures[off + 2])) & 0xFFFFFF == 0x808080


def test_tangent_formula_and_uv_flip(tmp_path):
    path = _write(tmp_path, "t.obj", "v 0 0 0\nv 2 1 0\nv 0 1 3\nvt 0 0\nvt 1 0.25\nvt 0.5 1\n"
                                     "f 1/1 2/2 3/3\n")
    _, sc = _load(path)
    pos, _, uv, tan = _tris(sc)
    e0 = np.array([2, 1, 0], np.float32)
    e1 = np.array([0, 1, 3], np.float32)
    # v flipped on load, but the tangent is computed from the file's coordinates
    du0, dv0, du1, dv1 = np.float32(1), np.float32(0.25), np.float32(0.5), np.float32(1)
    fr = np.float32(1) / np.float32(du0 * dv1 - du1 * dv0)
    t = np.array([fr * (dv1 * e0[0] - dv0 * e1[0]), fr * (dv1 * e0[1] - dv0 * e1[1]),
                  fr * (dv1 * e0[1] - dv0 * e1[1])], np.float64)
    for k in range(3):
        np.testing.assert_allclose(tan[0, k], t / np.linalg.norm(t), atol=1e-7)
    np.testing.assert_array_equal(uv[0], [[0, 1], [1, 0.75], [0.5, 0]])


def test_groups_split_meshes_and_name_the_mapper_call(tmp_path):
    # A triangle's group and material are those recorded for its first vertex when that vertex was first
    # met; a mesh ends where that group changes and takes the material of its last triangle.
    text = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nv 0 0 1\n"
            "g a\nusemtl red\nf 1 2 3\nusemtl blue\nf 2 4 3\n"
            "g b\nf 5 3 1\n"
            "g a\nusemtl green\nf 1 2 5\n")
    path = _write(tmp_path, "g.obj", text)
    calls = []
    with dfa.SceneBuilder("g") as b:
        black = b.constant((0, 0, 0))
        surf = {m: b.surface(3, b.constant(c), black, black)
                for m, c in (("red", (1, 0, 0)), ("blue", (0, 0, 1)), ("green", (0, 1, 0)))}

        def mapper(group, material):
            calls.append((group, material))
            return surf[material]

        n = b.load_obj(path, mapper)
        sc = b.compile()
    # meshes [f1, f2] (a; f2 starts at vertex 2, first met under usemtl red), [f3] (b; vertex 5, first met
    # under blue), [f4] (a; vertex 1, red): usemtl green never reaches a mesh
    assert n == 3 and calls == [("a", "red"), ("b", "blue"), ("a", "red")]
    assert sc.primitive_count == 3
    assert [int(x) for x in sc.primitives.reshape(-1, 3)[:, 2]] == [surf["red"] * 6, surf["blue"] * 6,
                                                                   surf["red"] * 6]
    # a face in a new group that starts at a vertex of the old group does not split
    text2 = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\ng a\nf 1 2 3\ng b\nf 1 3 4\nf 4 2 3\n"
    n2, sc2 = _load(_write(tmp_path, "g2.obj", text2))
    assert n2 == 2 and sc2.triangles.size // 12 == 3


def test_dict_surface_mapping_falls_back_to_its_default(tmp_path):
    path = _write(tmp_path, "d.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nv 1 0 1\nv 0 1 1\n"
                                     "g x\nusemtl glass\nf 1 2 3\ng y\nusemtl wood\nf 4 5 6\n")
    with dfa.SceneBuilder("d") as b:
        black = b.constant((0, 0, 0))
        glass = b.surface(2, black, black, black)
        other = b.surface(3, black, black, black)
        assert b.load_obj(path, {"glass": glass, None: other}) == 2
        sc = b.compile()
    so = sc.primitives.reshape(-1, 3)[:, 2]
    assert [int(sc.surfaces[s]) for s in so] == [2, 3]


def test_polygon_fans_crlf_and_ignored_lines(tmp_path):
    text = ("# comment\r\nv 0 0 0\r\nv  1 0 0\r\nv 1 1 0\r\nv 0.5 2 0\r\nv 0 1 0\r\n"
            "o name\r\ns off\r\nvp 1 2\r\n\r\nf 1 2 3 4 5\r\n")
    n, sc = _load(_write(tmp_path, "p.obj", text))
    pos = _tris(sc)[0]
    assert n == 1 and pos.shape[0] == 3
    np.testing.assert_array_equal(pos[2], [[0, 0, 0], [0.5, 2, 0], [0, 1, 0]])


def test_shared_positions_share_normals_across_uv_seams(tmp_path):
    # position 1 appears with two texture coordinates: two model vertices, one normal / tangent sum
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nvt 0 0\nvt 1 0\nvt 0 1\nvt 1 1\nf 1/1 2/2 3/3\nf 1/4 3/3 4/2\n"
    _, sc = _load(_write(tmp_path, "s.obj", text))
    _, nrm, uv, tan = _tris(sc)
    np.testing.assert_array_equal(nrm[0, 0], nrm[1, 0])
    np.testing.assert_array_equal(tan[0, 0], tan[1, 0])
    assert not np.array_equal(uv[0, 0], uv[1, 0])


def test_texture_coordinate_flag_is_file_wide(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.25 0.5\nvt 1 1\nf 1 2 3\nf 1/2 2/2 3/2\n"
    _, sc = _load(_write(tmp_path, "f.obj", text))
    uv = _tris(sc)[2]
    np.testing.assert_array_equal(uv[0], [[0.25, 0.5]] * 3)  # no index given: entry 0, flipped
    np.testing.assert_array_equal(uv[1], [[1.0, 0.0]] * 3)


def test_named_models_are_scaled_by_100(tmp_path):
    text = "v 0.5 0.25 -1\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0 0 1\nvn 0 0 1\nf 1//1 2//2 3//3\n"
    _, plain = _load(_write(tmp_path, "plain.obj", text))
    _, big = _load(_write(tmp_path, "aphroditegirl.obj", text))
    p0, n0 = _tris(plain)[:2]
    p1, n1 = _tris(big)[:2]
    np.testing.assert_array_equal(p1, p0 * np.float32(100))
    # normals by the transpose of the inverse scaling (1 / 100 up to the float inverse), not renormalized
    np.testing.assert_allclose(n1, n0 * 0.01, rtol=1e-6)


@pytest.mark.parametrize("text,why", [
    ("v 0 0 0\nv 1 0 0\nf 1 2 3\n", "missing vertex"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\ng\nf 1 2 3\n", "g without a name"),
    ("v 0 0 0\n", "no faces"),
    ("v 0 0 zero\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "bad number"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1 2/1 3/1\n", "missing texture coordinate"),
])
def test_malformed_files_are_refused(tmp_path, text, why):
    with dfa.SceneBuilder("bad") as b:
        with pytest.raises(RuntimeError):
            b.load_obj(_write(tmp_path, "bad.obj", text))
    with dfa.SceneBuilder("bad") as b:
        with pytest.raises(RuntimeError):
            b.load_obj(str(tmp_path / "does_not_exist.obj"))


def _rodrigues(axis, angle):
    k = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def test_transforms_match_rotation_scaling_translation():
    r = np.random.default_rng(4)
    with dfa.SceneBuilder("tr") as b:
        black = b.constant((0, 0, 0))
        s = b.surface(3, black, black, black)
        cases = []
        for _ in range(6):
            axis = r.normal(size=3)
            angle = r.uniform(-np.pi, np.pi)
            q = list(np.sin(angle / 2) * axis / np.linalg.norm(axis)) + [np.cos(angle / 2)]
            pos = r.uniform(-10, 10, 3)
            scale = r.uniform(0.5, 3.0, 3)
            b.sphere((0, 0, 0), 1.0, s, assets.transform(pos, q, scale))
            cases.append((axis, angle, pos, scale))
        b.sphere((0, 0, 0), 1.0, s, assets.transform((1.5, -2.0, 3.25)))
        sc = b.compile()
    o2w = sc.object_to_world.reshape(-1, 4, 4).astype(np.float64)
    w2o = sc.world_to_object.reshape(-1, 4, 4).astype(np.float64)
    for i, (axis, angle, pos, scale) in enumerate(cases):
        m = np.eye(4)
        m[:3, :3] = _rodrigues(axis, angle) @ np.diag(scale)
        m[:3, 3] = pos
        np.testing.assert_allclose(o2w[i], m, atol=2e-5 * max(1.0, np.abs(m).max()))
        np.testing.assert_allclose(o2w[i] @ w2o[i], np.eye(4), atol=1e-5)
    # translation only: exact
    np.testing.assert_array_equal(sc.object_to_world.reshape(-1, 4, 4)[-1], [[1, 0, 0, 1.5], [0, 1, 0, -2.0],
                                                                          [0, 0, 1, 3.25], [0, 0, 0, 1]])
    np.testing.assert_array_equal(sc.world_to_object.reshape(-1, 4, 4)[-1][:3, 3], [-1.5, 2.0, -3.25])


def test_singular_transform_is_refused():
    with dfa.SceneBuilder("s") as b:
        black = b.constant((0, 0, 0))
        s = b.surface(3, black, black, black)
        with pytest.raises(RuntimeError):
            b.sphere((0, 0, 0), 1.0, s, assets.transform(scale=(1.0, 0.0, 1.0)))
        with pytest.raises(RuntimeError):
            b.sphere((0, 0, 0), 1.0, s + 7)


def _pack(rgb):
    rgb = rgb.astype(np.int64)
    return ((rgb[..., 0] << 16) | (rgb[..., 1] << 8) | rgb[..., 2]).astype(np.int32).reshape(-1)


def test_png_images_decode_to_the_int_rgb_raster(tmp_path):
    from PIL import Image

    r = np.random.default_rng(2)
    rgb = r.integers(0, 256, (7, 5, 3), dtype=np.uint8)
    Image.fromarray(rgb, "RGB").save(tmp_path / "rgb.png")
    w, h, px = dfa.image_rgb(str(tmp_path / "rgb.png"))
    assert (w, h) == (5, 7)
    np.testing.assert_array_equal(px, _pack(rgb))
    # alpha: opaque copied, transparent black, partial round(a * c / 255)
    rgba = r.integers(0, 256, (6, 4, 4), dtype=np.uint8)
    rgba[0, :, 3] = 255
    rgba[1, :, 3] = 0
    Image.fromarray(rgba, "RGBA").save(tmp_path / "rgba.png")
    _, _, px = dfa.image_rgb(str(tmp_path / "rgba.png"))
    a = rgba[..., 3:4].astype(np.int64)
    expect = np.floor(rgba[..., :3].astype(np.int64) * a / 255.0 + 0.5).astype(np.int64)
    np.testing.assert_array_equal(px, _pack(expect))
    np.testing.assert_array_equal(px.reshape(6, 4)[0], _pack(rgba[0, :, :3]))
    assert np.all(px.reshape(6, 4)[1] == 0)
    # grey replicated; palette looked up
    g = r.integers(0, 256, (3, 3), dtype=np.uint8)
    Image.fromarray(g, "L").save(tmp_path / "g.png")
    np.testing.assert_array_equal(dfa.image_rgb(str(tmp_path / "g.png"))[2], _pack(np.stack([g] * 3, -1)))
    pal = Image.fromarray(rgb, "RGB").quantize(colors=8)
    pal.save(tmp_path / "p.png")
    np.testing.assert_array_equal(dfa.image_rgb(str(tmp_path / "p.png"))[2], _pack(np.asarray(pal.convert("RGB"))))
    # 16-bit grey is refused
    Image.fromarray(g.astype(np.uint16) * 257).save(tmp_path / "g16.png")
    with pytest.raises(ValueError):
        dfa.image_rgb(str(tmp_path / "g16.png"))


def test_jpeg_images_decode_close_to_the_source(tmp_path):
    from PIL import Image

    yy, xx = np.mgrid[0:32, 0:48]
    rgb = np.stack([xx * 5, yy * 7, (xx + yy) * 3], -1).astype(np.uint8)
    Image.fromarray(rgb, "RGB").save(tmp_path / "a.jpg", quality=95)
    w, h, px = dfa.image_rgb(str(tmp_path / "a.jpg"))
    assert (w, h) == (48, 32)
    got = np.stack([(px >> 16) & 255, (px >> 8) & 255, px & 255], -1).reshape(32, 48, 3)
    assert np.abs(got.astype(int) - rgb.astype(int)).mean() < 3.0


def test_image_texture_record_layout():
    data = np.arange(6, dtype=np.int32) * 0x010203
    with dfa.SceneBuilder("img") as b:
        t = b.image((3, 2, data), 90.0, 2.0, 0.5)
        black = b.constant((0, 0, 0))
        b.sphere((0, 0, 0), 1.0, b.surface(3, t, black, black))
        sc = b.compile()
    off = int(sc.surfaces[1])
    rec = sc.textures[off:off + 8 + 6]
    assert rec[0] == 6 and rec[1] == 14  # ImageTexture TYPE, size 8 + W*H (ImageTexture.java:337-356)
    assert rec[2] == np.float32(np.cos(np.float64(np.float32(np.pi / 2))))
    assert rec[3] == np.float32(np.sin(np.float64(np.float32(np.pi / 2))))
    np.testing.assert_array_equal(rec[4:8], [3, 2, 2.0, 0.5])
    np.testing.assert_array_equal(rec[8:].astype(np.int64), data)


def _synthetic_textures():
    r = np.random.default_rng(11)
    t2 = (64, 48, _pack(r.integers(0, 256, (48, 64, 3))))
    yy, xx = np.mgrid[0:40, 0:40]
    normal = np.stack([128 + 60 * np.sin(xx / 3.0), 128 + 60 * np.cos(yy / 4.0), np.full_like(xx, 230.0)], -1)
    return t2, (40, 40, _pack(np.clip(normal, 0, 255)))


def test_material_showcase_scene_compiles():
    t2, nrm = _synthetic_textures()
    sc = assets.material_showcase_scene(t2, nrm, width=160, height=90)
    prims = sc.primitives.reshape(-1, 3)
    assert len(prims) == 16 and prims[0, 0] == 1 and np.all(prims[1:, 0] == 2)
    assert (sc.width, sc.height) == (160, 90)
    tex_types = []
    i = 0
    while i < sc.textures.size:
        tex_types.append(int(sc.textures[i]))
        i += int(sc.textures[i + 1])
    assert tex_types.count(6) == 2
    # eye (295, 42, 332.6) with the default orientation
    np.testing.assert_array_equal(sc.camera[1:4], np.array([295.0, 42.0, 332.6], np.float32))


# ---------------------------------------------------------------- the reference's own asset files
# tests/golden/Texture_2.png, bricks2_normal.jpg, monkey0.obj and smoothMonkey0.obj are data files the reference
# ships (resources/distribution/resources/texture, .../model); Scenes.java:356,376 load the two images.

def _png_rgba(path):
    """A minimal independent PNG decoder (8-bit, non-interlaced; RGB, RGBA or palette with tRNS): the filtered
    scanlines of the concatenated IDAT stream unfiltered per the PNG specification, section 9."""
    import struct
    import zlib

    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    i, chunks, idat = 8, {}, b""
    while i < len(data):
        n, = struct.unpack(">I", data[i:i + 4])
        kind, body = data[i + 4:i + 8], data[i + 8:i + 8 + n]
        if kind == b"IDAT":
            idat += body
        else:
            chunks[kind] = body
        i += 12 + n
    w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert depth == 8 and interlace == 0 and ctype in (2, 3, 6)
    bpp = {2: 3, 3: 1, 6: 4}[ctype]
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * bpp)
    out = np.zeros((h, w * bpp), np.int64)
    for y in range(h):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int64)
        prev = out[y - 1] if y else np.zeros(w * bpp, np.int64)
        cur = np.zeros(w * bpp, np.int64)
        for x in range(w * bpp):
            a = cur[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            if f == 0:
                pred = 0
            elif f == 1:
                pred = a
            elif f == 2:
                pred = b
            elif f == 3:
                pred = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if pa <= pb and pa <= pc else b if pb <= pc else c
            cur[x] = (line[x] + pred) & 0xFF
        out[y] = cur
    if ctype == 3:
        pal = np.frombuffer(chunks[b"PLTE"], np.uint8).reshape(-1, 3).astype(np.int64)
        alpha = np.full(len(pal), 255, np.int64)
        trns = np.frombuffer(chunks.get(b"tRNS", b""), np.uint8)
        alpha[:len(trns)] = trns
        return w, h, pal[out], alpha[out]
    px = out.reshape(h, w, bpp)
    return w, h, px[:, :, :3], px[:, :, 3] if bpp == 4 else np.full((h, w), 255, np.int64)


def test_reference_png_texture_decodes_exactly():
    """Texture_2.png (an 8-bit palette image whose tRNS chunk makes palette entry 0 transparent; no pixel uses
    that entry) as ImageTexture.load keeps it: the TYPE_INT_RGB raster of the image drawn over black, i.e.
    opaque texels as stored, a texel of alpha a scaled by a / 255 rounded to nearest."""
    path = os.path.join(GOLDEN, "Texture_2.png")
    w, h, rgb, alpha = _png_rgba(path)
    assert (w, h) == (512, 512)
    over_black = np.where(alpha[..., None] == 255, rgb, (alpha[..., None] * rgb + 127) // 255)
    want = (over_black[..., 0] << 16) | (over_black[..., 1] << 8) | over_black[..., 2]
    gw, gh, px = assets.image_rgb(path)
    assert (gw, gh) == (w, h)
    np.testing.assert_array_equal(px, want.reshape(-1).astype(np.int32))
    assert np.all(alpha == 255) and len(np.unique(px)) > 8


def test_reference_jpeg_normal_map_decodes():
    """bricks2_normal.jpg: 512 x 512 baseline JPEG; a tangent-space normal map, so its blue channel dominates.
    (The decode's agreement with ImageIO's is unpinned: no JVM here.)"""
    w, h, px = assets.image_rgb(os.path.join(GOLDEN, "bricks2_normal.jpg"))
    assert (w, h) == (512, 512)
    r, g, b = (px >> 16) & 0xFF, (px >> 8) & 0xFF, px & 0xFF
    assert b.mean() > 200 and abs(r.mean() - 128) < 20 and abs(g.mean() - 128) < 20


@pytest.mark.parametrize("name", ["monkey0.obj", "smoothMonkey0.obj"])
def test_reference_obj_models_load(name):
    """Blender's Suzanne as the reference ships it, flat (monkey0) and smooth (smoothMonkey0): every face fanned
    from its first vertex in file order, positions and given normals kept as written, v flipped to 1 - v."""
    path = os.path.join(GOLDEN, name)
    n, sc = _load(path)
    v, vt, vn, faces = _obj_tokens(path)
    fan = [(f[0], f[k], f[k + 1]) for f in faces for k in range(1, len(f) - 1)]
    assert n == 1 and sc.primitive_count == 1
    pos, nrm, uv, _ = _tris(sc)
    assert pos.shape[0] == len(fan) == sum(len(f) - 2 for f in faces)
    for i, tri in enumerate(fan):
        for k, (vi, ti, ni) in enumerate(tri):
            np.testing.assert_array_equal(pos[i, k], v[vi])
            np.testing.assert_array_equal(nrm[i, k], vn[ni])
            assert uv[i, k, 1] == np.float32(np.float32(1.0) - vt[ti, 1])
    stats = dfa.scene.stats(sc)
    assert stats["triangles"] == len(fan)
