"""Scene files on the host side: the reference's PBRT subset
(crates/pyre/src/io/pbrt/{lexer,parser,builder}.rs), glTF meshes
(io/gltf.rs) and Radiance .hdr environments (io/hdr.rs), read through
pyre_preset_load_pbrt / pyre_preset_load_gltf / pyre_load_hdr.

What pins them:
  * the lexer/parser behaviours the reference's own tests assert
    (lexer.rs:184-207, parser.rs:389-441), through whole-file loads;
  * the Cornell round trip the reference ships scenes/cornell.pbrt for:
    the same scene written in PBRT loads to the built-in preset's vertices,
    indices, materials and light bit for bit, and to its normals as numbers
    (the sign of a zero component aside); the sphere normals are the
    preset's normalised once more (builder.rs:549 normalises
    transform_vector3(n)), which is the only difference;
  * transforms against an independent float32 restatement of the same glam
    products (translations, scales, matrices: exact) and a closed form
    (rotations: within 1e-6);
  * Radiance pixels against the format's definition, channel * 2^(E - 136),
    for flat, old run-length and new run-length scanlines written here;
  * glTF node transforms (T * R * S, parent * child) and inverse-transpose
    normals against float64 closed forms.
No GPU: these are host-only entry points.
"""
import base64
import json
import os
import struct

import numpy as np
import pytest

import pyre_amd
from pyre_amd import PyreError

F = np.float32


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _load(tmp_path, text, name="s.pbrt"):
    return pyre_amd.Preset.from_pbrt(_write(tmp_path, name, text))


def _quad(pts, n, indices="0 1 2 0 2 3"):
    P = " ".join(f"{v:g}" for p in pts for v in p)
    s = f'Shape "trianglemesh" "point P" [{P}] "integer indices" [{indices}]'
    if n is not None:
        s = s.replace('"integer', '"normal N" [' + " ".join(f"{v:g}" for _ in pts for v in n) + '] "integer')
    return s


def _cornell_pbrt():
    """The built-in Cornell box (main.rs:365-496) written in the PBRT subset."""
    walls = [
        ("0.73 0.73 0.73", [(-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1)], (0, 1, 0)),
        ("0.73 0.73 0.73", [(-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)], (0, -1, 0)),
        ("0.73 0.73 0.73", [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)], (0, 0, 1)),
        ("0.65 0.05 0.05", [(-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1)], (1, 0, 0)),
        ("0.12 0.45 0.15", [(1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)], (-1, 0, 0)),
    ]
    out = ["LookAt 0 0 3  0 0 0  0 1 0", 'Camera "perspective" "float fov" [45]',
           'Film "image" "integer xresolution" [512] "integer yresolution" [512]',
           'Sampler "halton" "integer pixelsamples" [64]', 'Integrator "path" "integer maxdepth" [8]', "WorldBegin"]
    for kd, pts, n in walls:
        out += ["AttributeBegin", f'  Material "matte" "rgb Kd" [{kd}]', "  " + _quad(pts, n), "AttributeEnd"]
    out += ["AttributeBegin", '  Material "metal" "rgb Kr" [1.0 0.766 0.336] "float roughness" [0.1]',
            "  Translate 0.35 -0.65 0.25", '  Shape "sphere" "float radius" [0.35]', "AttributeEnd"]
    out += ["AttributeBegin", '  Material "plastic" "rgb Kd" [0.72 0.72 0.72] "rgb Ks" [0.5 0.5 0.5]',
            '    "float roughness" [0.18]', "  Translate -0.35 -0.65 -0.3", '  Shape "sphere" "float radius" [0.35]',
            "AttributeEnd"]
    out += ["AttributeBegin", '  AreaLightSource "diffuse" "rgb L" [15 15 15]',
            "  " + _quad([(-0.3, 0.999, -0.3), (0.3, 0.999, -0.3), (0.3, 0.999, 0.3), (-0.3, 0.999, 0.3)], None),
            "AttributeEnd", "WorldEnd"]
    return "\n".join(out) + "\n"


def _normalize_f32(n):
    """Vec3::normalize in f32: v * (1 / sqrt((x*x + y*y) + z*z))."""
    n = n.astype(F)
    d = (n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2]
    r = F(1.0) / np.sqrt(d)
    return n * r[:, None]


def test_cornell_round_trip_matches_builtin_preset(tmp_path):
    got = _load(tmp_path, _cornell_pbrt(), "cornell.pbrt")
    ref = pyre_amd.Preset.cornell()
    a, b = got.arrays(), ref.arrays()
    assert len(a["meshes"]) == len(b["meshes"]) == 7
    for k, (ma, mb) in enumerate(zip(a["meshes"], b["meshes"])):
        assert np.array_equal(ma["positions"].view(np.uint32), mb["positions"].view(np.uint32)), k
        assert np.array_equal(ma["indices"], mb["indices"]), k
        # equal as numbers: the file's "0" is +0 where the preset negates to
        # -0, and transform_vector3's "+ 0 * y" turns -0 into +0
        want = mb["normals"] if k < 5 else _normalize_f32(mb["normals"])
        assert np.array_equal(ma["normals"], want), k
        assert a["materials"][ma["material_id"]] == b["materials"][mb["material_id"]], k
    # one material per Material directive (builder.rs:201-207)
    assert len(a["materials"]) == 7
    assert len(a["lights"]) == 1 and np.array_equal(a["lights"][0], b["lights"][0])
    assert a["env"] is None
    v = got.view
    assert v["camera"] == dict(origin=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), vfov_deg=45.0)
    assert v["film"] == (512, 512)
    assert got.camera(1.0) == ref.camera(1.0)


def test_lexer_and_parser_cases_of_the_reference_tests(tmp_path):
    # lexer.rs:185-190 / parser.rs:390-400: LookAt with 9 numbers, ".5" forms
    # parser.rs:403-418: "float fov" [39] with a bare integer
    p = _load(tmp_path, "# only a comment\nLookAt 3 4 1.5 .5 .5 0 0 0 1\nCamera \"perspective\" \"float fov\" [39]\n")
    assert p.view["camera"] == dict(origin=(3.0, 4.0, 1.5), target=(0.5, 0.5, 0.0), up=(0.0, 0.0, 1.0),
                                    vfov_deg=39.0)
    assert p.view["film"] is None
    # the bare-value form of a parameter
    p = _load(tmp_path, 'LookAt 0 0 3 0 0 0 0 1 0 Camera "perspective" "float fov" 60')
    assert p.view["camera"]["vfov_deg"] == 60.0
    # parser.rs:434-441: [0.5 0 1] promotes the integers to floats
    p = _load(tmp_path, 'WorldBegin Material "matte" "rgb Kd" [0.5 0 1]\n' + _quad([(0, 0, 0), (1, 0, 0), (1, 1, 0),
                                                                                    (0, 1, 0)], None))
    assert p.arrays()["materials"][0].base_color == (0.5, 0.0, 1.0)
    # a camera needs both LookAt and a fov; otherwise the scene is framed
    p = _load(tmp_path, 'Camera "perspective"\nWorldBegin\n' + _quad([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)],
                                                                     None))
    assert p.view["camera"] is None
    o, t = pyre_amd.auto_frame((0, 0, 0), (1, 1, 0), 45.0, 1.5)
    c = p.camera(1.5)
    assert c.origin == tuple(o) and c.target == tuple(t)
    # Film defaults (builder.rs:118-131)
    assert _load(tmp_path, 'Film "image" "string filename" ["x.png"]').view["film"] == (800, 600)


def test_materials_lights_and_defaults(tmp_path):
    text = "\n".join([
        "WorldBegin",
        'MakeNamedMaterial "gold" "string type" ["conductor"] "float roughness" [0.3]',
        'MakeNamedMaterial "grey"',
        'NamedMaterial "gold"',
        _quad([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], None),
        'Material "plastic" "rgb Kd" [0.2 0.3 0.4] "rgb Ks" [0.1 0.2 0.6]',
        _quad([(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], None),
        'Material "glass"',
        _quad([(0, 0, 2), (1, 0, 2), (1, 1, 2), (0, 1, 2)], None),
        'NamedMaterial "grey"',
        'AttributeBegin AreaLightSource "diffuse" "rgb L" [1 2 3]',
        '  Shape "trianglemesh" "point P" [0 0 3 1 0 3 0 1 3] "integer indices" [0 1 2]',  # not a quad: a primitive
        "AttributeEnd",
        'LightSource "infinite" "rgb L" [0.5 0.25 2] "float scale" [3]',
        'LightSource "point"',
        'Shape "cylinder"',
    ])
    a = _load(tmp_path, text).arrays()
    mats = a["materials"]
    gold, grey = mats[0], mats[1]
    assert (gold.kind, gold.base_color, gold.metallic, gold.roughness, gold.specular) == (
        pyre_amd.MAT_DISNEY, (1.0, F(0.766), F(0.336)), 1.0, F(0.3), 0.5)
    assert (grey.kind, grey.base_color) == (pyre_amd.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))
    plastic = mats[a["meshes"][1]["material_id"]]
    assert plastic.kind == pyre_amd.MAT_DISNEY and plastic.metallic == 0.0 and plastic.roughness == F(0.2)
    assert plastic.specular == (F(0.1) + F(0.2) + F(0.6)) / F(3)
    assert mats[a["meshes"][2]["material_id"]].base_color == (0.5, 0.5, 0.5)  # unsupported -> grey
    assert [m["material_id"] for m in a["meshes"]] == [0, 2, 3, 1]
    assert len(a["meshes"]) == 4 and not a["lights"]
    env = a["env"]
    assert (env["width"], env["height"], env["intensity"]) == (1, 1, 3.0)
    assert env["rgb"].reshape(-1).tolist() == [0.5, 0.25, 2.0]
    # a shape without a material gets a new grey Lambertian each time
    b = _load(tmp_path, "WorldBegin\n" + _quad([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], None) * 1 + "\n" +
              _quad([(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], None)).arrays()
    assert [m["material_id"] for m in b["meshes"]] == [0, 1] and len(b["materials"]) == 2
    # shapes before WorldBegin are ignored
    assert not _load(tmp_path, _quad([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], None)).arrays()["meshes"]


def _mat_f32(cols):
    return np.array(cols, F).reshape(4, 4)  # [column][row]


def _mul_f32(a, b):
    """glam Mat4 * Mat4 in f32: each result column is ((x*v0 + y*v1) + z*v2) + w*v3."""
    out = np.zeros((4, 4), F)
    for j in range(4):
        v = b[j]
        out[j] = ((a[0] * v[0] + a[1] * v[1]) + a[2] * v[2]) + a[3] * v[3]
    return out


def _point_f32(m, p):
    return ((m[0][:3] * F(p[0]) + m[1][:3] * F(p[1])) + m[2][:3] * F(p[2])) + m[3][:3]


def test_transforms_against_glam_restatement(tmp_path):
    T = lambda x, y, z: _mat_f32([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, x, y, z, 1])
    S = lambda x, y, z: _mat_f32([x, 0, 0, 0, 0, y, 0, 0, 0, 0, z, 0, 0, 0, 0, 1])
    M = [0.5, 0.25, 0, 0, -0.75, 1.5, 0.125, 0, 0.3, 0, 2.0, 0, 1.1, -2.2, 3.3, 1]
    tri = [(0.1, 0.2, 0.3), (1.7, -0.4, 0.9), (-0.6, 1.3, 2.5)]
    P = " ".join(f"{v!r}" for p in tri for v in p)
    shape = f'Shape "trianglemesh" "point P" [{P}] "integer indices" [0 1 2]'
    text = "\n".join([
        "Translate 9 9 9",  # WorldBegin resets the CTM
        "WorldBegin",
        "AttributeBegin", "Translate 0.5 -1.25 2", "Scale 2 0.5 3", shape, "AttributeEnd",
        "TransformBegin", "ConcatTransform [" + " ".join(map(repr, M)) + "]", "Translate 0.1 0.2 0.3", shape,
        "TransformEnd",
        shape,  # identity again
        "Scale 3 3 3", "Transform [" + " ".join(map(repr, M)) + "]", shape,
        "Identity", "Rotate 90 0 0 2", shape,
    ])
    a = _load(tmp_path, text).arrays()["meshes"]
    I = _mat_f32([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])
    want = [
        _mul_f32(_mul_f32(I, T(0.5, -1.25, 2)), S(2, 0.5, 3)),
        _mul_f32(_mul_f32(I, _mat_f32(M)), T(0.1, 0.2, 0.3)),
        I,
        _mat_f32(M),
    ]
    for k, m in enumerate(want):
        exp = np.stack([_point_f32(m, p) for p in tri])
        assert np.array_equal(a[k]["positions"].view(np.uint32), exp.view(np.uint32)), k
    # Rotate 90 degrees about +z (axis normalised): (x, y, z) -> (-y, x, z)
    exp = np.array([(-p[1], p[0], p[2]) for p in tri])
    assert np.abs(a[4]["positions"] - exp).max() < 1e-6


@pytest.mark.parametrize("text,msg", [
    ("AttributeEnd", "build: AttributeEnd without matching AttributeBegin (line 1)"),
    ("\nTransformEnd", "build: TransformEnd without matching TransformBegin (line 2)"),
    ('NamedMaterial "nope"', 'build: named material "nope" not declared'),
    ('WorldBegin Shape "trianglemesh" "integer indices" [0 1 2]',
     'build: Shape "trianglemesh" requires parameter "P"'),
    ('WorldBegin Shape "trianglemesh" "point P" [0 0 0 1 0] "integer indices" [0 1 2]',
     "build: malformed trianglemesh on line 1: P length not a multiple of 3"),
    ('WorldBegin Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 5]',
     "build: malformed trianglemesh on line 1: index 5 out of range"),
    ("LookAt 1 2 3", "build: malformed LookAt on line 1: expected 9 float positionals, got 3"),
    ("Transform [1 2 3]", "build: malformed Transform on line 1: expected a 16-float array"),
    ('Camera "perspective" "float fov [45]', "lex: line 1, col 22: unterminated string literal"),
    ("LookAt 1 2 3 4 5 6 7 8 9 {", "lex: line 1, col 26: unexpected character '{'"),
    ("Translate 1-2 0 0", 'lex: line 1, col 11: invalid number "1-2"'),
    ('Camera "perspective" "fov" [45]', ""),  # no space: a positional string, not an error
    ('Camera "perspective" "float  " [45]', 'parse: line 1, col 22: parameter header "float  " must have form'),
    ('Film "image" "integer xresolution" [1 "a"]', "parse: line 1, col 39: mixed types in parameter array"),
    ('Film "image" "integer xresolution" [1', "parse: line 1, col 37: unexpected end of input"),
    ("] LookAt", "parse: line 1, col 1: expected directive keyword, found `]`"),
    ('LightSource "infinite" "string mapname" ["env.exr"]',
     "build: HDRI load failed: unsupported environment-map extension: exr"),
])
def test_errors_carry_the_reference_messages(tmp_path, text, msg):
    path = _write(tmp_path, "e.pbrt", text)
    if not msg:
        pyre_amd.Preset.from_pbrt(path)
        return
    with pytest.raises(PyreError) as e:
        pyre_amd.Preset.from_pbrt(path)
    assert e.value.code == -1 and msg in str(e.value)


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(PyreError) as e:
        pyre_amd.Preset.from_pbrt(str(tmp_path / "none.pbrt"))
    assert e.value.code == -6 and "io:" in str(e.value)


# ---------------------------------------------------------------- Radiance


def _rgbe_header(w, h, fmt="32-bit_rle_rgbe", dims=None):
    head = f"#?RADIANCE\n# written by the test\nFORMAT={fmt}\nEXPOSURE=1.0\n\n"
    return (head + (dims or f"-Y {h} +X {w}") + "\n").encode()


def _rle_plane(vals):
    """New-style run-length coding of one component plane."""
    out, i = bytearray(), 0
    while i < len(vals):
        j = i
        while j < len(vals) and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), vals[i]])
            i = j
        else:
            k = i
            while k < len(vals) and k - i < 128 and not (k + 2 < len(vals) and vals[k] == vals[k + 1] == vals[k + 2]):
                k += 1
            out += bytes([k - i]) + bytes(vals[i:k])
            i = k
    return bytes(out)


def _expected(px):
    px = px.astype(np.int64)
    s = np.where(px[..., 3] == 0, 0.0, np.ldexp(1.0, (px[..., 3] - 136).clip(-200, 200)))
    return (px[..., :3] * s[..., None]).astype(F)


def test_hdr_flat_and_run_length_scanlines(tmp_path):
    rng = np.random.default_rng(7)
    w, h = 40, 5
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    px[..., 3] = rng.integers(100, 160, (h, w))
    px[0, :7, 3] = 0  # black regardless of the mantissas
    px[1, 3:30] = px[1, 3]  # a long run
    px[2, :, 3] = 1  # subnormal scales, still exact
    body = bytearray()
    for y in range(h):
        if y % 2 == 0:  # new RLE: 2 2 w_hi w_lo, then four planes
            body += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                body += _rle_plane(list(px[y, :, c]))
        else:  # flat, with an old-style run marker (1 1 1 n) after the first pixel
            row = px[y].copy()
            row[1:4] = row[0]
            body += bytes(row[0]) + bytes([1, 1, 1, 3]) + b"".join(bytes(p) for p in row[4:])
            px[y] = row
    f = tmp_path / "t.hdr"
    f.write_bytes(_rgbe_header(w, h) + bytes(body))
    img = pyre_amd.load_hdr(str(f))
    assert img.shape == (h, w, 3)
    assert np.array_equal(img.view(np.uint32), _expected(px).view(np.uint32))


@pytest.mark.parametrize("data,msg", [
    (b"#?RGBE\n\n-Y 1 +X 1\n\x80\x80\x80\x80", "signature"),
    (None, "dimensions"),
    ("xyze", "pixel format"),
    ("short", "end of pixel data"),
])
def test_hdr_rejects_what_the_decoder_rejects(tmp_path, data, msg):
    if data is None:
        data = _rgbe_header(1, 1, dims="+Y 1 +X 1") + b"\x80\x80\x80\x80"
    elif data == "xyze":
        data = _rgbe_header(1, 1, fmt="32-bit_rle_xyze") + b"\x80\x80\x80\x80"
    elif data == "short":
        data = _rgbe_header(3, 1) + b"\x80\x80\x80\x80"
    f = tmp_path / "bad.hdr"
    f.write_bytes(data)
    with pytest.raises(PyreError) as e:
        pyre_amd.load_hdr(str(f))
    assert msg in str(e.value)


def test_pbrt_infinite_light_reads_the_map_next_to_the_file(tmp_path):
    px = np.array([[[128, 64, 32, 129], [255, 255, 255, 120]]], np.uint8)
    (tmp_path / "maps").mkdir()
    (tmp_path / "maps" / "sky.hdr").write_bytes(_rgbe_header(2, 1) + px.tobytes())
    p = _load(tmp_path, 'LightSource "infinite" "string mapname" ["maps/sky.hdr"] "float scale" [0.5]')
    env = p.arrays()["env"]
    assert (env["width"], env["height"], env["intensity"]) == (2, 1, 0.5)
    assert np.array_equal(env["rgb"], _expected(px))
    # the CLI's --env replaces the scene's environment (main.rs:162-166)
    q = pyre_amd.Preset.cornell()
    q.set_env_hdr(str(tmp_path / "maps" / "sky.hdr"), 2.0)
    assert q.arrays()["env"]["intensity"] == 2.0


# -------------------------------------------------------------------- glTF


def _gltf_doc(positions, normals, indices, nodes, scene_nodes, mode=None, with_positions=True):
    blob = positions.astype("<f4").tobytes() + normals.astype("<f4").tobytes() + indices.astype("<u2").tobytes()
    npos = len(positions)
    attrs = {"POSITION": 0, "NORMAL": 1} if with_positions else {"NORMAL": 1}
    prim = {"attributes": attrs, "indices": 2}
    if mode is not None:
        prim["mode"] = mode
    doc = {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": scene_nodes}],
        "nodes": nodes,
        "meshes": [{"primitives": [prim, {"attributes": {"POSITION": 0}}]}],
        "buffers": [{"byteLength": len(blob)}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 12 * npos},
                        {"buffer": 0, "byteOffset": 12 * npos, "byteLength": 12 * npos},
                        {"buffer": 0, "byteOffset": 24 * npos, "byteLength": 2 * len(indices)}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": npos, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": npos, "type": "VEC3"},
                      {"bufferView": 2, "componentType": 5123, "count": len(indices), "type": "SCALAR"}],
    }
    return doc, blob


def _quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _trs(t, q, s):
    m = np.eye(4)
    m[:3, :3] = _quat_mat(q) @ np.diag(s)
    m[:3, 3] = t
    return m


@pytest.mark.parametrize("container", ["gltf", "glb"])
def test_gltf_nodes_transforms_and_normals(tmp_path, container):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    nrm = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0], [1, 1, 1]], np.float64)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    c, s = np.cos(0.3), np.sin(0.3)
    q = (0.0, s, 0.0, c)  # 0.6 rad about +y
    m_child = [2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0.5, 0, 0.25, 0.5, 0.75, 1]
    nodes = [
        {"translation": [1, -2, 3], "rotation": list(q), "scale": [1, 3, 1], "children": [1]},
        {"matrix": m_child, "mesh": 0},
        {"mesh": 0},  # not in the scene
    ]
    doc, blob = _gltf_doc(pos, nrm, idx, nodes, [0])
    if container == "gltf":
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
        path = _write(tmp_path, "m.gltf", json.dumps(doc))
    else:
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        blob += b"\0" * (-len(blob) % 4)
        data = (struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(blob)) +
                struct.pack("<I4s", len(js), b"JSON") + js + struct.pack("<I4s", len(blob), b"BIN\0") + blob)
        path = str(tmp_path / "m.glb")
        open(path, "wb").write(data)
    p = pyre_amd.Preset.from_gltf(path)
    a = p.arrays()
    assert len(a["meshes"]) == 2 and len(a["materials"]) == 1
    assert a["materials"][0].kind == pyre_amd.MAT_LAMBERTIAN and a["materials"][0].base_color == (F(0.73),) * 3
    world = _trs((1, -2, 3), q, (1, 3, 1)) @ np.array(m_child, np.float64).reshape(4, 4).T
    want = (world[:3, :3] @ pos.T).T + world[:3, 3]
    m0, m1 = a["meshes"]
    assert np.abs(m0["positions"] - want).max() < 1e-5
    assert m0["indices"].reshape(-1).tolist() == idx.tolist()
    n = (np.linalg.inv(world[:3, :3]).T @ nrm.T).T
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.abs(m0["normals"] - n).max() < 1e-5
    # a primitive without indices lists its vertices in order (4 vertices:
    # one triangle, the partial one dropped); no normals
    assert m1["indices"].reshape(-1).tolist() == [0, 1, 2]
    assert m1["normals"] is None
    assert p.view == dict(camera=None, film=None)


def test_gltf_errors(tmp_path):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    doc, blob = _gltf_doc(pos, pos, np.array([0, 1, 2], np.uint16), [{"mesh": 0}], [0], mode=1)
    doc["buffers"][0]["uri"] = "b.bin"
    (tmp_path / "b.bin").write_bytes(blob)
    with pytest.raises(PyreError, match="non-triangle primitive"):
        pyre_amd.Preset.from_gltf(_write(tmp_path, "a.gltf", json.dumps(doc)))
    doc, blob = _gltf_doc(pos, pos, np.array([0, 1, 2], np.uint16), [{"mesh": 0}], [0], with_positions=False)
    doc["buffers"][0]["uri"] = "b.bin"
    with pytest.raises(PyreError, match="no positions"):
        pyre_amd.Preset.from_gltf(_write(tmp_path, "b.gltf", json.dumps(doc)))
    with pytest.raises(PyreError, match="invalid JSON"):
        pyre_amd.Preset.from_gltf(_write(tmp_path, "c.gltf", "{\"asset\": "))
    # a node with two parents (or a cycle) is not a tree
    doc, blob = _gltf_doc(pos, pos, np.array([0, 1, 2], np.uint16), [{"children": [1, 1]}, {"mesh": 0}], [0])
    doc["buffers"][0]["uri"] = "b.bin"
    with pytest.raises(PyreError, match="more than one parent"):
        pyre_amd.Preset.from_gltf(_write(tmp_path, "d.gltf", json.dumps(doc)))


def _crafted(tmp_path, name, edit):
    """A valid one-triangle glTF whose document `edit` then corrupts."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    doc, blob = _gltf_doc(pos, pos, np.array([0, 1, 2], np.uint16), [{"mesh": 0}], [0])
    doc["meshes"][0]["primitives"] = doc["meshes"][0]["primitives"][:1]
    doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
    edit(doc)
    return _write(tmp_path, name, json.dumps(doc))


@pytest.mark.parametrize("edit,msg", [
    # byteOffset near 2^64: the old sum view_off + view_len wrapped below the buffer size
    (lambda d: d["bufferViews"][0].update(byteOffset=2.0 ** 64 - 2048, byteLength=4096), "bufferView byteOffset"),
    (lambda d: d["bufferViews"][0].update(byteOffset=-8), "bufferView byteOffset"),
    (lambda d: d["bufferViews"][0].update(byteLength=1e300), "bufferView byteLength"),
    (lambda d: d["bufferViews"][0].update(byteLength=3.5), "bufferView byteLength"),
    (lambda d: d["bufferViews"][1].update(byteOffset=50, byteLength=40), "bufferView exceeds its buffer"),
    # count * stride chosen to wrap: 2^62 elements of 12 bytes
    (lambda d: d["accessors"][0].update(count=2.0 ** 62), "accessor count"),
    (lambda d: d["accessors"][0].update(count=2 ** 30 + 1), "accessor exceeds its bufferView"),
    (lambda d: d["accessors"][0].update(count=-1), "accessor count"),
    (lambda d: d["accessors"][0].update(byteOffset=2 ** 63), "accessor byteOffset"),
    (lambda d: d["accessors"][0].update(byteOffset=8), "accessor exceeds its bufferView"),
    (lambda d: d["bufferViews"][0].update(byteStride=4), "byteStride below"),
    (lambda d: d["bufferViews"][0].update(byteStride=256), "bufferView byteStride"),
    (lambda d: d["accessors"][2].update(componentType=-5123), "accessor componentType"),
    (lambda d: d["buffers"][0].update(byteLength=2.0 ** 70), "buffer byteLength"),
])
def test_gltf_rejects_crafted_accessors(tmp_path, edit, msg):
    """Accessor and bufferView fields that are negative, fractional, huge or
    point outside their buffer are refused before any read."""
    with pytest.raises(PyreError, match=msg):
        pyre_amd.Preset.from_gltf(_crafted(tmp_path, "x.gltf", edit))


def test_gltf_crafted_baseline_loads(tmp_path):
    p = pyre_amd.Preset.from_gltf(_crafted(tmp_path, "ok.gltf", lambda d: None))
    m = p.arrays()["meshes"][0]
    assert m["positions"].tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    assert m["indices"].reshape(-1).tolist() == [0, 1, 2]
    # an explicit stride equal to the element size, and a strided view
    p = pyre_amd.Preset.from_gltf(_crafted(tmp_path, "s.gltf", lambda d: d["bufferViews"][0].update(byteStride=12)))
    assert p.arrays()["meshes"][0]["positions"].tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
