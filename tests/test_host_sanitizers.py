"""The host-side scene loaders and film sinks (pyre_amd/csrc/scene_io.cpp,
film_io.cpp, gen.cpp) built with AddressSanitizer and UndefinedBehavior-
Sanitizer (g++, host code only) and run over valid scene files, the crafted
glTF files of test_scene_io.py, truncated and corrupted Radiance files, and
seeded byte mutations of all of them. The files come from untrusted sources
(the reference's loaders return errors for malformed input:
io/pbrt/lexer.rs:55, io/pbrt/parser.rs:110, io/gltf.rs:11, io/hdr.rs:11), so
every one must load or fail with a status, with no sanitizer report."""
import base64
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import test_scene_io as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pyre_amd", "csrc")
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{os.path.join(ROOT, 'include')}"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are missing")
    d = tmp_path_factory.mktemp("asan")
    srcs = [os.path.join(CSRC, f) for f in ("scene_io.cpp", "film_io.cpp", "gen.cpp")]
    srcs.append(os.path.join(ROOT, "tests", "sanitize_driver.cpp"))
    exe = str(d / "sanitize_driver")
    r = subprocess.run(["g++", *FLAGS, *srcs, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def _corpus(tmp_path):
    files = []

    def add(kind, name, data):
        p = tmp_path / name
        p.write_bytes(data if isinstance(data, bytes) else data.encode())
        files.append((kind, str(p)))

    add("pbrt", "cornell.pbrt", T._cornell_pbrt())
    from tests.test_scene_io_emitters import SCENE as EMITTERS
    add("pbrt", "emitters.pbrt", EMITTERS)
    w, h = 16, 8
    px = (np.arange(w * h * 4) % 251).astype(np.uint8).reshape(h, w, 4)
    px[..., 3] = 130
    add("hdr", "flat.hdr", T._rgbe_header(w, h) + px.tobytes())
    rle = b"".join(bytes([2, 2, 0, w]) + b"".join(T._rle_plane(list(px[y, :, c])) for c in range(4))
                   for y in range(h))
    add("hdr", "rle.hdr", T._rgbe_header(w, h) + rle)
    add("hdr", "short.hdr", T._rgbe_header(w, h) + px.tobytes()[:100])
    add("hdr", "huge.hdr", T._rgbe_header(1 << 30, 1 << 30) + px.tobytes())
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    edits = [lambda d: None,
             lambda d: d["bufferViews"][0].update(byteOffset=2.0 ** 64 - 2048, byteLength=4096),
             lambda d: d["accessors"][0].update(count=2.0 ** 62),
             lambda d: d["accessors"][0].update(count=2 ** 30 + 1),
             lambda d: d["bufferViews"][0].update(byteStride=4),
             lambda d: d["accessors"][2].update(count=4),
             lambda d: d["accessors"][2].update(componentType=5125),
             lambda d: d["nodes"][0].update(children=[0]),
             lambda d: d["nodes"][0].update(matrix=[1e38] * 16)]
    for k, e in enumerate(edits):
        doc, blob = T._gltf_doc(pos, pos, np.array([0, 1, 2], np.uint16), [{"mesh": 0}], [0])
        doc["meshes"][0]["primitives"] = doc["meshes"][0]["primitives"][:1]
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
        e(doc)
        add("gltf", f"g{k}.gltf", json.dumps(doc))
    # seeded mutations of every file above: flipped bytes, cut tails, repeats
    rng = np.random.default_rng(7)
    for kind, path in list(files):
        data = open(path, "rb").read()
        for m in range(40):
            b = bytearray(data)
            op = m % 3
            if op == 0:
                for _ in range(1 + m % 5):
                    b[int(rng.integers(len(b)))] = int(rng.integers(256))
            elif op == 1:
                b = b[:int(rng.integers(len(b)))]
            else:
                i = int(rng.integers(len(b)))
                b = b[:i] + b[i:i + 64] * 3 + b[i:]
            add(kind, f"m{m}_{os.path.basename(path)}", bytes(b))
    return files


def test_loaders_and_sinks_under_asan_ubsan(driver, tmp_path):
    files = _corpus(tmp_path)
    args = [driver, str(tmp_path)]
    for kind, path in files:
        args += [kind, path]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout.splitlines()
    assert r.returncode == 0, (out[-3:], r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    rc = {line.split()[1]: int(line.split("rc=")[1]) for line in out if line.split()[0] in ("pbrt", "gltf", "hdr")}
    assert len(rc) == len(files)
    name = lambda n: str(tmp_path / n)
    # the valid files load; the crafted and truncated ones fail with a status
    for good in ("cornell.pbrt", "emitters.pbrt", "flat.hdr", "rle.hdr", "g0.gltf"):
        assert rc[name(good)] == 0, good
    for bad in ("short.hdr", "huge.hdr", "g1.gltf", "g2.gltf", "g3.gltf", "g4.gltf", "g5.gltf", "g7.gltf"):
        assert rc[name(bad)] < 0, bad
    assert all(v in (0, -1, -6) for v in rc.values())
    assert "sinks rc=0" in out
