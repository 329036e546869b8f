"""Film sinks (crates/pyre/src/film.rs:117-264) on the host: the conversions
of save_png / save_aov_pngs / save_exr through pyre_film_pack, and the PNG /
EXR files of pyre_write_png / pyre_write_exr read back with independent
readers written here (zlib for the PNG stream, the OpenEXR layout by hand).

Pinned by: exact float32 restatements of the linear conversions (normal,
depth, EXR planes), the special values of the gamma conversion (NaN,
negatives, 0, 1, > 1) and a float64 restatement of powf for the rest (a
one-unit difference is allowed where powf's last-ulp rounding meets a
truncation boundary); the file formats by decoding them.
"""
import struct
import zlib

import numpy as np
import pytest

import pyre_amd
from pyre_amd import _lib as L

F = np.float32


def _film(w=37, h=23, seed=5):
    f = pyre_amd.Film(w, h)
    rng = np.random.default_rng(seed)
    s = (rng.random((w * h, 10), dtype=np.float32) * 1.4 - 0.2).astype(F)
    s[:, 6:9] = rng.uniform(-1.2, 1.2, (w * h, 3)).astype(F)
    s[:, 9] = np.where(rng.random(w * h) < 0.2, 0.0, rng.uniform(0.5, 9.0, w * h)).astype(F)
    s[0, :3] = [np.nan, -1.0, 2.0]
    s[1, :3] = [0.0, 1.0, 1e-30]
    f.samples[:] = s
    return f


def _u8(x):
    """Rust's `f32 as u8`."""
    x = np.asarray(x, np.float64)
    return np.where(np.isnan(x) | (x <= 0), 0, np.minimum(np.floor(x), 255)).astype(np.uint8)


def _clamp01(x):
    """glam's clamp: NaN-ignoring max / min."""
    return np.fmin(np.fmax(x, F(0)), F(1))


def test_gamma_conversion(tmp_path):
    f = _film()
    for what, sl in ((L.PACK_BEAUTY_PNG, slice(0, 3)), (L.PACK_ALBEDO_PNG, slice(3, 6))):
        got = f.pack(what).reshape(-1, 3)
        c = _clamp01(f.samples[:, sl].astype(F))
        # powf(c, 1/2.2f) correctly rounded to f32, times 255 in f32
        p = np.power(c.astype(np.float64), np.float64(F(1.0) / F(2.2))).astype(F)
        want = _u8(p * F(255))
        diff = np.abs(got.astype(int) - want.astype(int))
        assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    got = f.pack(L.PACK_BEAUTY_PNG).reshape(-1, 3)
    assert got[0].tolist() == [0, 0, 255] and got[1].tolist() == [0, 255, 0]


def test_linear_conversions_and_planes():
    f = _film()
    s = f.samples
    got = f.pack(L.PACK_NORMAL_PNG).reshape(-1, 3)
    assert np.array_equal(got, _u8(_clamp01(s[:, 6:9] * F(0.5) + F(0.5)) * F(255)))
    d = s[:, 9]
    hit = d > 0
    lo, hi = d[hit].min(), d[hit].max()
    rng_ = np.fmax(hi - lo, F(1e-6))
    v = np.where(hit, _clamp01(F(1) - (d - lo) / rng_), F(0)).astype(F)
    assert np.array_equal(f.pack(L.PACK_DEPTH_PNG).reshape(-1), _u8(v * F(255)))
    pl = f.pack(L.PACK_EXR).reshape(10, -1)
    order = [2, 1, 6, 7, 8, 0, 9, 5, 4, 3]  # B G N.X N.Y N.Z R Z albedo.B albedo.G albedo.R
    for k, col in enumerate(order):
        assert np.array_equal(pl[k].view(np.uint32), s[:, col].view(np.uint32)), k
    # a film with no hit: depth all black
    f.samples[:, 9] = 0
    assert not f.pack(L.PACK_DEPTH_PNG).any()


def _read_png(path):
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    i, chunks = 8, {}
    while i < len(b):
        n, = struct.unpack(">I", b[i:i + 4])
        typ, data = b[i + 4:i + 8], b[i + 8:i + 8 + n]
        crc, = struct.unpack(">I", b[i + 8 + n:i + 12 + n])
        assert crc == zlib.crc32(typ + data) & 0xFFFFFFFF
        chunks.setdefault(typ, b"")
        chunks[typ] += data
        i += 12 + n
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert (depth, comp, filt, inter) == (8, 0, 0, 0) and b"IEND" in chunks
    ch = {0: 1, 2: 3}[ctype]
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * ch)
    assert not rows[:, 0].any()  # filter 0
    px = rows[:, 1:].reshape(h, w, ch)
    return px[..., 0] if ch == 1 else px


def test_png_files(tmp_path):
    f = _film(w=300, h=250)  # more than one 64 KiB stored block
    f.save_png(tmp_path / "a.png")
    assert np.array_equal(_read_png(tmp_path / "a.png"), f.pack(L.PACK_BEAUTY_PNG))
    f.save_aov_pngs(tmp_path / "a.png")
    for name, what in (("albedo", L.PACK_ALBEDO_PNG), ("normal", L.PACK_NORMAL_PNG), ("depth", L.PACK_DEPTH_PNG)):
        assert np.array_equal(_read_png(tmp_path / f"a.{name}.png"), f.pack(what)), name


def _read_exr(path):
    b = open(path, "rb").read()
    assert b[:4] == b"\x76\x2f\x31\x01" and b[4] == 2
    i, attrs = 8, {}
    while b[i] != 0:
        j = b.index(b"\0", i)
        name = b[i:j].decode()
        k = b.index(b"\0", j + 1)
        typ = b[j + 1:k].decode()
        n, = struct.unpack("<i", b[k + 1:k + 5])
        attrs[name] = (typ, b[k + 5:k + 5 + n])
        i = k + 5 + n
    i += 1
    typ, cl = attrs["channels"]
    assert typ == "chlist"
    names, p = [], 0
    while cl[p] != 0:
        q = cl.index(b"\0", p)
        names.append(cl[p:q].decode())
        ptype, = struct.unpack("<i", cl[q + 1:q + 5])
        assert ptype == 2  # FLOAT
        p = q + 17
    assert attrs["compression"] == ("compression", b"\0")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack(f"<{h}Q", b[i:i + 8 * h])
    img = np.zeros((len(names), h, w), np.float32)
    for y, o in enumerate(offs):
        yy, size = struct.unpack("<ii", b[o:o + 8])
        assert yy == y and size == 4 * w * len(names)
        img[:, y] = np.frombuffer(b[o + 8:o + 8 + size], "<f4").reshape(len(names), w)
    return names, img


@pytest.mark.parametrize("aovs", [(True, True, True), (False, False, False), (True, False, True)])
def test_exr_files(tmp_path, aovs):
    f = _film()
    f.save_exr(tmp_path / "a.exr", *aovs)
    names, img = _read_exr(tmp_path / "a.exr")
    all_names = ["B", "G", "N.X", "N.Y", "N.Z", "R", "Z", "albedo.B", "albedo.G", "albedo.R"]
    want = [n for n in all_names if n in ("B", "G", "R")
            or (aovs[0] and n.startswith("albedo")) or (aovs[1] and n.startswith("N.")) or (aovs[2] and n == "Z")]
    assert names == want == sorted(want)
    planes = f.pack(L.PACK_EXR)
    for k, n in enumerate(names):
        assert np.array_equal(img[k].view(np.uint32), planes[all_names.index(n)].view(np.uint32)), n


def test_sinks_reject_bad_arguments(tmp_path):
    with pytest.raises(pyre_amd.PyreError):
        pyre_amd.write_png(tmp_path / "x.png", np.zeros((0, 4, 3), np.uint8))
    with pytest.raises(pyre_amd.PyreError) as e:
        pyre_amd.write_png(tmp_path / "nodir" / "x.png", np.zeros((2, 2, 3), np.uint8))
    assert e.value.code == -6
