"""EXR output and input (output.rs:37-193), checked without a GPU against an
independent parser written here from the OpenEXR file-format document
(attribute header, offset table, scanline and tile chunks, the ZIP/RLE byte
interleave and delta predictor), against numpy's IEEE half conversion, and
against a golden image the reference's own tests hold
(crates/cenote-pbrt/tests/golden/cornell-box.exr: tiled, RLE, f32 RGBA)."""

import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from cenote_amd import output

GOLDEN = Path(__file__).resolve().parent / "golden" / "cornell-box.exr"
ACESCG = (0.713, 0.293, 0.165, 0.830, 0.128, 0.044, 0.32168, 0.33767)  # output.rs:17-23


# --- an independent reader --------------------------------------------------------

def _cstr(b, at):
    end = b.index(b"\0", at)
    return b[at:end].decode(), end + 1


def parse(path):
    b = Path(path).read_bytes()
    magic, version = struct.unpack_from("<II", b, 0)
    assert magic == 20000630
    at, attrs = 8, {}
    while True:
        name, at = _cstr(b, at)
        if not name:
            break
        typ, at = _cstr(b, at)
        (size,) = struct.unpack_from("<i", b, at)
        attrs[name] = (typ, b[at + 4:at + 4 + size])
        at += 4 + size
    chans, raw = [], attrs["channels"][1]
    i = 0
    while raw[i] != 0:
        name, i = _cstr(raw, i)
        ptype, _, xs, ys = struct.unpack_from("<iIii", raw, i)
        i += 16
        chans.append((name, ptype))
    comp = attrs["compression"][1][0]
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    tiled = bool(version & 0x200)
    return b, attrs, chans, comp, w, h, tiled, at


def _unfilter(data):
    t = np.frombuffer(data, np.uint8).astype(np.int64)
    d = t.copy()
    d[1:] = t[1:] - 128
    t = np.cumsum(d) % 256
    n = t.size
    out = np.empty(n, np.uint8)
    half = (n + 1) // 2
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle(data, expect):
    out, i = bytearray(), 0
    while i < len(data):
        c = struct.unpack_from("b", data, i)[0]
        i += 1
        if c < 0:
            out += data[i:i - c]
            i -= c
        else:
            out += bytes([data[i]]) * (c + 1)
            i += 1
    assert len(out) == expect
    return bytes(out)


def _decode(comp, data, expect):
    if len(data) == expect:
        return data
    if comp == 1:
        return _unfilter(_rle(data, expect))
    if comp in (2, 3):
        return _unfilter(zlib.decompress(data))
    raise AssertionError(f"compression {comp}")


def channels(path):
    """Every channel as a numpy array of its stored type, (H, W)."""
    b, attrs, chans, comp, w, h, tiled, at = parse(path)
    dtype = {0: np.uint32, 1: np.float16, 2: np.float32}
    out = {n: np.zeros((h, w), dtype[t]) for n, t in chans}
    size = {0: 4, 1: 2, 2: 4}

    def store(x0, y0, span, lines, raw):
        off = 0
        for ly in range(lines):
            for n, t in chans:
                nb = span * size[t]
                out[n][y0 + ly, x0:x0 + span] = np.frombuffer(raw[off:off + nb], dtype[t])
                off += nb

    px = sum(size[t] for _, t in chans)
    if tiled:
        tw, th, _ = struct.unpack("<IIB", attrs["tiles"][1])
        count = ((w + tw - 1) // tw) * ((h + th - 1) // th)
        for k in range(count):
            (o,) = struct.unpack_from("<Q", b, at + 8 * k)
            dx, dy, _, _, sz = struct.unpack_from("<iiiii", b, o)
            x0, y0 = dx * tw, dy * th
            span, lines = min(tw, w - x0), min(th, h - y0)
            store(x0, y0, span, lines, _decode(comp, b[o + 20:o + 20 + sz], span * lines * px))
    else:
        lines = 16 if comp == 3 else 1
        for k in range((h + lines - 1) // lines):
            (o,) = struct.unpack_from("<Q", b, at + 8 * k)
            y, sz = struct.unpack_from("<ii", b, o)
            n = min(lines, h - y)
            store(0, y, w, n, _decode(comp, b[o + 8:o + 8 + sz], n * w * px))
    return attrs, [n for n, _ in chans], out


# --- the reference's tests, restated ----------------------------------------------

def test_exr_round_trips_exactly(tmp_path):
    """output.rs:200-220: f32 channels are lossless."""
    pixels = (np.arange(3 * 2 * 4, dtype=np.float32) * 0.125).reshape(2, 3, 4)
    path = tmp_path / "roundtrip.exr"
    output.write_exr(path, pixels)
    back = output.read_exr(path)
    assert back.shape == (2, 3, 4)
    assert np.array_equal(back, pixels)


def test_aov_exr_layers_carry_the_named_channels(tmp_path):
    """output.rs:222-283."""
    beauty = (np.arange(8, dtype=np.float32) * 0.25).reshape(1, 2, 4)
    albedo = np.full((1, 2, 4), 0.5, np.float32)
    normal = np.full((1, 2, 4), 0.25, np.float32)
    depth = np.array([[1.5, np.inf]], np.float32)
    path = tmp_path / "aov.exr"
    output.write_aov_exr(path, beauty, albedo, normal, depth)
    attrs, names, ch = channels(path)
    assert names == ["A", "B", "G", "R", "Z", "albedo.B", "albedo.G", "albedo.R", "normal.X", "normal.Y", "normal.Z"]
    assert struct.unpack("<8f", attrs["chromaticities"][1]) == tuple(np.float32(v) for v in ACESCG)
    for n in names:
        assert ch[n].dtype == (np.float32 if n == "Z" else np.float16), n
    assert np.array_equal(ch["R"][0], np.array([0.0, 1.0], np.float16))
    assert np.array_equal(ch["albedo.G"][0], np.array([0.5, 0.5], np.float16))
    assert np.array_equal(ch["normal.Z"][0], np.array([0.25, 0.25], np.float16))
    assert ch["Z"][0, 0].view(np.uint32) == np.float32(1.5).view(np.uint32)
    assert np.isposinf(ch["Z"][0, 1])


def test_written_headers_carry_acescg_chromaticities(tmp_path):
    """output.rs:285-295."""
    path = tmp_path / "chroma.exr"
    output.write_exr(path, np.zeros((1, 1, 4), np.float32))
    attrs, _, _ = channels(path)
    assert attrs["chromaticities"][0] == "chromaticities"
    assert struct.unpack("<8f", attrs["chromaticities"][1]) == tuple(np.float32(v) for v in ACESCG)


# --- what the format and IEEE fix -------------------------------------------------

def test_half_channels_round_to_nearest_even(tmp_path):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.normal(size=4000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 4000).astype(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 65504.0, 65520.0, 65519.99, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8,
                  1e-10, -3e-6, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11], np.float32),
    ])
    n = vals.size
    beauty = np.zeros((1, n, 4), np.float32)
    beauty[0, :, 0] = vals
    path = tmp_path / "halves.exr"
    output.write_aov_exr(path, beauty, beauty, beauty, np.zeros((1, n), np.float32))
    _, _, ch = channels(path)
    with np.errstate(over="ignore"):  # the overflowing values must round to inf, as numpy's cast does
        want = vals.astype(np.float16)
    assert np.array_equal(ch["R"][0].view(np.uint16), want.view(np.uint16))


def test_large_zip_images_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    rgba = rng.normal(size=(37, 53, 4)).astype(np.float32)  # not a multiple of the 16-line block
    path = tmp_path / "big.exr"
    output.write_exr(path, rgba)
    assert np.array_equal(output.read_exr(path), rgba)
    _, names, ch = channels(path)
    assert names == ["A", "B", "G", "R"]
    assert np.array_equal(ch["G"], rgba[..., 1])


def test_reads_the_reference_golden_like_the_independent_parser():
    rgba = output.read_exr(GOLDEN)
    attrs, names, ch = channels(GOLDEN)
    assert rgba.shape == (256, 256, 4)
    assert names == ["A", "B", "G", "R"]
    for k, n in enumerate("RGBA"):
        assert np.array_equal(rgba[..., k], ch[n].astype(np.float32))
    assert np.all(np.isfinite(rgba)) and rgba[..., :3].min() >= 0.0
    assert 0.05 < float(rgba[..., :3].mean()) < 5.0


def test_bad_files_fail_with_a_status(tmp_path):
    import cenote_amd as cn

    bad = tmp_path / "bad.exr"
    bad.write_bytes(b"not an exr at all")
    with pytest.raises(cn.CenoteError):
        output.read_exr(bad)
    with pytest.raises(cn.CenoteError):
        output.read_exr(tmp_path / "missing.exr")
    truncated = tmp_path / "cut.exr"
    truncated.write_bytes(GOLDEN.read_bytes()[:300])
    with pytest.raises(cn.CenoteError):
        output.read_exr(truncated)
