"""Field loaders, PNG codec, result writers and test runner -- CPU checks.

The PNG decoder is checked pixel for pixel against Pillow (an independent decoder) on
images written here with every colour type, bit depth, filter type and Adam7 interlacing,
and on two of the reference's own images (tests/golden/7_ring_*.png, copied from its
data/ folder).  The conversions applied after decoding restate stb_image's documented
behaviour (grey = (77 r + 150 g + 29 b) >> 8, 16-bit -> high byte, 8-bit -> float as
(v/255)^2.2); stb itself is not available here, so those three are checked against their
published definitions only (parity with stb unpinned).
"""
import io
import os
import struct
import zlib

import numpy as np
import pytest
from PIL import Image as PILImage

import ghcm_amd as g
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- a PNG writer for the tests (independent of the library's encoder) ----------------

def _chunk(t: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)


def _filter_rows(raw_rows, bpp, rng):
    out, prev = [], None
    for row in raw_rows:
        f = int(rng.integers(0, 5))
        cur = np.frombuffer(row, dtype=np.uint8).astype(np.int32)
        up = np.zeros_like(cur) if prev is None else prev
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]]) if len(cur) > bpp else np.zeros_like(cur)
        ul = np.concatenate([np.zeros(bpp, np.int32), up[:-bpp]]) if len(cur) > bpp else np.zeros_like(cur)
        if f == 0:
            enc = cur
        elif f == 1:
            enc = cur - left
        elif f == 2:
            enc = cur - up
        elif f == 3:
            enc = cur - (left + up) // 2
        else:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
            enc = cur - pred
        out.append(bytes([f]) + (enc & 0xFF).astype(np.uint8).tobytes())
        prev = cur
    return out


def _pack_rows(samples, depth):
    """samples: (h, w*channels) ints -> list of packed row bytes."""
    rows = []
    for r in samples:
        if depth == 16:
            rows.append(r.astype(">u2").tobytes())
        elif depth == 8:
            rows.append(r.astype(np.uint8).tobytes())
        else:
            per = 8 // depth
            n = (len(r) + per - 1) // per
            b = np.zeros(n, np.uint8)
            for i, v in enumerate(r):
                b[i // per] |= (int(v) << (8 - depth * (i % per + 1))) & 0xFF
            rows.append(b.tobytes())
    return rows


def write_png(samples, w, h, ctype, depth, interlace=False, palette=None, trns=None, seed=0):
    """samples: (h, w, channels) array of raw sample values (palette indices for type 3)."""
    rng = np.random.default_rng(seed)
    ch = samples.shape[2]
    bpp = max(1, ch * depth // 8)
    data = b""
    if not interlace:
        data = b"".join(_filter_rows(_pack_rows(samples.reshape(h, w * ch), depth), bpp, rng))
    else:
        for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4),
                               (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)):
            sub = samples[y0::dy, x0::dx]
            if sub.size == 0:
                continue
            ph, pw = sub.shape[:2]
            data += b"".join(_filter_rows(_pack_rows(sub.reshape(ph, pw * ch), depth), bpp, rng))
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        out += _chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(data, 9)
    out += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])   # split IDAT
    return out + _chunk(b"IEND", b"")


def pil_decode(png: bytes) -> np.ndarray:
    im = PILImage.open(io.BytesIO(png))
    im.load()
    if im.mode == "P":
        im = im.convert("RGBA" if "transparency" in im.info else "RGB")
    elif im.mode == "1":
        im = im.convert("L")
    a = np.asarray(im)
    if a.ndim == 2:
        a = a[:, :, None]
    return a


CASES = [  # (colour type, depth, channels)
    (0, 1, 1), (0, 2, 1), (0, 4, 1), (0, 8, 1), (0, 16, 1),
    (2, 8, 3), (2, 16, 3), (4, 8, 2), (4, 16, 2), (6, 8, 4), (6, 16, 4),
    (3, 1, 1), (3, 2, 1), (3, 4, 1), (3, 8, 1),
]


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,depth,ch", CASES)
def test_png_decode_matches_pillow(ctype, depth, ch, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    w, h = 37, 23                                   # odd sizes exercise partial bytes and passes
    if ctype == 3:
        npal = min(256, 1 << depth)
        palette = rng.integers(0, 256, (npal, 3))
        samples = rng.integers(0, npal, (h, w, 1))
        trns = list(rng.integers(0, 256, npal // 2)) if depth >= 2 else None
        png = write_png(samples, w, h, 3, depth, interlace, palette=palette, trns=trns)
    else:
        top = (1 << depth) - 1
        # smooth ramps plus noise so every filter type gets used with real data
        yy, xx = np.mgrid[0:h, 0:w]
        base = ((xx * 7 + yy * 3)[:, :, None] * (np.arange(ch) + 1)) % (top + 1)
        samples = (base + rng.integers(0, 3, (h, w, ch))) % (top + 1)
        png = write_png(samples, w, h, ctype, depth, interlace)
    ours = g.png_decode8(png)
    ref = pil_decode(png)
    if ref.dtype != np.uint8:                       # 16-bit grey: Pillow keeps 16 bits
        ref = (ref.astype(np.uint32) >> 8).astype(np.uint8)
    if ch == 2 and ref.shape[2] == 4:               # Pillow widens 16-bit grey+alpha to RGBA
        ref = ref[:, :, [0, 3]]
    assert ours.shape == ref.shape, (ours.shape, ref.shape)
    assert np.array_equal(ours, ref)


@pytest.mark.parametrize("name", ["7_ring_32_2d.png", "7_ring_256_2d.png"])
def test_reference_images_decode_like_pillow(name):
    png = open(os.path.join(GOLDEN, name), "rb").read()
    ours = g.png_decode8(png)
    ref = pil_decode(png)
    assert np.array_equal(ours, ref)


def test_stb_channel_conversions():
    rng = np.random.default_rng(5)
    rgba = rng.integers(0, 256, (9, 11, 4))
    png = write_png(rgba, 11, 9, 6, 8)
    grey = g.png_decode8(png, 1)[:, :, 0].astype(np.int64)
    r, gg, b = rgba[..., 0], rgba[..., 1], rgba[..., 2]
    assert np.array_equal(grey, (r * 77 + gg * 150 + b * 29) >> 8)
    assert np.array_equal(g.png_decode8(png, 3), rgba[..., :3])
    # 16-bit: conversion at 16 bits, then the high byte
    rgb16 = rng.integers(0, 65536, (5, 6, 3))
    png16 = write_png(rgb16, 6, 5, 2, 16)
    grey16 = g.png_decode8(png16, 1)[:, :, 0].astype(np.int64)
    y16 = (rgb16[..., 0] * 77 + rgb16[..., 1] * 150 + rgb16[..., 2] * 29) >> 8
    assert np.array_equal(grey16, y16 >> 8)
    # grey source -> 3 channels replicates
    gr = rng.integers(0, 256, (4, 5, 1))
    assert np.array_equal(g.png_decode8(write_png(gr, 5, 4, 0, 8), 3), np.repeat(gr, 3, axis=2))


def test_ldr_to_float_mapping(tmp_path):
    ramp = np.arange(256, dtype=np.int64).reshape(16, 16, 1)
    p = tmp_path / "ramp.png"
    p.write_bytes(write_png(ramp, 16, 16, 0, 8))
    f = g.image_load_f32(p, 1)[:, :, 0].ravel()
    assert f[0] == 0.0 and f[255] == 1.0
    assert np.all(np.diff(f) > 0)
    expect = np.array([np.float32(np.float64(np.float32(v) / np.float32(255.0)) ** np.float64(np.float32(2.2)))
                       for v in range(256)], dtype=np.float32)
    assert np.array_equal(f, expect)
    # field2d_img_monochannel is the same with one channel
    assert np.array_equal(g.field_load_mono(p).ravel(), f)


def test_png_writer_round_trip(tmp_path):
    rng = np.random.default_rng(9)
    for ch in (1, 2, 3, 4):
        img = rng.integers(0, 256, (33, 47, ch)).astype(np.uint8)
        img[10:20] = 200                             # long runs for the matcher
        p = tmp_path / f"w{ch}.png"
        g.png_write8(p, img)
        png = p.read_bytes()
        assert np.array_equal(pil_decode(png), img)
        assert np.array_equal(g.png_decode8(png), img)


def test_time_field_heat_map(tmp_path):
    rng = np.random.default_rng(3)
    t = rng.uniform(0, 10, (20, 30)).astype(np.float32)
    t[3, 4] = np.inf
    p = tmp_path / "t.png"
    g.write_time_field_png(p, t)
    img = pil_decode(p.read_bytes())
    fin = np.isfinite(t)
    lo, hi = t[fin].min(), t[fin].max()
    h = np.float32(5.0 / 6.0) * (np.where(fin, t, lo) - lo) / np.float32(hi - lo)
    # hue -> RGB of src/util_files.cpp:13-27
    s = np.fmod(h, np.float32(1)) * np.float32(6)
    x = ((np.float32(1) - np.abs(np.fmod(s, np.float32(2)) - np.float32(1))) * np.float32(255)).astype(np.uint8)
    k = s.astype(np.int32)
    z, f = np.zeros_like(x), np.full_like(x, 255)
    table = [(f, x, z), (x, f, z), (z, f, x), (z, x, f), (x, z, f), (f, z, x)]
    exp = np.zeros((20, 30, 3), np.uint8)
    for i in range(6):
        m = (k == i) & fin
        for c in range(3):
            exp[..., c][m] = table[i][c][m]
    assert np.array_equal(img, exp)
    assert np.all(img[3, 4] == 0)
    # explicit range, and an all-INF field stays black
    g.write_time_field_png(p, np.full((4, 4), np.inf, np.float32))
    assert not pil_decode(p.read_bytes()).any()


def _read_hdr(data: bytes):
    head, _, rest = data.partition(b"\n\n")
    dims, _, body = rest.partition(b"\n")
    _, h, _, w = dims.split()
    h, w = int(h), int(w)
    out = np.zeros((h, w), np.float64)
    pos = 0
    for y in range(h):
        if w >= 8 and body[pos] == 2 and body[pos + 1] == 2:
            pos += 4
            comps = []
            for c in range(4):
                vals = []
                while len(vals) < w:
                    n = body[pos]
                    pos += 1
                    if n > 128:
                        vals += [body[pos]] * (n - 128)
                        pos += 1
                    else:
                        vals += list(body[pos:pos + n])
                        pos += n
                comps.append(np.array(vals))
            r, e = comps[0], comps[3]
        else:
            px = np.frombuffer(body[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
            r, e = px[:, 0], px[:, 3]
        out[y] = np.where(e > 0, (r + 0.5) * np.ldexp(1.0, e.astype(np.int64) - 136), 0.0)
    return out


@pytest.mark.parametrize("w", [5, 64])
def test_hdr_writer(tmp_path, w):
    rng = np.random.default_rng(w)
    f = rng.uniform(0.01, 100, (7, w)).astype(np.float32)
    f[:, 10 % w:] = 3.5                            # a run
    f[0, 0] = 0.0
    p = tmp_path / "f.hdr"
    g.write_field_hdr(p, f)
    back = _read_hdr(p.read_bytes())
    nz = f > 0
    assert np.all(np.abs(back[nz] - f[nz]) <= f[nz] / 128)
    assert back[0, 0] == 0.0


def test_blob_layout(tmp_path):
    f = np.arange(12, dtype=np.float32).reshape(3, 4)
    p = tmp_path / "r.blob"
    g.blob_write(p, f)
    raw = p.read_bytes()
    x, y = struct.unpack("=QQ", raw[:16])
    assert (x, y) == (4, 3) and len(raw) == 16 + 48
    assert np.array_equal(np.frombuffer(raw[16:], np.float32).reshape(3, 4), f)
    assert np.array_equal(g.blob_read(p), f)


def test_blob_with_a_wrapping_size_header_is_refused(tmp_path):
    """2^31 x 2^31 floats: x * y * 4 wraps to 0 in 64 bits, which an empty payload would
    match; the header must not be believed."""
    p = tmp_path / "bad.blob"
    x = y = 1 << 31
    assert (x * y * 4) % (1 << 64) == 0
    p.write_bytes(struct.pack("=QQ", x, y))
    with pytest.raises(g.GhcmError):
        g.blob_read(p)


def _ihdr_png(w, h, idat: bytes) -> bytes:
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0))
            + _chunk(b"IDAT", idat) + _chunk(b"IEND", b""))


def test_png_with_huge_header_and_tiny_data_is_refused():
    """A 40,000 x 40,000 grey header (1.6e9 samples, under the sample cap) over a few bytes of
    data: an error, not an allocation."""
    png = _ihdr_png(40000, 40000, zlib.compress(bytes(64)))
    with pytest.raises(g.GhcmError, match="too short"):
        g.png_decode8(png)


def test_png_above_the_sample_cap_is_refused_before_inflating():
    """A 16M x 16M grey header (2^48 samples) over a stream that would inflate to 64 MiB:
    refused from the header alone (more than 2^31 samples), before any inflation."""
    png = _ihdr_png(1 << 24, 1 << 24, zlib.compress(bytes(64 << 20), 9))
    assert len(png) < 200_000
    with pytest.raises(g.GhcmError, match="too large"):
        g.png_decode8(png)
    # just above the cap: 46,341^2 grey samples > 2^31
    with pytest.raises(g.GhcmError, match="too large"):
        g.png_decode8(_ihdr_png(46341, 46341, zlib.compress(bytes(64))))


def test_png_inflation_stops_at_the_image_size():
    """A stream that inflates far past what the header needs (a small zlib bomb) decodes
    the image it describes and ignores the rest, as before the bound."""
    w, h = 16, 4
    rows = b"".join(b"\x00" + bytes(range(y, y + w)) for y in range(h))
    png = _ihdr_png(w, h, zlib.compress(rows + bytes(50 << 20), 9))
    assert len(png) < 200_000
    out = g.png_decode8(png)
    assert out.shape == (h, w, 1)
    assert np.array_equal(out[:, :, 0], np.array([list(range(y, y + w)) for y in range(h)], np.uint8))


def test_test_run_outputs_and_skip_rule(tmp_path):
    speed = g.gen_ring_speed(64, 48, 3, 1e-3, 0.25)
    cost = g.cost_from_speed(speed)
    seeds = [[10, 20]]
    g.test_run("fmm", cost, seeds, tmp_path, solver_name="fmm", job_name="rings", trials=3)
    csv = (tmp_path / "fmm_rings.csv").read_text().splitlines()
    assert len(csv) == 2
    for line, kind in zip(csv, ("device", "total")):
        parts = line.split(",")
        assert parts[:3] == ["fmm", "rings", " " + kind] and len(parts) == 6
        assert all(int(v) >= 0 for v in parts[3:])
    blob = g.blob_read(tmp_path / "fmm_rings.blob")
    assert np.array_equal(blob, oracle.fmm_f32(cost, seeds))
    assert pil_decode((tmp_path / "fmm_rings.png").read_bytes()).shape == (48, 64, 3)
    # an existing, non-empty CSV skips the job
    (tmp_path / "fmm_rings.blob").unlink()
    g.test_run("fmm", cost, seeds, tmp_path, solver_name="fmm", job_name="rings", trials=1)
    assert not (tmp_path / "fmm_rings.blob").exists()
    g.test_run("fim", cost, seeds, tmp_path, solver_name="fim", job_name="rings", trials=1,
               fim_minimum_change=1e-3)
    ref, _ = oracle.fim_f32(cost, seeds, 1e-3)
    assert np.array_equal(g.blob_read(tmp_path / "fim_rings.blob"), ref)


def test_parametric_fields():
    w, h = 40, 30
    sc = g.field_cost_sin_checkerboard(w, h, 0.75, 20.0)
    k = np.float32(np.float64(np.float32(20.0) * np.float32(2)) * np.pi)
    xx = np.arange(w, dtype=np.float32) / np.float32(w)
    yy = np.arange(h, dtype=np.float32) / np.float32(h)
    exp = np.float32(1) + np.float32(0.75) * np.sin(k * xx)[None, :] * np.sin(k * yy)[:, None]
    assert np.allclose(sc, exp, rtol=2e-6, atol=2e-6)
    # minstd_rand0 (x <- 16807 x mod 2^31-1), in row-major order
    r = g.field_cost_random(w, h, 42).ravel()
    x, vals = 42, []
    for _ in range(w * h):
        x = (16807 * x) % 2147483647
        vals.append(np.float32(1) + np.float32(np.float32(x - 1) / np.float32(2147483646)) * np.float32(5))
    assert np.array_equal(r, np.array(vals, np.float32))
