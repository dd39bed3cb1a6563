"""The host side of the encode step and of process()'s decode: PNG decode
(the png crate's EXPAND decode for U8, png/decode_full.rs:101-108, 470-520)
and PNG encode (png/encode.rs at a zlib level), through the C-ABI with no
GPU call.

Pins: Pillow's PNG decoder on the files Pillow writes (every colour type,
bit depths 1, 2, 4, 8, palettes with and without tRNS, grey and RGB tRNS
keys), an Adam7 writer in this file whose output Pillow decodes to the same
pixels, and lossless round trips of our encoder through Pillow.  The
reference's own PNG bytes come from the png crate's encoder (a different
deflate), so encode parity is the decoded pixels, not the bytes."""
import io
import struct
import zlib

import numpy as np
import pytest

import viprs_amd as v

Image = pytest.importorskip("PIL.Image")


def _pil_png(img, **kw) -> bytes:
    b = io.BytesIO()
    img.save(b, "PNG", **kw)
    return b.getvalue()


def _chunk(t: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xffffffff)


def _adam7_png(px: np.ndarray, ctype: int) -> bytes:
    """An interlaced 8-bit PNG (filter 1 on every row of every pass)."""
    h, w, ch = px.shape
    raw = bytearray()
    for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                           (0, 1, 1, 2)):
        sub = px[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for row in sub.reshape(sub.shape[0], -1):
            r = row.astype(np.int32)
            sub_f = (r - np.concatenate([np.zeros(ch, np.int32), r[:-ch]])) & 255
            raw += b"\x01" + sub_f.astype(np.uint8).tobytes()
    ihdr = struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 1)
    return b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(bytes(raw))) + \
        _chunk(b"IEND", b"")


def _rng_image(shape, seed):
    r = np.random.default_rng(seed)
    base = np.linspace(0, 200, shape[1])[None, :, None] + r.integers(0, 56, size=shape)
    return base.astype(np.uint8)


@pytest.mark.parametrize("mode,bands", [("L", 1), ("LA", 2), ("RGB", 3), ("RGBA", 4)])
@pytest.mark.parametrize("shape", [(1, 1), (7, 13), (64, 33)])
def test_decode_8bit(mode, bands, shape):
    a = _rng_image(shape + (bands,), bands)
    img = Image.fromarray(a[..., 0] if bands == 1 else a, mode)
    for kw in ({}, {"optimize": True}, {"compress_level": 0}):
        got = v.png_decode(_pil_png(img, **kw))
        assert got.shape == a.shape and np.array_equal(got, a), kw


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
@pytest.mark.parametrize("trns", [False, True])
def test_decode_palette(bits, trns):
    r = np.random.default_rng(bits)
    n = 1 << bits
    pal = r.integers(0, 256, size=(n, 3), dtype=np.uint8)
    idx = r.integers(0, n, size=(19, 29), dtype=np.uint8)
    img = Image.fromarray(idx, "P")
    img.putpalette(pal.ravel().tolist())
    kw = {"bits": bits}
    if trns:
        kw["transparency"] = bytes(int(x) for x in r.integers(0, 256, size=max(1, n // 2)))
    data = _pil_png(img, **kw)
    got = v.png_decode(data)
    exp = np.array(Image.open(io.BytesIO(data)).convert("RGBA" if trns else "RGB"))
    assert got.shape == exp.shape and np.array_equal(got, exp)


@pytest.mark.parametrize("bits", [1, 2, 4])
def test_decode_low_bit_grey(bits):
    # Grey samples scaled to 0..255 by 255 / (2^bits - 1).
    r = np.random.default_rng(bits)
    levels = r.integers(0, 1 << bits, size=(11, 21))
    ihdr = struct.pack(">IIBBBBB", 21, 11, bits, 0, 0, 0, 0)
    raw = bytearray()
    for row in levels:
        bitstr = "".join(format(int(x), f"0{bits}b") for x in row)
        bitstr += "0" * (-len(bitstr) % 8)
        raw += b"\x00" + int(bitstr, 2).to_bytes(len(bitstr) // 8, "big")
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(bytes(raw))) + \
        _chunk(b"IEND", b"")
    got = v.png_decode(data)
    assert np.array_equal(got[..., 0], (levels * (255 // ((1 << bits) - 1))).astype(np.uint8))
    pil = np.array(Image.open(io.BytesIO(data)).convert("L"))
    assert np.array_equal(got[..., 0], pil)


def test_decode_trns_keys():
    # grey and RGB tRNS keys become an alpha band: 0 on the key, else 255
    a = _rng_image((9, 17, 3), 3)
    a[2, 3] = (10, 20, 30)
    data = _pil_png(Image.fromarray(a, "RGB"), transparency=(10, 20, 30))
    got = v.png_decode(data)
    assert got.shape == (9, 17, 4)
    assert np.array_equal(got[..., :3], a)
    key = np.all(a == (10, 20, 30), axis=-1)
    assert np.array_equal(got[..., 3], np.where(key, 0, 255).astype(np.uint8))
    g = a[..., 0].copy()
    data = _pil_png(Image.fromarray(g, "L"), transparency=int(g[4, 4]))
    got = v.png_decode(data)
    assert got.shape == (9, 17, 2)
    assert np.array_equal(got[..., 1], np.where(g == g[4, 4], 0, 255).astype(np.uint8))


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (9, 13), (40, 41)])
@pytest.mark.parametrize("bands,ctype", [(1, 0), (3, 2), (4, 6)])
def test_decode_adam7(shape, bands, ctype):
    a = _rng_image(shape + (bands,), shape[0] + bands)
    data = _adam7_png(a, ctype)
    pil = np.array(Image.open(io.BytesIO(data)))
    assert np.array_equal(pil.reshape(a.shape), a)  # pins the writer above
    got = v.png_decode(data)
    assert np.array_equal(got, a)


def test_decode_errors():
    a = _rng_image((5, 6, 3), 0)
    good = _pil_png(Image.fromarray(a, "RGB"))
    with pytest.raises(v.VgpuError) as e:
        v.png_decode(b"\xff\xd8\xff" + good[3:])
    assert e.value.status == v.api.E_UNSUPPORTED_FORMAT
    bad = bytearray(good)
    bad[40] ^= 0x55  # inside IDAT: the chunk CRC no longer matches
    with pytest.raises(v.VgpuError):
        v.png_decode(bytes(bad))
    with pytest.raises(v.VgpuError):
        v.png_decode(good[: len(good) // 2])
    wide = _pil_png(Image.fromarray(a[..., 0].astype(np.uint16) * 257))
    with pytest.raises(v.VgpuError) as e:
        v.png_decode(wide)
    assert e.value.status == v.api.E_UNSUPPORTED_FORMAT  # 16-bit file, U8 decode


@pytest.mark.parametrize("bands", [1, 2, 3, 4])
@pytest.mark.parametrize("level", [0, 1, 6, 9])
def test_encode_round_trip(bands, level):
    a = _rng_image((23, 31, bands), bands * 10 + level)
    data = v.png_encode(a, level)
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pil = np.array(Image.open(io.BytesIO(data)))
    assert np.array_equal(pil.reshape(a.shape), a)
    assert np.array_equal(v.png_decode(data), a)


def test_encode_levels_and_errors():
    a = _rng_image((64, 64, 3), 5)
    assert len(v.png_encode(a, 9)) < len(v.png_encode(a, 0))
    with pytest.raises(v.VgpuError):
        v.png_encode(a, 10)
    with pytest.raises(v.VgpuError):
        v.png_encode(np.zeros((4, 4, 5), np.uint8), 6)
