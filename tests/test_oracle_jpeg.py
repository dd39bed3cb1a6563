"""The oracle's baseline JPEG encoder (oracle/jpeg.c), the checker of the
device encode after the path (jpeg/encode.rs:169-258 through turbojpeg over
libjpeg-turbo), pinned by:

  * the committed golden streams (tests/golden/jpeg_vectors.npz, libjpeg-
    turbo's bytes through Pillow): byte for byte;
  * a sweep against Pillow where it is importable: grey, 4:4:4 and 4:2:0,
    sizes that are and are not whole MCUs (dummy blocks, padded rows),
    quality 0..100 (0 encodes as 1), RGBA with the alpha dropped;
  * what the format fixes without any library: a constant image has no AC
    coefficients and decodes to itself, the SOF carries the size, the stream
    is SOI ... EOI with every 0xFF in the entropy data stuffed.
"""
import io
import pathlib

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "jpeg_vectors.npz"


def _golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        n = sum(1 for k in z.files if k.startswith("px"))
        return [(z[f"px{i}"], bytes(z[f"jpeg{i}"]), tuple(int(x) for x in z[f"args{i}"])) for i in range(n)]


def test_golden_streams():
    cases = _golden()
    assert len(cases) >= 10
    for px, exp, (q, ss) in cases:
        assert O.jpeg_encode(px, q, ss) == exp, (px.shape, q, ss)


def _segments(data: bytes):
    """(marker, payload) up to SOS, then the entropy-coded bytes."""
    assert data[:2] == b"\xff\xd8" and data[-2:] == b"\xff\xd9"
    p, segs = 2, []
    while True:
        assert data[p] == 0xff
        m = data[p + 1]
        n = int.from_bytes(data[p + 2:p + 4], "big")
        segs.append((m, data[p + 4:p + 2 + n]))
        p += 2 + n
        if m == 0xda:
            return segs, data[p:-2]


@pytest.mark.parametrize("h,w,b,ss", [(8, 8, 1, 0), (21, 35, 3, 1), (21, 35, 3, 2), (40, 48, 4, 0)])
def test_stream_structure(h, w, b, ss):
    px = np.random.default_rng(h * w).integers(0, 256, size=(h, w, b), dtype=np.uint8)
    segs, ecs = _segments(O.jpeg_encode(px, 75, ss))
    markers = [m for m, _ in segs]
    ncomp = 1 if b == 1 else 3
    assert markers == [0xe0] + [0xdb] * (1 if ncomp == 1 else 2) + [0xc0] + [0xc4] * (2 * (1 if ncomp == 1 else 2)) + \
        [0xda]
    sof = dict(segs)[0xc0]
    assert int.from_bytes(sof[1:3], "big") == h and int.from_bytes(sof[3:5], "big") == w and sof[5] == ncomp
    if ncomp == 3:
        assert sof[7] == (0x22 if ss != 1 else 0x11)  # auto at 75 is 4:2:0
    # every 0xFF inside the entropy-coded data is followed by a stuffed 0x00
    idx = np.flatnonzero(np.frombuffer(ecs, np.uint8) == 0xff)
    assert all(ecs[i + 1] == 0 for i in idx)


def test_quantisation_tables():
    # T.81 K.1 scaled: quality 50 is the table itself, 100 is all ones,
    # quality 0 is quality 1 (clamped to 255).
    px = np.zeros((8, 8, 1), np.uint8)
    dqt = lambda q: dict(_segments(O.jpeg_encode(px, q))[0])[0xdb][1:]
    zz = [0, 1, 8, 16, 9, 2, 3, 10]
    k1 = [16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14]  # rows 0, 1 and (2, 0)
    assert [dqt(50)[k] for k in range(8)] == [k1[n] for n in zz]
    assert set(dqt(100)) == {1}
    assert dqt(0) == dqt(1) and max(dqt(1)) == 255


def test_constant_image_has_only_dc():
    # A flat block has no AC energy: each block codes its DC difference, then
    # EOB (1010, 4 bits).  Quality 50 keeps K.1's DC step 16, so the first
    # DC is round(64 * (77 - 128) / (8 * 16)) = -26: category 5 (code 3
    # bits + 5 bits); the other 63 differences are 0 (code 00).  Total
    # 12 + 63 * 6 = 390 bits -> 49 bytes.
    px = np.full((64, 64, 1), 77, np.uint8)
    _, ecs = _segments(O.jpeg_encode(px, 50))
    assert len(ecs) == 49
    assert ecs[-1] == 0b00101011  # the last block's 00 + 1010, then two padding ones


def test_errors():
    px = np.zeros((4, 4, 2), np.uint8)
    with pytest.raises(O.OracleError):
        O.jpeg_encode(px, 75)  # two bands: not encodable (jpeg/encode.rs:155-157)
    with pytest.raises(O.OracleError):
        O.jpeg_encode(np.zeros((4, 4, 3), np.uint8), 101)


Image = None
try:  # Pillow (libjpeg-turbo) is the second, wider pin when it is present
    from PIL import Image  # noqa: F811
except ImportError:  # pragma: no cover
    pass


def _pillow(px, q, ss):
    b = px.shape[2]
    img = Image.fromarray(px[..., 0] if b == 1 else px[..., :3])
    kw = {} if b == 1 else {"subsampling": {0: 2 if max(q, 1) < 90 else 0, 1: 0, 2: 2}[ss]}
    buf = io.BytesIO()
    img.save(buf, "JPEG", quality=max(q, 1), optimize=False, **kw)
    return buf.getvalue()


@pytest.mark.skipif(Image is None, reason="Pillow not importable")
@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (6, 6), (8, 8), (15, 17), (16, 16), (23, 37), (8, 16), (10, 16),
                                   (4, 20), (31, 2), (130, 9), (40, 56)])
@pytest.mark.parametrize("b", [1, 3, 4])
def test_against_libjpeg_turbo(shape, b):
    r = np.random.default_rng(shape[0] * 1000 + shape[1] + b)
    px = (r.integers(0, 128, size=shape + (b,)) + np.linspace(0, 120, shape[1])[None, :, None]).astype(np.uint8)
    for q in (0, 1, 5, 20, 50, 75, 89, 90, 95, 100):
        for ss in ((0,) if b == 1 else (0, 1, 2)):
            assert O.jpeg_encode(px, q, ss) == _pillow(px, q, ss), (shape, b, q, ss)


@pytest.mark.skipif(Image is None, reason="Pillow not importable")
def test_decodes_back():
    r = np.random.default_rng(9)
    y, x = np.mgrid[0:96, 0:128]
    px = np.stack([x * 2, y * 2, (x + y)], -1).astype(np.uint8)
    px = np.clip(px.astype(int) + r.integers(-3, 4, size=px.shape), 0, 255).astype(np.uint8)
    for q, tol in ((95, 3.0), (75, 6.0)):
        back = np.array(Image.open(io.BytesIO(O.jpeg_encode(px, q, 1))).convert("RGB")).astype(float)
        assert np.abs(back - px).mean() < tol, q
