"""The encode step after the path on the device and the bytes-in / bytes-out
process() (process.rs:147-260), through the C-ABI:

  * vgpu_jpeg_encode: per-block colour conversion, downsampling, DCT and
    quantisation on the device, entropy coding on the host; the stream must
    equal the oracle's byte for byte (the oracle equals libjpeg-turbo's
    bytes: the committed golden streams are checked here too);
  * vgpu_process: PNG in, the caller's pipeline on the device, JPEG or PNG
    out; JPEG bytes equal the oracle's encode of the oracle's pipeline
    output, PNG output decodes to that output; option validation, errors
    from the build callback, cancellation, and the result record.
"""
import ctypes as C
import io
import pathlib
import zlib
import struct

import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "jpeg_vectors.npz"


def _image(shape, seed):
    r = np.random.default_rng(seed)
    return (r.integers(0, 128, size=shape) + np.linspace(0, 120, shape[1])[None, :, None]).astype(np.uint8)


def _device_jpeg(ctx, px, q, ss):
    d = v.DeviceBuffer.from_numpy(ctx, px)
    return v.jpeg_encode(ctx, d, q, ss)


def _png(px: np.ndarray) -> bytes:
    """A plain 8-bit PNG (filter 0) written here, so no codec under test
    produces the input."""
    h, w, b = px.shape
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(h))
    ct = {1: 0, 2: 4, 3: 2, 4: 6}[b]

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xffffffff)

    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ct, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")


def test_golden_streams_on_device(ctx):
    with np.load(GOLDEN, allow_pickle=False) as z:
        n = sum(1 for k in z.files if k.startswith("px"))
        for i in range(n):
            q, ss = (int(x) for x in z[f"args{i}"])
            assert _device_jpeg(ctx, z[f"px{i}"], q, ss) == bytes(z[f"jpeg{i}"]), (i, z[f"px{i}"].shape, q, ss)


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (8, 8), (15, 17), (16, 16), (10, 16), (31, 2), (130, 9),
                                   (67, 91)])
@pytest.mark.parametrize("b", [1, 3, 4])
def test_jpeg_matches_oracle(ctx, shape, b):
    px = _image(shape + (b,), shape[0] * 7 + shape[1] + b)
    for q in (0, 5, 50, 75, 90, 100):
        for ss in ((0,) if b == 1 else (0, 1, 2)):
            assert _device_jpeg(ctx, px, q, ss) == O.jpeg_encode(px, q, ss), (shape, b, q, ss)


def test_jpeg_large(ctx):
    # A 2000 x 1500 photo-sized image: many blocks per launch, 4:2:0 and 4:4:4.
    y, x = np.mgrid[0:1500, 0:2000]
    px = np.stack([(x // 7) % 256, (y // 5) % 256, (x * y // 97) % 256], -1).astype(np.uint8)
    for q, ss in ((75, 0), (92, 0)):
        assert _device_jpeg(ctx, px, q, ss) == O.jpeg_encode(px, q, ss), (q, ss)


def test_jpeg_errors(ctx):
    d = v.DeviceBuffer.from_numpy(ctx, np.zeros((4, 4, 2), np.uint8))
    with pytest.raises(v.VgpuError) as e:
        v.jpeg_encode(ctx, d, 75)
    assert e.value.status == v.api.E_UNSUPPORTED_FORMAT  # 2 bands (jpeg/encode.rs:155-157)
    d = v.DeviceBuffer.from_numpy(ctx, np.zeros((4, 4, 3), np.uint8))
    for q, ss in ((101, 0), (-1, 0), (75, 3)):
        with pytest.raises(v.VgpuError) as e:
            v.jpeg_encode(ctx, d, q, ss)
        assert e.value.status == v.api.E_INVALID_ARGUMENT
    # host memory is refused before any launch
    host = np.zeros((4, 4, 3), np.uint8)
    n = C.c_size_t()
    rc = ctx._lib.vgpu_jpeg_encode(ctx.handle, C.c_void_p(host.ctypes.data), 4, 4, 3, 75, 0, None, 0, C.byref(n))
    assert rc == v.api.E_INVALID_ARGUMENT


def test_process_thumbnail_to_jpeg(ctx):
    src = _image((301, 413, 3), 1)
    out, res = v.process(ctx, _png(src), lambda p: p.thumbnail(v.WIDTH, 100), quality=80)
    exp_px = O.Pipeline(413, 301, 3).thumbnail(O.WIDTH, 100).run(src)
    assert out == O.jpeg_encode(exp_px, 80, 0)
    assert res == {"input_format": "png", "input_dimensions": (413, 301), "output_format": "jpeg",
                   "bytes_written": len(out)}


def test_process_to_png(ctx):
    src = _image((120, 90, 4), 2)
    out, res = v.process(ctx, _png(src), lambda p: p.resize(0.37, 0.5, v.LANCZOS3).gauss_blur(1.2),
                         fmt=v.ENCODE_PNG, compression=9)
    exp_px = O.Pipeline(90, 120, 4).resize(0.37, 0.5, O.LANCZOS3).gauss_blur(1.2).run(src)
    assert np.array_equal(v.png_decode(out), exp_px)
    assert res["output_format"] == "png" and res["input_dimensions"] == (90, 120)


def test_process_identity_and_grey(ctx):
    src = _image((33, 47, 1), 3)
    out, _ = v.process(ctx, _png(src), None, quality=95)
    assert out == O.jpeg_encode(src, 95, 0)
    out, _ = v.process(ctx, _png(src), None, fmt=v.ENCODE_PNG, compression=0)
    assert np.array_equal(v.png_decode(out), src)


def test_process_errors(ctx):
    data = _png(_image((20, 30, 3), 4))
    for kw in ({"quality": 0}, {"quality": 101}, {"fmt": v.ENCODE_PNG, "compression": 10}):
        with pytest.raises(v.VgpuError) as e:
            v.process(ctx, data, None, **kw)
        assert e.value.status == v.api.E_INVALID_ARGUMENT, kw
    with pytest.raises(v.VgpuError) as e:
        v.process(ctx, b"\xff\xd8\xff\xe0" + data[4:], None)
    assert e.value.status == v.api.E_UNSUPPORTED_FORMAT  # only PNG input is decoded
    with pytest.raises(v.VgpuError) as e:
        v.process(ctx, _png(_image((20, 30, 2), 5)), None)
    assert e.value.status == v.api.E_UNSUPPORTED_FORMAT  # grey + alpha cannot be JPEG

    def bad_build(p):
        p.reduce_h(0.5)  # a reduce factor below 1 is refused by the builder

    with pytest.raises(v.VgpuError) as e:
        v.process(ctx, data, bad_build)
    assert e.value.status == v.api.E_INVALID_ARGUMENT
    tok = v.CancelToken()
    tok.cancel()
    with pytest.raises(v.VgpuError) as e:
        v.process(ctx, data, None, token=tok)
    assert e.value.status == v.api.E_CANCELLED
    # the context still works afterwards
    out, _ = v.process(ctx, data, lambda p: p.shrink(2, 2))
    assert out[:2] == b"\xff\xd8"
