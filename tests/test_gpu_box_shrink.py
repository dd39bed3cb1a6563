"""The PNG source's box pre-shrink on the device (vgpu_pipeline_box_shrink,
vgpu_pipeline_thumbnail_shrink_on_load): software_box_shrink_generic
(sources/decoder_source/shrink.rs:47-117) for an eager backing, the u8 row
decoder's ceil form (codecs/src/png/decode_full.rs:320-452), and the
thumbnail planned again over the shrunk image (builder/resample.rs:871-906).
Every result is compared bit for bit with the oracle."""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import float_image, noise_u16, noise_u8

pytestmark = pytest.mark.gpu
FMT = {np.dtype(np.uint8): v.U8, np.dtype(np.uint16): v.U16, np.dtype(np.float32): v.F32}


def _run(ctx, p, src):
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    return out.numpy()


def _same(got, exp, what):
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}"


@pytest.mark.parametrize("bands", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("f", [2, 3, 4, 5, 8, 16])
@pytest.mark.parametrize("ceil", [False, True])
def test_box_shrink_u8(ctx, bands, f, ceil):
    h, w = 7 * f + 3, 11 * f + 5
    print(f"[box] u8 {w}x{h}x{bands} f={f} ceil={ceil}", flush=True)
    src = noise_u8(h, w, bands, f * 10 + bands)
    p = v.Pipeline(ctx, w, h, bands).box_shrink(f, ceil)
    got = _run(ctx, p, src)
    _same(got, O.box_shrink(src, f, ceil), f"box u8 f={f} bands={bands} ceil={ceil}")
    assert p.stage_kernels() == ["box_shrink"]


@pytest.mark.parametrize("w,h,f", [(3, 5, 4), (5, 40, 8), (40, 2, 16), (1, 1, 2)])
def test_box_shrink_window_cut(ctx, w, h, f):
    # Sources smaller than the factor: one output pixel on that axis, its
    # window the whole axis (shrink.rs:58-59,70-73).
    src = noise_u8(h, w, 3, w * h)
    _same(_run(ctx, v.Pipeline(ctx, w, h, 3).box_shrink(f), src), O.box_shrink(src, f), f"cut {w}x{h} f={f}")


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
@pytest.mark.parametrize("f", [2, 3, 16])
def test_box_shrink_u16_f32(ctx, dtype, f):
    h, w, b = 5 * f + 1, 9 * f + 2, 3
    src = noise_u16(h, w, b, f) if dtype == np.uint16 else float_image(h, w, b, f)
    p = v.Pipeline(ctx, w, h, b, FMT[np.dtype(dtype)]).box_shrink(f)
    _same(_run(ctx, p, src), O.box_shrink(src, f), f"box {np.dtype(dtype).name} f={f}")


def test_box_shrink_fixture_images(ctx):
    # The reference's PNG decode tests: 17x18x3 at f = 16 (partial tail
    # blocks) and 4x4 grey + alpha at f = 2, in the row decoder's form.
    y, x = np.mgrid[0:18, 0:17].astype(np.uint32)
    tail = np.stack([(x * 13 + y) & 255, (y * 7 + x * 3) & 255, x ^ ((y * 11) & 255)], -1).astype(np.uint8)
    _same(_run(ctx, v.Pipeline(ctx, 17, 18, 3).box_shrink(16, True), tail), O.box_shrink(tail, 16, True), "tail")
    y, x = np.mgrid[0:4, 0:4].astype(np.int64)
    ga = np.stack([(x * 17 + y * 9) & 255, 255 - ((x * 21 + y * 5) & 255)], -1).astype(np.uint8)
    _same(_run(ctx, v.Pipeline(ctx, 4, 4, 2).box_shrink(2, True), ga), O.box_shrink(ga, 2, True), "ga")


def test_box_shrink_only_first(ctx):
    p = v.Pipeline(ctx, 64, 64, 3).gauss_blur(1.0)
    with pytest.raises(v.VgpuError):
        p.box_shrink(2)
    with pytest.raises(v.VgpuError):
        v.Pipeline(ctx, 64, 64, 3, v.F32).box_shrink(2, True)  # the row decoder's form is u8 only


@pytest.mark.parametrize("w,h,bands,target,tw,th,crop", [
    (2000, 1500, 3, v.WIDTH, 100, 0, False),    # hint 16: 125 x 93, SmallTile Affine tail
    (3000, 2000, 4, v.FITBOX, 300, 300, False),  # hint 4, premultiplied RGBA
    (1201, 901, 3, v.WIDTH, 250, 0, False),     # hint 2
    (1000, 800, 1, v.HEIGHT, 48, 0, False),     # hint 16, one band
    (900, 700, 3, v.WIDTH, 300, 0, False),      # no hint
    (3001, 2003, 3, v.FITBOX, 200, 150, True),  # crop
    (6000, 4000, 3, v.WIDTH, 400, 0, False),    # hint 4 over a large source
])
def test_thumbnail_shrink_on_load(ctx, w, h, bands, target, tw, th, crop):
    print(f"[box] thumbnail {w}x{h}x{bands} -> {target} {tw}x{th} crop={crop}", flush=True)
    src = noise_u8(h, w, bands, w + h + bands)
    p = v.Pipeline(ctx, w, h, bands).thumbnail_shrink_on_load(target, tw, th, crop=crop)
    exp = O.Pipeline(w, h, bands).thumbnail_shrink_on_load(target, tw, th, crop=crop).run(src)
    _same(_run(ctx, p, src), exp, f"thumbnail {w}x{h}")
    f = O.lib().orc_thumbnail_shrink_on_load(w, h, target, tw, th, int(crop))
    assert (p.stage_kernels()[0] == "box_shrink") == bool(f)
    # The no-alloc run of the same plan.
    out2 = p.output_buffer()
    p.run_noalloc(v.DeviceBuffer.from_numpy(ctx, src), out2)
    ctx.synchronize()
    _same(out2.numpy(), exp, f"thumbnail {w}x{h} no-alloc")


def test_thumbnail_shrink_on_load_16bit_and_float(ctx):
    for dtype in (np.uint16, np.float32):
        src = noise_u16(1100, 1500, 3, 5) if dtype == np.uint16 else float_image(1100, 1500, 3, 5)
        fmt = FMT[np.dtype(dtype)]
        p = v.Pipeline(ctx, 1500, 1100, 3, fmt).thumbnail_shrink_on_load(v.WIDTH, 90)
        exp = O.Pipeline(1500, 1100, 3, fmt).thumbnail_shrink_on_load(O.WIDTH, 90).run(src)
        _same(_run(ctx, p, src), exp, np.dtype(dtype).name)
