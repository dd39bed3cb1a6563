"""The I16, U32, I32 and F64 band formats (band.rs:19-35) through every node of
the path on the device, bit for bit against the oracle: ReduceH/V
(fixed-point i16 taps with the signed from_fixed_i64 rounding, the FMA
chain for F64), ReduceOp, ShrinkH/V and Shrink (from_f64_clamped), the
Gaussian and ConvSep (to F32), thumbnail and resize plans (Affine and Zoom
tails), the PNG box pre-shrink, and the op-level boundary."""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FORMATS = [(v.I16, np.int16), (v.U32, np.uint32), (v.I32, np.int32), (v.F64, np.float64)]
IDS = ["i16", "u32", "i32", "f64"]


def _image(dtype, shape, seed):
    r = np.random.default_rng(seed)
    if dtype == np.float64:
        return r.normal(0.0, 1000.0, size=shape) * np.exp2(r.integers(-12, 12, size=shape))
    info = np.iinfo(dtype)
    return r.integers(int(info.min), int(info.max) + 1, size=shape, dtype=np.int64).astype(dtype)


def _run(ctx, p, src):
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    return out.numpy()


def _same(got, exp, what):
    assert got.shape == exp.shape and got.dtype == exp.dtype, (what, got.shape, exp.shape, got.dtype, exp.dtype)
    bad = np.argwhere(~((got == exp) | (np.isnan(got) & np.isnan(exp)) if got.dtype.kind == "f" else got == exp))
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}"


def _both(ctx, w, h, b, fmt, src, build):
    """The same builder calls on the device pipeline and the oracle's."""
    got = _run(ctx, build(v.Pipeline(ctx, w, h, b, fmt)), src)
    exp = build(O.Pipeline(w, h, b, fmt)).run(src)
    return got, exp


@pytest.mark.parametrize("fmt,dtype", FORMATS, ids=IDS)
@pytest.mark.parametrize("bands", [1, 3])
def test_reduce_h_v(ctx, fmt, dtype, bands):
    h, w = 37, 53
    src = _image(dtype, (h, w, bands), fmt * 10 + bands)
    for k, f in ((v.LANCZOS3, 2.5), (v.BICUBIC, 1.7), (v.BILINEAR, 3.0)):
        got, exp = _both(ctx, w, h, bands, fmt, src, lambda p: p.reduce_h(f, k).reduce_v(f, k))
        _same(got, exp, f"reduce_hv {k} {f}")
        got, exp = _both(ctx, w, h, bands, fmt, src, lambda p: p.reduce_v(f, k))
        _same(got, exp, f"reduce_v {k} {f}")


@pytest.mark.parametrize("fmt,dtype", FORMATS, ids=IDS)
def test_reduce_op(ctx, fmt, dtype):
    src = _image(dtype, (29, 41, 3), fmt)
    got, exp = _both(ctx, 41, 29, 3, fmt, src, lambda p: p.reduce(2.3, 1.6, v.LANCZOS3))
    _same(got, exp, "reduce")


@pytest.mark.parametrize("fmt,dtype", FORMATS, ids=IDS)
@pytest.mark.parametrize("f", [2, 3, 5])
def test_shrink(ctx, fmt, dtype, f):
    src = _image(dtype, (8 * f + 1, 9 * f + 2, 3), fmt + f)
    h, w = src.shape[:2]
    got, exp = _both(ctx, w, h, 3, fmt, src, lambda p: p.shrink_h(f).shrink_v(f))
    _same(got, exp, "shrink_h, shrink_v")
    got, exp = _both(ctx, w, h, 3, fmt, src, lambda p: p.shrink(f, f + 1))
    _same(got, exp, "shrink")


@pytest.mark.parametrize("fmt,dtype", FORMATS, ids=IDS)
def test_gauss_and_convsep_to_f32(ctx, fmt, dtype):
    src = _image(dtype, (33, 47, 3), 3 * fmt)
    for sigma in (0.8, 2.5):
        got, exp = _both(ctx, 47, 33, 3, fmt, src, lambda p: p.gauss_blur(sigma))
        assert got.dtype == np.float32
        _same(got, exp, f"gauss {sigma}")
    coeffs = np.array([0.1, 0.3, 0.35, 0.3, 0.1])
    got, exp = _both(ctx, 47, 33, 3, fmt, src, lambda p: p.convsep(coeffs, 1.15, -2.0))
    _same(got, exp, "convsep")


@pytest.mark.parametrize("fmt,dtype", FORMATS, ids=IDS)
def test_thumbnail_resize_and_box(ctx, fmt, dtype):
    src = _image(dtype, (301, 413, 3), 7 * fmt)
    for tw in (100, 37):
        got, exp = _both(ctx, 413, 301, 3, fmt, src, lambda p: p.thumbnail(v.WIDTH, tw))
        _same(got, exp, f"thumbnail {tw}")
    for hs, vs in ((0.37, 0.5), (2.5, 1.7)):
        got, exp = _both(ctx, 413, 301, 3, fmt, src, lambda p: p.resize(hs, vs, v.LANCZOS3))
        _same(got, exp, f"resize {hs} {vs}")
    got, exp = _both(ctx, 413, 301, 3, fmt, src, lambda p: p.box_shrink(4))
    _same(got, exp, "box_shrink")
    got, exp = _both(ctx, 413, 301, 3, fmt, src, lambda p: p.thumbnail_shrink_on_load(v.WIDTH, 60))
    _same(got, exp, "thumbnail_shrink_on_load")


@pytest.mark.parametrize("fmt,dtype", FORMATS, ids=IDS)
def test_op_level(ctx, fmt, dtype):
    # DynOperation::dyn_process_region on host tiles, whole-image regions.
    src = _image(dtype, (21, 34, 2), 11 * fmt)
    op = v.Op.reduce_h(ctx, 2.2, v.BICUBIC, 2, 34, fmt)
    ow, oh = op.output_size(34, 21)
    r = v.Region(0, 0, ow, oh)
    in_r = op.required_input_region(r)
    ys = np.clip(np.arange(in_r.y, in_r.y + in_r.height), 0, 20)
    xs = np.clip(np.arange(in_r.x, in_r.x + in_r.width), 0, 33)
    tile = np.ascontiguousarray(src[ys][:, xs, :])
    dst = np.zeros((oh, ow, 2), dtype)
    op.process_region_host(tile, dst, in_r, r)
    _same(dst, O.Pipeline(34, 21, 2, fmt).reduce_h(2.2, O.BICUBIC).run(src), "op reduce_h")
