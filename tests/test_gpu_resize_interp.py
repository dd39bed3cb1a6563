"""Resize with the LBB and VSQBS interpolators on the device (resize.rs:
176-284; affine/core.rs:588-610,699-725; affine/interpolation.rs:170-441;
kernel.rs:863-923), bit for bit against the oracle: LBB as one Affine for
down- and upscales, VSQBS upscales without the Zoom, four-band images
through the premultiplied stencil, every sample format."""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NP = {v.U8: np.uint8, v.U16: np.uint16, v.F32: np.float32, v.I16: np.int16, v.U32: np.uint32,
      v.I32: np.int32, v.F64: np.float64}


def _image(fmt, shape, seed):
    r = np.random.default_rng(seed)
    dt = np.dtype(NP[fmt])
    if dt.kind == "f":
        return (r.normal(0.5, 0.3, size=shape) * 100).astype(dt)
    info = np.iinfo(dt)
    return r.integers(int(info.min), int(info.max) + 1, size=shape, dtype=np.int64).astype(dt)


def _resize(ctx, src, fmt, hs, vs, kernel):
    h, w, b = src.shape
    p = v.Pipeline(ctx, w, h, b, fmt).resize(hs, vs, kernel)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    exp = O.Pipeline(w, h, b, fmt).resize(hs, vs, kernel).run(src)
    return out.numpy(), exp, p


def _same(got, exp, what):
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}: {got[tuple(bad[0])]} vs " \
                          f"{exp[tuple(bad[0])]}"


def test_vsqbs_golden_on_device(ctx):
    src = np.array([[0, 64], [128, 255]], np.uint8)[..., None]
    got, exp, _ = _resize(ctx, src, v.U8, 2.0, 2.0, v.VSQBS)
    assert list(got.ravel()) == [57, 96, 115, 118, 124, 169, 198, 201, 159, 214, 245, 249, 164, 219, 251, 255]
    _same(got, exp, "golden")


@pytest.mark.parametrize("bands", [1, 3, 4, 5])
@pytest.mark.parametrize("kernel,hs,vs", [(v.LBB, 0.6, 0.6), (v.LBB, 2.0, 2.0), (v.LBB, 1.37, 0.71),
                                          (v.VSQBS, 2.0, 2.0), (v.VSQBS, 3.3, 1.5), (v.VSQBS, 1.0, 2.7)])
def test_resize_u8(ctx, bands, kernel, hs, vs):
    src = _image(v.U8, (67, 91, bands), bands * 7 + int(hs * 10))
    got, exp, p = _resize(ctx, src, v.U8, hs, vs, kernel)
    _same(got, exp, f"resize u8 b={bands} k={kernel} {hs}x{vs}")
    assert "affine" in p.stage_kernels()[-1]


@pytest.mark.parametrize("fmt", [v.U16, v.F32, v.I16, v.U32, v.I32, v.F64])
@pytest.mark.parametrize("kernel", [v.LBB, v.VSQBS])
def test_resize_formats(ctx, fmt, kernel):
    src = _image(fmt, (41, 57, 4 if fmt in (v.U16, v.F32) else 3), fmt + kernel)
    got, exp, _ = _resize(ctx, src, fmt, 1.8, 2.3, kernel)
    _same(got, exp, f"resize fmt={fmt} k={kernel}")
    if kernel == v.LBB:
        got, exp, _ = _resize(ctx, src, fmt, 0.45, 0.8, kernel)
        _same(got, exp, f"resize down fmt={fmt}")


def test_vsqbs_downscale_refused(ctx):
    p = v.Pipeline(ctx, 100, 100, 3)
    with pytest.raises(v.VgpuError) as e:
        p.resize(0.5, 0.5, v.VSQBS)
    assert e.value.status == v.api.E_INVALID_KERNEL
    # the pipeline is left as it was
    assert p.output_info()[:2] == (100, 100)


def test_lbb_large_downscale(ctx):
    # one Affine over a 1500 x 1100 source to 317 x 233
    src = _image(v.U8, (1100, 1500, 3), 99)
    got, exp, _ = _resize(ctx, src, v.U8, 317 / 1500, 233 / 1100, v.LBB)
    _same(got, exp, "lbb large")
