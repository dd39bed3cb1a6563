"""Fast float paths for f32 images against the FP64 oracle.

The reference accumulates f32 images in FP64 (our default mode reproduces it
bit for bit, see test_gpu_parity).  The fast modes compute in plain FP32
(mode 1) or with an FP16 hi/lo split on MFMA and FP32 accumulation (mode 2);
they must stay as accurate against an FP64 evaluation of the same separable
math as stated here:
  - max |err| <= 2e-6 for pixel values in [-0.05, 1.05] (about 16 ulp of 1.0),
  - max relative error <= 1e-6 * (sum |w| over the taps of both passes),
    relative to max(|ref|, 1e-2),
and the fast modes' errors are compared with the reference-order error."""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import float_image

pytestmark = pytest.mark.gpu

MODES = [v.Pipeline.FLOAT_FP32, v.Pipeline.FLOAT_FP16X3]


KERNEL = {0: "reduce_hv", v.Pipeline.FLOAT_FP32: "reduce_f32", v.Pipeline.FLOAT_FP16X3: "sep_mfma_f16x3"}


def errors(got, ref):
    err = np.abs(got.astype(np.float64) - ref)
    rel = err / np.maximum(np.abs(ref), 1e-2)
    return float(err.max()), float(rel.max())


def where(got, ref):
    """First worst position and count of large errors, for failure messages."""
    err = np.abs(got.astype(np.float64) - ref)
    return np.unravel_index(int(np.argmax(err)), err.shape), int((err > 1e-4).sum())


def run_mode(ctx, src, steps, mode, kernel=None):
    h, w, b = src.shape
    p = v.Pipeline(ctx, w, h, b, v.F32).set_float_mode(mode)
    for name, args in steps:
        getattr(p, name)(*args)
    if kernel is not None:
        assert p.stage_kernels() == [kernel], (mode, p.stage_kernels())
    return p.run_host(src)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("bands", [1, 3, 4])
def test_float_reduce_against_fp64(ctx, mode, bands):
    src = float_image(331, 257, bands, 7 + bands)
    ref = O.f64_reduce_hv(src, 2.5, O.LANCZOS3)
    steps = [("reduce_h", (2.5, v.LANCZOS3)), ("reduce_v", (2.5, v.LANCZOS3))]
    exact = run_mode(ctx, src, steps, 0)
    fast = run_mode(ctx, src, steps, mode, KERNEL[mode])
    assert fast.shape == ref.shape
    e_abs, e_rel = errors(fast, ref)
    x_abs, _ = errors(exact, ref)
    assert e_abs <= 2e-6, (e_abs, x_abs, where(fast, ref))
    assert e_rel <= 1e-6 * 1.6 ** 2 * 2, (e_rel, x_abs)
    # The reference's order (mode 0) rounds to f32 twice: <= ~2 half-ulps of 1.0.
    assert x_abs <= 2.5e-7, x_abs


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("sigma", [0.8, 2.0, 8.0])
def test_float_gauss_against_fp64(ctx, mode, sigma):
    src = float_image(211, 307, 3, 40)
    ref = O.f64_gauss(src, sigma)
    kernel = "conv_pair_f32" if mode == v.Pipeline.FLOAT_FP32 else "sep_mfma_f16x3"
    fast = run_mode(ctx, src, [("gauss_blur", (sigma,))], mode, kernel)
    exact = run_mode(ctx, src, [("gauss_blur", (sigma,))], 0)
    e_abs, e_rel = errors(fast, ref)
    x_abs, _ = errors(exact, ref)
    assert e_abs <= 2e-6, (e_abs, x_abs, where(fast, ref))
    assert e_rel <= 2e-6, (e_rel, x_abs)
    assert x_abs <= 2.5e-7, x_abs


def test_float_mode_keeps_integer_images_exact(ctx):
    # The float modes apply to f32 images only; u8 results stay bit-exact.
    src = (float_image(97, 131, 3, 1) * 200).clip(0, 255).astype(np.uint8)
    p = v.Pipeline(ctx, 131, 97, 3).set_float_mode(v.Pipeline.FLOAT_FP32).reduce_h(1.7).reduce_v(1.7)
    exp = O.Pipeline(131, 97, 3).reduce_h(1.7).reduce_v(1.7).run(src)
    assert np.array_equal(p.run_host(src), exp)


@pytest.mark.parametrize("what", ["reduce", "gauss"])
def test_fp16x3_no_less_accurate_than_fp32(ctx, what):
    # The FP16 hi/lo MFMA path stays within a max relative error of 1e-6
    # against FP64, and within 10 % of the plain FP32 kernel's errors.
    src = float_image(256, 256, 3, 99)
    if what == "reduce":
        ref = O.f64_reduce_hv(src, 2.5, O.LANCZOS3)
        steps = [("reduce_h", (2.5, v.LANCZOS3)), ("reduce_v", (2.5, v.LANCZOS3))]
    else:
        ref = O.f64_gauss(src, 8.0)
        steps = [("gauss_blur", (8.0,))]
    e32 = errors(run_mode(ctx, src, steps, v.Pipeline.FLOAT_FP32), ref)
    e16 = errors(run_mode(ctx, src, steps, v.Pipeline.FLOAT_FP16X3), ref)
    assert e16[1] <= 1e-6, e16
    assert e16[0] <= 1.1 * e32[0] and e16[1] <= 1.1 * e32[1], (e16, e32)
