"""vgpu_pipeline_run_noalloc: the pipeline run without device allocation, its
tables and intermediates in the library's fixed scratch array.  Its output
must equal vgpu_pipeline_run's (and the oracle's) bit for bit, on two calls in
a row, the second with a smaller input, and a plan that does not fit must be
rejected before anything is enqueued."""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import float_image, noise_u8

pytestmark = pytest.mark.gpu


def _both(ctx, src, steps):
    h, w, b = src.shape
    fmt = {np.dtype(np.uint8): v.U8, np.dtype(np.float32): v.F32}[src.dtype]
    p = v.Pipeline(ctx, w, h, b, fmt)
    for name, args in steps:
        getattr(p, name)(*args)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    ref, got = p.output_buffer(), p.output_buffer()
    p.run(d_src, ref)
    p.run_noalloc(d_src, got)
    ctx.synchronize()
    return p, got.numpy(), ref.numpy()


def test_noalloc_matches_run_twice_second_smaller(ctx):
    cases = [
        (noise_u8(777, 1031, 3, 1), [("thumbnail", (v.WIDTH, 200, 0, v.LANCZOS3)), ("gauss_blur", (0.8,))]),
        (noise_u8(203, 301, 4, 2), [("thumbnail", (v.WIDTH, 64, 0, v.LANCZOS3)), ("gauss_blur", (1.5,))]),
    ]
    for src, steps in cases:
        p, got, ref = _both(ctx, src, steps)
        assert p.stage_count() >= 2  # at least one intermediate lives in the scratch array
        assert np.array_equal(got, ref)
        h, w, b = src.shape
        op = O.Pipeline(w, h, b)
        for name, args in steps:
            getattr(op, name)(*args)
        assert np.array_equal(got, op.run(src))


def test_noalloc_float_and_repeat_calls(ctx):
    src = float_image(131, 157, 3, 5)
    steps = [("reduce_h", (2.5, v.LANCZOS3)), ("gauss_blur", (2.0,)), ("reduce_v", (1.7, v.BICUBIC))]
    p, got, ref = _both(ctx, src, steps)
    assert np.array_equal(got, ref)
    # A second call on the same pipeline after another pipeline used the array.
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    again = p.output_buffer()
    p.run_noalloc(d_src, again)
    ctx.synchronize()
    assert np.array_equal(again.numpy(), ref)


def test_noalloc_rejects_plans_beyond_the_scratch_array(ctx):
    # A Gaussian whose 9000 x 8000 intermediate (72 MB) feeds a ReduceH.
    w, h = 9000, 8000
    assert w * h > v.NOALLOC_SCRATCH_BYTES
    p = v.Pipeline(ctx, w, h, 1).gauss_blur(2.0).reduce_h(1.5, v.BILINEAR)
    d_src = v.DeviceBuffer(ctx, (h, w, 1), np.uint8)
    d_dst = p.output_buffer()
    with pytest.raises(v.VgpuError) as e:
        p.run_noalloc(d_src, d_dst)
    assert e.value.status == v.E_SCRATCH_LIMIT
    # The context stays usable: a fitting run right after succeeds.
    src = noise_u8(64, 80, 3, 3)
    _, got, ref = _both(ctx, src, [("gauss_blur", (1.0,)), ("reduce_h", (1.5, v.BILINEAR))])
    assert np.array_equal(got, ref)
