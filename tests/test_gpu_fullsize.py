"""Every measured workload at its full stated size, checked against the
oracle (integer images bit for bit, float images bit for bit in the
reference's FP64 order and within stated bounds of an FP64 evaluation in the
fast modes):

  * 8192x8192x3 uchar thumbnail Width(400), Lanczos3 (the headline), and the
    same from a PNG source that box-shrinks on load;
  * 4096x4096x4 uchar gauss_blur(2.0), in both the VALU and the int8-limb
    MFMA kernel;
  * 4096x4096x3 float32 Lanczos3 reduce by 2.5 and gauss_blur(8.0), in the
    three float modes;
  * 512 uchar images, W, H ~ U{256..6000}, 3 or 4 bands, FitBox(400, 400),
    one batch, every image checked;
  * the serving loop at 2000 requests/s with requests from the serving
    distribution (P ~ logU[0.1, 24] MP), each result equal to the same
    request run alone and to the oracle.
"""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import float_image, gradient_noise_u8

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[fullsize] {msg}", flush=True)


def _same(got, exp, what):
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    n = int((got != exp).sum())
    assert n == 0, f"{what}: {n} samples differ"


def test_headline_thumbnail_8192x8192x3(ctx):
    _say("thumbnail 8192x8192x3 -> Width(400)")
    src = gradient_noise_u8(8192, 8192, 3, 2024)
    p = v.Pipeline(ctx, 8192, 8192, 3).thumbnail(v.WIDTH, 400)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    exp = O.Pipeline(8192, 8192, 3).thumbnail(O.WIDTH, 400).run(src)
    assert exp.shape == (400, 400, 3)
    _same(out.numpy(), exp, "thumbnail 8192^2")
    # The no-alloc run of the same plan gives the same bytes.
    out2 = p.output_buffer()
    p.run_noalloc(d_src, out2)
    ctx.synchronize()
    _same(out2.numpy(), exp, "thumbnail 8192^2 no-alloc")


@pytest.mark.parametrize("w,h,b", [(5845, 5551, 4), (5846, 5551, 4), (4321, 4355, 3), (2734, 4686, 3)])
def test_thumbnail_unaligned_row_pitch(ctx, w, h, b):
    # Row pitches that are not a multiple of 16 bytes: the single-image box
    # shrink splits each output row's source rows over up to four wavefronts
    # (16 wavefronts per SIMD) and sums their parts in LDS.
    _say(f"thumbnail {w}x{h}x{b} (pitch % 16 = {w * b % 16}) -> FitBox(400, 400)")
    src = gradient_noise_u8(w, h, b, 5000 + w)
    p = v.Pipeline(ctx, w, h, b).thumbnail(v.FITBOX, 400, 400)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    exp = O.Pipeline(w, h, b).thumbnail(O.FITBOX, 400, 400).run(src)
    _same(out.numpy(), exp, f"thumbnail {w}x{h}x{b}")


def test_png_thumbnail_box_shrink_8192x8192x3(ctx):
    # The PNG source: box pre-shrink by the hint (16) on the device, then the
    # thumbnail over the 512 x 512 image.
    _say("PNG thumbnail 8192x8192x3 -> Width(400), box pre-shrink")
    src = gradient_noise_u8(8192, 8192, 3, 77)
    p = v.Pipeline(ctx, 8192, 8192, 3).thumbnail_shrink_on_load(v.WIDTH, 400)
    assert p.stage_kernels()[0] == "box_shrink"
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    exp = O.Pipeline(8192, 8192, 3).thumbnail_shrink_on_load(O.WIDTH, 400).run(src)
    assert exp.shape == (400, 400, 3)
    _same(out.numpy(), exp, "PNG thumbnail 8192^2")


@pytest.mark.parametrize("form", ["valu", "mfma"])
def test_gauss_u8_4096x4096x4(ctx, form):
    _say(f"gauss_blur(2.0) 4096x4096x4 {form}")
    d_src = v.DeviceBuffer(ctx, (4096, 4096, 4), np.uint8)
    d_src.fill_random(4242)
    host = O.synthetic_bytes_fast(4096 * 4096 * 4, 4242).reshape(4096, 4096, 4)
    p = v.Pipeline(ctx, 4096, 4096, 4).gauss_blur(2.0)
    if form == "mfma":
        p.set_option(v.api.OPT_GAUSS_ROWS, -1)
    out = p.output_buffer()
    p.run(d_src, out)
    ctx.synchronize()
    kernels = p.stage_kernels()
    assert kernels == (["gauss_u8"] if form == "valu" else ["sep_mfma_i8"]), kernels
    _same(out.numpy(), O.Pipeline(4096, 4096, 4).gauss_blur(2.0).run(host), f"gauss {form}")


@pytest.mark.parametrize("what", ["reduce", "gauss"])
def test_float_4096x4096x3_modes(ctx, what):
    _say(f"float32 4096x4096x3 {what}")
    src = float_image(4096, 4096, 3, 31337)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    if what == "reduce":
        build = lambda p: p.reduce_h(2.5, v.LANCZOS3).reduce_v(2.5, v.LANCZOS3)
        exact = O.Pipeline(4096, 4096, 3, O.F32).reduce_h(2.5).reduce_v(2.5).run(src)
        ref = O.f64_reduce_hv(src, 2.5, O.LANCZOS3)
    else:
        build = lambda p: p.gauss_blur(8.0)
        exact = O.Pipeline(4096, 4096, 3, O.F32).gauss_blur(8.0).run(src)
        ref = O.f64_gauss(src, 8.0)
    for mode in (0, v.Pipeline.FLOAT_FP32, v.Pipeline.FLOAT_FP16X3):
        p = build(v.Pipeline(ctx, 4096, 4096, 3, v.F32).set_float_mode(mode))
        out = p.output_buffer()
        p.run(d_src, out)
        ctx.synchronize()
        got = out.numpy()
        if mode == 0:
            _same(got, exact, f"{what} reference order")
        err = np.abs(got.astype(np.float64) - ref)
        rel = float((err / np.maximum(np.abs(ref), 1e-2)).max())
        # max |err| <= 2e-6 on values in [-0.05, 1.05]; max rel <= 1e-6.
        assert float(err.max()) <= (2.5e-7 if mode == 0 else 2e-6), (mode, float(err.max()))
        assert rel <= 1e-6, (mode, rel)
        del p, out, got, err


def test_batch_512_uneven_every_image(ctx):
    _say("batch 512 uneven FitBox(400, 400)")
    r = np.random.default_rng(4004)
    sizes = [(int(r.integers(256, 6001)), int(r.integers(256, 6001)), int(r.choice([3, 4]))) for _ in range(512)]
    srcs = []
    for i, (w, h, b) in enumerate(sizes):
        d = v.DeviceBuffer(ctx, (h, w, b), np.uint8)
        d.fill_random(7_000_000 + i)
        srcs.append(d)
    batch = v.Batch(ctx, sizes, v.FITBOX, 400, 400, v.LANCZOS3)
    dsts = batch.output_buffers()
    batch.run(srcs, dsts)
    ctx.synchronize()
    bad = []
    for i, (w, h, b) in enumerate(sizes):
        host = O.synthetic_bytes_fast(w * h * b, 7_000_000 + i).reshape(h, w, b)
        exp = O.Pipeline(w, h, b).thumbnail(O.FITBOX, 400, 400).run(host)
        if not np.array_equal(dsts[i].numpy(), exp):
            bad.append((i, sizes[i]))
        if i % 128 == 127:
            _say(f"batch: {i + 1} images checked")
    assert not bad, bad[:5]


def test_serving_2000_per_s_alone_vs_mixed(ctx):
    _say("serving 2000 req/s, serving distribution")
    g = np.random.default_rng(2000)
    region = v.DeviceBuffer(ctx, (128 << 20,), np.uint8)
    region.fill_random(9_000_000)
    ctx.synchronize()
    host_region = O.synthetic_bytes_fast(128 << 20, 9_000_000)
    base = region.data_ptr()
    n = 120
    arrivals = (np.cumsum(g.exponential(1.0 / 2000.0, size=n)) * 1e9).astype(np.int64)
    specs = []
    for _ in range(n):
        P = np.exp(g.uniform(np.log(0.1e6), np.log(24e6)))
        a = g.uniform(0.5, 2.0)
        w, h = int(round(np.sqrt(P * a))), int(round(np.sqrt(P / a)))
        off = int(g.integers(0, ((128 << 20) - w * h * 3) // 256 + 1)) * 256
        specs.append((w, h, off, (256, 400, 800)[int(g.integers(0, 3))]))
    reqs = [v.Server.request(base + off, w, h, 3, on_device=True, target=v.WIDTH, target_w=tw, blur_sigma=0.8)
            for (w, h, off, tw) in specs]
    srv = v.Server(ctx, hbm_budget=4 << 30, max_inflight=64, streams=4, planner_threads=4)
    # Mixed: the open-loop trace, requests overlapping on the device.
    ids = []
    import time
    t0 = time.perf_counter_ns()
    for req, at in zip(reqs, arrivals):
        while time.perf_counter_ns() - t0 < at:
            pass
        ids.append(srv.submit(req))
    mixed = [srv.fetch(i) for i in ids]
    assert srv.stats()["peak_inflight"] >= 1
    for i in ids:
        srv.release(i)
    # Alone: each request by itself, and the oracle on a sample.
    for k, (req, (w, h, off, tw)) in enumerate(zip(reqs, specs)):
        rid = srv.submit(req)
        alone = srv.fetch(rid)
        srv.release(rid)
        _same(mixed[k], alone, f"request {k} mixed vs alone")
        if k % 12 == 0:
            src = host_region[off:off + w * h * 3].reshape(h, w, 3)
            exp = O.Pipeline(w, h, 3).thumbnail(O.WIDTH, tw).gauss_blur(0.8).run(src)
            _same(alone, exp, f"request {k} vs oracle")
    srv.close()
