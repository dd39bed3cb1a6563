"""Operation-level boundary (DynOperation::dyn_process_region,
crates/viprs-core/src/op/dynamic_op.rs:222-229) driven the way the reference's
tile scheduler drives it: for each output strip, the input region comes from
required_input_region, the source tile is read with clamp-to-edge
(sources/memory.rs:149-235), and chained ops see each other's tiles over the
regions the backward walk asks for (execute.rs:413-480).  Host tiles go
through vgpu_op_process_region_host, device tiles through
vgpu_op_process_region; every strip must equal the oracle's whole-image run.
Also: cancellation of pipeline runs and server requests."""
import threading

import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import float_image, noise_u8, noise_u16

pytestmark = pytest.mark.gpu
FMT = {np.dtype(np.uint8): v.U8, np.dtype(np.uint16): v.U16, np.dtype(np.float32): v.F32}
NP = {v.U8: np.uint8, v.U16: np.uint16, v.F32: np.float32}


def clamp_read(src, r):
    h, w, _ = src.shape
    ys = np.clip(np.arange(r.y, r.y + r.height), 0, h - 1)
    xs = np.clip(np.arange(r.x, r.x + r.width), 0, w - 1)
    return np.ascontiguousarray(src[ys][:, xs, :])


def run_chain_tiled(ops, src, out_w, out_h, strip, host=True, ctx=None):
    """The reference scheduler's per-tile walk over a chain of ops: backward
    regions from each output strip, a clamp-to-edge source read, then each
    op's process_region forward."""
    b = src.shape[2]
    out = np.zeros((out_h, out_w, b), NP[ops[-1].out_fmt])
    for y0 in range(0, out_h, strip):
        regions = [v.Region(0, y0, out_w, min(strip, out_h - y0))]
        for op in reversed(ops):
            regions.insert(0, op.required_input_region(regions[0]))
        tile = clamp_read(src, regions[0])
        for k, op in enumerate(ops):
            o = regions[k + 1]
            dst = np.zeros((o.height, o.width, b), NP[op.out_fmt])
            if host:
                op.process_region_host(tile, dst, regions[k], o)
            else:
                d_dst = v.DeviceBuffer(ctx, dst.shape, dst.dtype)
                op.process_region(v.DeviceBuffer.from_numpy(ctx, tile), d_dst, regions[k], o)
                ctx.synchronize()
                dst = d_dst.numpy()
            tile = dst
        out[y0:y0 + tile.shape[0]] = tile
    return out


def _assert_same(got, exp, what):
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}"


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("host", [True, False])
def test_op_shrink_h_and_v_strips(ctx, dtype, host):
    src = {np.uint8: noise_u8, np.uint16: noise_u16}.get(dtype, lambda h, w, b, s: float_image(h, w, b, s))(
        61, 97, 3, 5)
    fmt = FMT[np.dtype(dtype)]
    for f in (2, 3, 5):
        ops = [v.Op.shrink_h(ctx, f, 3, fmt=fmt), v.Op.shrink_v(ctx, f, 3, fmt=fmt)]
        exp = O.Pipeline(97, 61, 3, fmt).shrink_h(f).shrink_v(f).run(src)
        got = run_chain_tiled(ops, src, 97 // f, 61 // f, strip=4, host=host, ctx=ctx)
        _assert_same(got, exp, f"shrink {f} host={host}")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_op_reduce_v_strips(ctx, dtype):
    # ReduceV's x86 plan is direct per row, so strips equal the whole image.
    src = {np.uint8: noise_u8, np.uint16: noise_u16}.get(dtype, lambda h, w, b, s: float_image(h, w, b, s))(
        173, 41, 2, 9)
    fmt = FMT[np.dtype(dtype)]
    for f, k in ((2.3, v.LANCZOS3), (1.5, v.BICUBIC), (4.0, v.BILINEAR)):
        op = v.Op.reduce_v(ctx, f, k, 2, 173, fmt)
        ow, oh = op.output_size(41, 173)
        exp = O.Pipeline(41, 173, 2, fmt).reduce_v(f, k).run(src)
        for strip in (16, 5):
            got = run_chain_tiled([op], src, ow, oh, strip=strip)
            _assert_same(got, exp, f"reduce_v {f} strip {strip}")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("bands", [1, 3, 4])
def test_op_gauss_blur_h_then_v_strips(ctx, dtype, bands):
    # GaussBlurH -> GaussBlurV as two bridges (the reference's own
    # parity-test shape, parity/convolution.rs:48-90): U8 stays U8, other
    # formats go to F32 after the H pass.
    src = {np.uint8: noise_u8, np.uint16: noise_u16}.get(dtype, lambda h, w, b, s: float_image(h, w, b, s))(
        57, 83, bands, 3 + bands)
    fmt = FMT[np.dtype(dtype)]
    for sigma in (0.8, 2.0):
        gh = v.Op.gauss_blur_h(ctx, sigma, bands, fmt)
        gv = v.Op.gauss_blur_v(ctx, sigma, bands, gh.out_fmt)
        exp = O.Pipeline(83, 57, bands, fmt).gauss_blur(sigma).run(src)
        for strip in (16, 7):
            got = run_chain_tiled([gh, gv], src, 83, 57, strip=strip)
            _assert_same(got, exp, f"gauss {sigma} strip {strip}")


def test_op_reduce_h_tiles_swap_plans(ctx):
    # A region-dependent plan (u8 ReduceH accumulates from the tile origin)
    # re-uploads per region without draining the stream; alternating regions
    # and repeating one must keep every result right.
    src = noise_u8(12, 500, 3, 44)
    op = v.Op.reduce_h(ctx, 2.7, v.LANCZOS3, 3, 500)
    regions = [v.Region(0, 0, 90, 12), v.Region(50, 0, 80, 12), v.Region(0, 0, 90, 12), v.Region(-3, 0, 40, 12)]
    outs = []
    for r in regions * 3:
        in_r = op.required_input_region(r)
        tile = clamp_read(src, in_r)
        d_dst = v.DeviceBuffer(ctx, (r.height, r.width, 3), np.uint8)
        op.process_region(v.DeviceBuffer.from_numpy(ctx, tile), d_dst, in_r, r)
        outs.append(d_dst)
    ctx.synchronize()
    for i, r in enumerate(regions * 3):
        in_r = op.required_input_region(r)
        tile = clamp_read(src, in_r)
        ref = np.zeros((r.height, r.width, 3), np.uint8)
        op.process_region_host(tile, ref, in_r, r)
        _assert_same(outs[i].numpy(), ref, f"region {i}")
    whole = O.Pipeline(500, 12, 3).reduce_h(2.7).run(src)
    _assert_same(outs[0].numpy(), whole[:, :90], "origin tile vs oracle")


def test_op_host_entry_from_threads(ctx):
    # The unchanged RayonScheduler calls one op from many workers at once.
    src = noise_u8(96, 128, 3, 8)
    gh = v.Op.gauss_blur_h(ctx, 1.5, 3)
    gv = v.Op.gauss_blur_v(ctx, 1.5, 3)
    exp = O.Pipeline(128, 96, 3).gauss_blur(1.5).run(src)
    results, errors = {}, []

    def worker(i):
        try:
            results[i] = run_chain_tiled([gh, gv], src, 128, 96, strip=8)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(4):
        _assert_same(results[i], exp, f"thread {i}")


def test_op_device_entry_rejects_host_pointers(ctx):
    op = v.Op.gauss_blur_h(ctx, 1.0, 1)
    r = v.Region(0, 0, 8, 4)
    in_r = op.required_input_region(r)
    tile = np.zeros((in_r.height, in_r.width, 1), np.uint8)
    out = np.zeros((4, 8, 1), np.uint8)
    import ctypes as C
    rc = op._lib.vgpu_op_process_region(op._h, C.c_void_p(tile.ctypes.data), tile.nbytes,
                                        C.c_void_p(out.ctypes.data), out.nbytes, in_r, r)
    assert rc == v.api.E_INVALID_ARGUMENT
    op.process_region_host(tile, out, in_r, r)  # the host entry takes them


def test_pipeline_run_cancellable(ctx):
    src = noise_u8(300, 400, 3, 2)
    p = v.Pipeline(ctx, 400, 300, 3).thumbnail(v.WIDTH, 100).gauss_blur(1.0)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    tok = v.CancelToken()
    out = p.output_buffer()
    p.run_cancellable(d_src, out, tok)
    ctx.synchronize()
    exp = O.Pipeline(400, 300, 3).thumbnail(O.WIDTH, 100).gauss_blur(1.0).run(src)
    _assert_same(out.numpy(), exp, "uncancelled run")
    tok.cancel()
    assert tok.is_cancelled()
    with pytest.raises(v.VgpuError) as e:
        p.run_cancellable(d_src, out, tok)
    assert e.value.status == v.api.E_CANCELLED


def test_server_cancel(ctx):
    s = v.Server(ctx, hbm_budget=256 << 20, max_inflight=1, streams=1, planner_threads=1)
    src = noise_u8(1500, 2000, 3, 6)
    d_src = v.DeviceBuffer.from_numpy(ctx, src)
    req = v.Server.request(d_src, 2000, 1500, 3, target=v.WIDTH, target_w=200, blur_sigma=0.8)
    ids = [s.submit(req) for _ in range(6)]
    s.cancel(ids[-1])
    results = [s.wait(i) for i in ids]
    assert results[-1]["status"] == v.api.E_CANCELLED
    exp = O.Pipeline(2000, 1500, 3).thumbnail(O.WIDTH, 200).gauss_blur(0.8).run(src)
    assert results[0]["status"] == v.api.OK
    assert np.array_equal(s.fetch(ids[0]), exp)
    with pytest.raises(v.VgpuError):
        s.cancel(ids[0])  # already finished
    for i in ids:
        s.release(i)
