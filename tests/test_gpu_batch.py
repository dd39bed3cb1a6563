"""Batched thumbnails through the C ABI: every image of a batch must equal the
oracle's thumbnail of that image alone, bit for bit, whatever the launch shape
(one block per task or a persistent grid walking the task list, graph on or
off, any reduce tile), on repeated graph replays and after the caller's
buffers change."""
import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import noise_u8, splitmix_bytes

pytestmark = pytest.mark.gpu


def random_sizes(n, seed, lo=20, hi=900):
    r = np.random.default_rng(seed)
    return [(int(r.integers(lo, hi)), int(r.integers(lo, hi)), int(r.integers(1, 5))) for _ in range(n)]


def sources(sizes, seed):
    out = []
    for i, (w, h, b) in enumerate(sizes):
        s = noise_u8(h, w, b, seed * 1000 + i)
        if b in (2, 4):  # some fully transparent pixels
            s[..., -1] = np.where(np.random.default_rng(i).random((h, w)) < 0.1, 0, s[..., -1])
        out.append(s)
    return out


def expected(srcs, target, tw, th, kernel=O.LANCZOS3):
    exp = []
    for s in srcs:
        h, w, b = s.shape
        exp.append(O.Pipeline(w, h, b).thumbnail(target, tw, th, kernel).run(s))
    return exp


def check_outputs(dsts, exp, what):
    for i, (d, e) in enumerate(zip(dsts, exp)):
        got = d.numpy()
        assert got.shape == e.shape, (what, i, got.shape, e.shape)
        if not np.array_equal(got, e):
            bad = np.argwhere(got != e)
            raise AssertionError(f"{what}: image {i} {e.shape}: {len(bad)} mismatches, first {bad[0]}")


@pytest.mark.parametrize("target,tw,th", [(v.FITBOX, 150, 150), (v.WIDTH, 97, 0), (v.HEIGHT, 0, 64)])
def test_batch_matches_oracle(ctx, target, tw, th):
    sizes = random_sizes(24, 7 + target)
    srcs = sources(sizes, target)
    exp = expected(srcs, target, tw, th)
    batch = v.Batch(ctx, sizes, target, tw, th)
    info = batch.info()
    # Only empty plans (images already inside the target) leave the batched launches.
    empty = sum(1 for (w, h, b) in sizes if not O.thumbnail_plan(w, h, b, target, tw, th)[0])
    assert info["images"] == 24 and info["fallback_images"] == empty, info
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    dst = batch.output_buffers()
    for i, d in enumerate(dst):
        assert d.shape == exp[i].shape
    batch.run(dsrc, dst)
    ctx.synchronize()
    check_outputs(dst, exp, "first run")


def test_batch_launch_shapes_and_replays(ctx):
    sizes = random_sizes(16, 99, 40, 1300) + [(64, 48, 3), (150, 150, 4), (2, 900, 1), (900, 3, 2)]
    srcs = sources(sizes, 5)
    exp = expected(srcs, v.FITBOX, 120, 120)
    batch = v.Batch(ctx, sizes, v.FITBOX, 120, 120)
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    for grid_cap, graph, tile, rgrid in [(0, True, 0, None), (0, False, 0, None), (64, True, 0, None),
                                         (300, True, 2, None), (0, True, 3, None), (7, False, 1, None),
                                         (0, True, 0, 5), (9, True, 0, 0)]:
        batch.configure(grid_cap, graph, tile, rgrid)
        dst = batch.output_buffers()
        for rep in range(3):  # graph replays, host work and syncs in between
            for d in dst:
                d.upload(np.zeros(d.shape, np.uint8))
            batch.run(dsrc, dst)
            ctx.synchronize()
            check_outputs(dst, exp, f"grid {grid_cap}/{rgrid} graph {graph} tile {tile} replay {rep}")


@pytest.mark.parametrize("reduce_chain", [2, 1, 0])
@pytest.mark.parametrize("form", [0, 1])
def test_batch_chain_layouts_keep_results(ctx, reduce_chain, form):
    # Reduce stages on a chain of their own (each launch after the shrink
    # launch that wrote its inputs) or on their images' chains, on either
    # form of the MFMA pair: the layout moves work between streams only.
    sizes = random_sizes(20, 321, 200, 2600) + [(64, 48, 3), (150, 150, 4), (300, 2000, 2)]
    srcs = sources(sizes, 17)
    exp = expected(srcs, v.FITBOX, 160, 160)
    batch = v.Batch(ctx, sizes, v.FITBOX, 160, 160)
    batch.set_option(v.api.OPT_BATCH_REDUCE_CHAIN, reduce_chain)
    batch.set_option(v.api.OPT_SEP_FORM, form)
    assert batch.info()["reduce_chain"] == reduce_chain
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    dst = batch.output_buffers()
    for rep in range(3):
        for d in dst:
            d.upload(np.zeros(d.shape, np.uint8))
        batch.run(dsrc, dst)
        ctx.synchronize()
        check_outputs(dst, exp, f"reduce chain {reduce_chain} form {form} replay {rep}")
    with pytest.raises(v.VgpuError):
        batch.set_option(v.api.OPT_BATCH_REDUCE_CHAIN, 3)


def test_batch_outputs_equal_single_pipelines(ctx):
    sizes = [(1031, 777, 3), (640, 1000, 4), (300, 200, 1), (1500, 97, 2)]
    srcs = sources(sizes, 11)
    batch = v.Batch(ctx, sizes, v.WIDTH, 200, 0)
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    dst = batch.output_buffers()
    batch.run(dsrc, dst)
    ctx.synchronize()
    for i, ((w, h, b), s) in enumerate(zip(sizes, srcs)):
        alone = v.Pipeline(ctx, w, h, b).thumbnail(v.WIDTH, 200, 0).run_host(s)
        assert np.array_equal(dst[i].numpy(), alone), i


def test_batch_identity_and_upsize_free_plans(ctx):
    # Images inside the box keep their size (the thumbnail never upsizes):
    # their plan is empty and the batch copies them.
    sizes = [(100, 80, 3), (150, 150, 4), (1, 1, 1), (149, 3, 2)]
    srcs = sources(sizes, 3)
    exp = expected(srcs, v.FITBOX, 150, 150)
    batch = v.Batch(ctx, sizes, v.FITBOX, 150, 150)
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    dst = batch.output_buffers()
    batch.run(dsrc, dst)
    ctx.synchronize()
    check_outputs(dst, exp, "identity")


class _View:
    """A device pointer with an offset, for argument checks (never launched)."""

    def __init__(self, base, offset, nbytes):
        self._p = base.data_ptr() + offset
        self.nbytes = nbytes

    def data_ptr(self):
        return self._p


def test_batch_rejects_bad_buffers(ctx):
    sizes = [(300, 200, 3), (50, 60, 4)]
    batch = v.Batch(ctx, sizes, v.FITBOX, 40, 40)
    srcs = [v.DeviceBuffer(ctx, (h, w, b), np.uint8) for (w, h, b) in sizes]
    dst = batch.output_buffers()
    with pytest.raises(v.VgpuError) as e:
        batch.run(srcs, dst[:1] + [v.DeviceBuffer(ctx, (1, 1, 4), np.uint8)])
    assert e.value.status == v.api.E_BUFFER_SIZE
    big = v.DeviceBuffer(ctx, (200 * 300 * 3 + 16,), np.uint8)
    with pytest.raises(v.VgpuError) as e:
        batch.run([_View(big, 1, 200 * 300 * 3), srcs[1]], dst)
    assert e.value.status == v.api.E_INVALID_ARGUMENT
    with pytest.raises(v.VgpuError):
        v.Batch(ctx, [(10, 10, 5)], v.FITBOX, 4, 4)
    with pytest.raises(v.VgpuError):
        v.Batch(ctx, [(10, 10, 3)], v.FITBOX, 4, 4, v.LBB)
    # The batch still runs after the rejected calls.
    srcs_np = sources(sizes, 1)
    for d, s in zip(srcs, srcs_np):
        d.upload(s)
    batch.run(srcs, dst)
    ctx.synchronize()
    check_outputs(dst, expected(srcs_np, v.FITBOX, 40, 40), "after rejects")


def test_batch_tune_keeps_results(ctx):
    sizes = random_sizes(12, 1234, 100, 1200)
    srcs = sources(sizes, 8)
    exp = expected(srcs, v.FITBOX, 100, 100)
    batch = v.Batch(ctx, sizes, v.FITBOX, 100, 100)
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    dst = batch.output_buffers()
    ms = batch.tune(dsrc, dst, iterations=2)
    assert ms > 0
    info = batch.info()
    assert info["use_graph"] == 1
    for d in dst:
        d.upload(np.zeros(d.shape, np.uint8))
    batch.run(dsrc, dst)
    ctx.synchronize()
    check_outputs(dst, exp, f"tuned {info}")


def test_device_fill_matches_host_generator(ctx):
    for n, seed in ((1001, 5), (64, 2**63 + 11), (7, 0)):
        d = v.DeviceBuffer(ctx, (n,), np.uint8)
        d.fill_random(seed)
        assert np.array_equal(d.numpy(), splitmix_bytes(n, seed)), (n, seed)


def test_batch_uneven_sizes_fitbox_400(ctx):
    # The uneven-size batch at reduced count: W, H ~ U{256..3000}, 3 or 4
    # bands, seeded device-made pixels, FitBox(400, 400) Lanczos3.
    r = np.random.default_rng(4004)
    sizes = [(int(r.integers(256, 3001)), int(r.integers(256, 3001)), int(r.choice([3, 4]))) for _ in range(10)]
    srcs = []
    for i, (w, h, b) in enumerate(sizes):
        d = v.DeviceBuffer(ctx, (h, w, b), np.uint8)
        d.fill_random(9000 + i)
        srcs.append(d)
    batch = v.Batch(ctx, sizes, v.FITBOX, 400, 400, v.LANCZOS3)
    assert batch.info()["fallback_images"] == 0
    dst = batch.output_buffers()
    batch.run(srcs, dst)
    ctx.synchronize()
    for i, (w, h, b) in enumerate(sizes):
        host = splitmix_bytes(w * h * b, 9000 + i).reshape(h, w, b)
        exp = O.Pipeline(w, h, b).thumbnail(O.FITBOX, 400, 400).run(host)
        assert np.array_equal(dst[i].numpy(), exp), (i, sizes[i])


def test_batch_affine_tails_run_stage_by_stage(ctx):
    # ForceExact plans with an Affine tail (upscales, 1-pixel inputs, and u8
    # ReduceH ahead of a SmallTile Affine, planned per 128-column tile column)
    # leave the batched launches and run stage by stage, bit-exact.
    sizes = [(300, 200, 3), (400, 300, 1), (500, 120, 4), (1, 700, 3), (900, 40, 2), (2000, 1500, 3)]
    srcs = sources(sizes, 21)
    exp = expected(srcs, v.FORCE_EXACT, 333, 555)
    batch = v.Batch(ctx, sizes, v.FORCE_EXACT, 333, 555)
    tails = sum(1 for (w, h, b) in sizes
                if "Affine" in [n for n, _ in O.thumbnail_plan(w, h, b, O.FORCE_EXACT, 333, 555)[0]])
    assert tails == 5 and batch.info()["fallback_images"] == tails
    dsrc = [v.DeviceBuffer.from_numpy(ctx, s) for s in srcs]
    dst = batch.output_buffers()
    for rep in range(2):
        batch.run(dsrc, dst)
        ctx.synchronize()
        check_outputs(dst, exp, f"affine tails, run {rep}")
