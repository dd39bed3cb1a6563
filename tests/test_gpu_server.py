"""The request loop through the C ABI: every request's output equals the
oracle's pipeline on that image and the same request run alone, whether it
ran alone or mixed with others, whatever the source memory (device, pinned or
pageable host); admission holds the arena budget, rejects a request larger
than the whole budget and keeps queued requests waiting until memory is
released."""
import ctypes as C

import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import noise_u8

pytestmark = pytest.mark.gpu

MB = 1 << 20


def c5_specs(n, seed, lo=0.02, hi=1.5):
    """C5-like requests at reduced size: P ~ logU[lo, hi] MP, aspect ~ U[0.5, 2],
    3 bands (every 5th: 1 or 4), Width(256 | 400 | 800), then gauss 0.8."""
    r = np.random.default_rng(seed)
    out = []
    for i in range(n):
        P = np.exp(r.uniform(np.log(lo * 1e6), np.log(hi * 1e6)))
        a = r.uniform(0.5, 2.0)
        w, h = max(int(round(np.sqrt(P * a))), 1), max(int(round(np.sqrt(P / a))), 1)
        b = 3 if i % 5 else (1, 4)[(i // 5) % 2]
        out.append((w, h, b, (256, 400, 800)[i % 3]))
    return out


def expected(src, tw, sigma=0.8):
    h, w, b = src.shape
    return O.Pipeline(w, h, b).thumbnail(O.WIDTH, tw).gauss_blur(sigma).run(src)


def req_for(srv, src, tw, mode, pinned_bufs):
    h, w, b = src.shape
    if mode == "device":
        d = v.DeviceBuffer.from_numpy(srv.ctx, src)
        return v.Server.request(d, w, h, b, target=v.WIDTH, target_w=tw, blur_sigma=0.8), d
    if mode == "pinned":
        p = C.c_void_p()
        assert srv._lib.vgpu_host_alloc(src.nbytes, C.byref(p)) == 0
        C.memmove(p.value, src.ctypes.data, src.nbytes)
        pinned_bufs.append(p.value)
        return v.Server.request(p.value, w, h, b, target=v.WIDTH, target_w=tw, blur_sigma=0.8,
                                on_device=False, pinned=True), None
    return v.Server.request(src, w, h, b, target=v.WIDTH, target_w=tw, blur_sigma=0.8), src


def test_server_results_alone_and_mixed(ctx):
    specs = c5_specs(18, 7)
    srcs = [noise_u8(h, w, b, 100 + i) for i, (w, h, b, _) in enumerate(specs)]
    exp = [expected(s, tw) for s, (_, _, _, tw) in zip(srcs, specs)]
    srv = v.Server(ctx, hbm_budget=256 * MB, max_inflight=16, streams=4, planner_threads=3)
    pinned = []
    try:
        alone = []
        for i, (s, (_, _, _, tw)) in enumerate(zip(srcs, specs)):
            req, keep = req_for(srv, s, tw, "device", pinned)
            rid = srv.submit(req, keep)
            out = srv.fetch(rid)
            srv.release(rid)
            assert np.array_equal(out, exp[i]), (i, specs[i])
            alone.append(out)
        # All at once, three source kinds interleaved, each submitted twice.
        ids = []
        for rep in range(2):
            for i, (s, (_, _, _, tw)) in enumerate(zip(srcs, specs)):
                mode = ("device", "pageable", "pinned")[(i + rep) % 3]
                req, keep = req_for(srv, s, tw, mode, pinned)
                ids.append((i, srv.submit(req, keep)))
        for i, rid in ids:
            out = srv.fetch(rid)
            srv.release(rid)
            assert np.array_equal(out, alone[i]), (i, specs[i])
        st = srv.stats()
        assert st["completed"] >= 3 * len(specs) and st["rejected"] == 0, st
        assert st["arena_peak"] <= st["arena_bytes"], st
        assert st["peak_inflight"] <= 16, st
    finally:
        srv.close()
        for p in pinned:
            srv._lib.vgpu_host_free(C.c_void_p(p))


def test_server_admission_budget_and_rejection(ctx):
    # Each 600x500x3 request holds about 0.9 MB of input copy plus its
    # intermediates: a 3 MB arena holds two or three at once.
    src = noise_u8(500, 600, 3, 3)
    exp = expected(src, 256)
    srv = v.Server(ctx, hbm_budget=3 * MB, max_inflight=64, streams=2, planner_threads=2)
    try:
        big = noise_u8(1200, 1500, 3, 4)  # 5.4 MB input copy: over the whole budget
        rid_big = srv.submit(v.Server.request(big, 1500, 1200, 3, target=v.WIDTH, target_w=256,
                                              blur_sigma=0.8), big)
        ids = [srv.submit(v.Server.request(src, 600, 500, 3, target=v.WIDTH, target_w=256, blur_sigma=0.8),
                          src) for _ in range(12)]
        r = srv.wait(rid_big)
        assert r["status"] == v.api.E_ADMISSION, r
        for rid in ids:
            out = srv.fetch(rid)
            srv.release(rid)
            assert np.array_equal(out, exp)
        st = srv.stats()
        assert st["rejected"] == 1, st
        assert st["arena_peak"] <= 3 * MB, st
        assert st["peak_inflight"] >= 1, st
    finally:
        srv.close()


def test_server_waits_for_released_memory(ctx):
    src = noise_u8(500, 600, 3, 5)
    exp = expected(src, 400)
    probe = v.Server(ctx, hbm_budget=64 * MB, max_inflight=4, streams=1, planner_threads=1)
    try:
        rid = probe.submit(v.Server.request(src, 600, 500, 3, target=v.WIDTH, target_w=400, blur_sigma=0.8), src)
        chunk = probe.wait(rid)["arena_bytes"]
        probe.release(rid)
    finally:
        probe.close()
    # Room for exactly two requests: the third waits until one is released.
    srv = v.Server(ctx, hbm_budget=2 * chunk + 256, max_inflight=8, streams=2, planner_threads=1)
    try:
        mk = lambda: v.Server.request(src, 600, 500, 3, target=v.WIDTH, target_w=400, blur_sigma=0.8)
        a, b, c = (srv.submit(mk(), src) for _ in range(3))
        assert srv.wait(a)["status"] == 0 and srv.wait(b)["status"] == 0
        with pytest.raises(v.VgpuError) as e:
            srv.wait(c, timeout_us=200_000)
        assert e.value.status == v.api.E_NOT_READY
        assert srv.stats()["queued"] == 1
        out_a = srv.fetch(a)
        srv.release(a)
        out_c = srv.fetch(c)
        assert np.array_equal(out_a, exp) and np.array_equal(out_c, exp)
        srv.release(b)
        srv.release(c)
    finally:
        srv.close()


def test_server_single_slot_and_identity(ctx):
    # One slot: requests run one at a time; an image already inside the
    # target (no resampling) is copied, then blurred.
    specs = [(300, 200, 3, 256), (200, 150, 4, 800), (1000, 90, 1, 400)]
    srcs = [noise_u8(h, w, b, 40 + i) for i, (w, h, b, _) in enumerate(specs)]
    srv = v.Server(ctx, hbm_budget=64 * MB, max_inflight=1, streams=3, planner_threads=2)
    try:
        ids = [srv.submit(v.Server.request(s, s.shape[1], s.shape[0], s.shape[2], target=v.WIDTH,
                                           target_w=tw, blur_sigma=0.8), s)
               for s, (_, _, _, tw) in zip(srcs, specs)]
        for rid, s, (_, _, _, tw) in zip(ids, srcs, specs):
            out = srv.fetch(rid)
            srv.release(rid)
            assert np.array_equal(out, expected(s, tw))
        assert srv.stats()["peak_inflight"] == 1
    finally:
        srv.close()


def test_server_replay_open_loop(ctx):
    specs = c5_specs(8, 11)
    bufs = []
    for i, (w, h, b, _) in enumerate(specs):
        d = v.DeviceBuffer(ctx, (h, w, b), np.uint8)
        d.fill_random(500 + i)
        bufs.append(d)
    ctx.synchronize()
    r = np.random.default_rng(3)
    n = 300
    arrivals = np.cumsum(r.exponential(1.0 / 2000.0, size=n)) * 1e9
    reqs = []
    for k in range(n):
        i = int(r.integers(0, len(specs)))
        w, h, b, _ = specs[i]
        reqs.append(v.Server.request(bufs[i], w, h, b, target=v.WIDTH, target_w=(256, 400, 800)[k % 3],
                                     blur_sigma=0.8))
    srv = v.Server(ctx, hbm_budget=512 * MB, max_inflight=32, streams=4, planner_threads=4)
    try:
        lat, status, wall = srv.replay(reqs, arrivals.astype(np.int64))
        assert (status == 0).all(), status
        assert (lat > 0).all() and wall > 0
        st = srv.stats()
        assert st["completed"] == n and st["arena_used"] == 0, st
    finally:
        srv.close()
