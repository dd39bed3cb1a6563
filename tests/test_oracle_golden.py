"""Pin the CPU oracle before trusting it (CPU only).

Checks the oracle against (1) the reference's committed fixture files,
(2) the reference's in-crate known answers, (3) invariants and closed forms
that a dropped term, a wrong sign/index or a transposed operand would break.
"""
from fractions import Fraction
import pathlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import (grayscale_source, noise_u8, rgb_source, smooth_grayscale_source,
                           structural_source)

FIX = pathlib.Path(__file__).resolve().parent / "golden" / "reference_fixtures"


def fixture_u8(rel):
    return np.fromfile(FIX / rel, np.uint8)


# ---------------------------------------------------------------- fixtures

def _gray8x8_prefix(w, h):
    # grayscale_source() is 8x8; the tests take its first w*h samples
    # (tests/functional/golden/libvips/resample.rs:124-132, 157-165).
    return grayscale_source(8, 8).ravel()[: w * h].reshape(h, w, 1)


@pytest.mark.parametrize("rel,w,h,b,op,factor,kernel,src", [
    ("reduceh_libvips/grayscale_factor1_5_lanczos3.bin", 8, 6, 1, "h", 1.5, O.LANCZOS3, "gray_prefix"),
    ("reducev_libvips/grayscale_factor1_5_lanczos3.bin", 6, 8, 1, "v", 1.5, O.LANCZOS3, "gray_prefix"),
    ("reduceh_libvips/grayscale_factor1_5_cubic.bin", 8, 6, 1, "h", 1.5, O.BICUBIC, "gray"),
    ("reducev_libvips/grayscale_factor1_5_cubic.bin", 6, 8, 1, "v", 1.5, O.BICUBIC, "gray"),
    ("reduceh_libvips/rgb_factor2_linear.bin", 8, 6, 3, "h", 2.0, O.BILINEAR, "rgb"),
    ("reducev_libvips/rgb_factor2_linear.bin", 6, 8, 3, "v", 2.0, O.BILINEAR, "rgb"),
    ("libvips_parity_resample/test_reduceh_factor_1_5_lanczos3.bin", 8, 6, 1, "h", 1.5, O.LANCZOS3, "smooth"),
    ("libvips_parity_resample/test_reducev_factor_1_5_lanczos3.bin", 6, 8, 1, "v", 1.5, O.LANCZOS3, "smooth"),
])
def test_reduce_matches_reference_fixture(rel, w, h, b, op, factor, kernel, src):
    img = {"gray_prefix": lambda: _gray8x8_prefix(w, h), "gray": lambda: grayscale_source(w, h),
           "rgb": lambda: rgb_source(w, h), "smooth": lambda: smooth_grayscale_source(w, h)}[src]()
    p = O.Pipeline(w, h, b)
    (p.reduce_h if op == "h" else p.reduce_v)(factor, kernel)
    assert np.array_equal(p.run(img).ravel(), fixture_u8(rel))


def test_reduce_v_then_h_edge_fixture():
    # reduce_4x4_edge_pixels_match_libvips (golden/libvips/resample.rs:214-248)
    src = np.arange(16, dtype=np.uint8).reshape(4, 4, 1)
    out = O.Pipeline(4, 4, 1).reduce_v(2.0, O.BILINEAR).reduce_h(2.0, O.BILINEAR).run(src)
    assert np.array_equal(out.ravel(), fixture_u8("reduce_libvips/edge_4x4_factor2_linear.bin"))


@pytest.mark.parametrize("rel,op", [("shrinkh/grayscale_factor_2.bin", "h"),
                                    ("shrinkv/grayscale_factor_2.bin", "v")])
def test_shrink_factor2_fixture(rel, op):
    p = O.Pipeline(10, 16, 1)
    (p.shrink_h if op == "h" else p.shrink_v)(2)
    assert np.array_equal(p.run(structural_source(10, 16)).ravel(), fixture_u8(rel))


@pytest.mark.parametrize("rel", ["shrink_libvips/rgb_factor4.bin",
                                 "libvips_parity_resample/test_shrink_factor4_rgb.bin"])
def test_shrink_factor4_rgb_fixture(rel):
    out = O.Pipeline(16, 16, 3).shrink_v(4).shrink_h(4).run(rgb_source(16, 16))
    assert np.array_equal(out.ravel(), fixture_u8(rel))


def test_gauss_f32_uniform_fixture():
    # gauss_blur_matches_libvips_fixture (golden/basic.rs:303-315), tolerance 1e-3
    exp = np.fromfile(FIX / "gauss_blur/uniform_sigma_1_5.bin", np.float32)
    src = np.full((64, 64, 1), 37.5, np.float32)
    out = O.Pipeline(64, 64, 1, O.F32).gauss_blur(1.5).run(src).ravel()
    assert np.max(np.abs(out - exp)) <= 1e-3


def test_thumbnail_libvips_fixture_within_reference_tolerance():
    # libvips_parity_upstream_resample_thumbnail_width_target
    # (parity/resample.rs:231-275): 512^2 -> 128^2, tolerance 16 vs libvips.
    src = smooth_grayscale_source(512, 512)
    out = O.Pipeline(512, 512, 1).thumbnail(O.WIDTH, 128).run(src)
    exp = fixture_u8("libvips_parity_resample/test_thumbnail_width_128_square.bin")
    assert out.shape == (128, 128, 1)
    assert np.max(np.abs(out.ravel().astype(int) - exp.astype(int))) <= 16


# --------------------------------------------------------- known answers

def test_reduce_known_answers():
    # reduceh.rs:596-610, reducev.rs:532-545
    grad = np.arange(8, dtype=np.uint8)
    assert list(O.Pipeline(8, 1, 1).reduce_h(2.0, O.BILINEAR).run(grad.reshape(1, 8, 1)).ravel()) == [1, 3, 5, 6]
    assert list(O.Pipeline(1, 8, 1).reduce_v(2.0, O.BILINEAR).run(grad.reshape(8, 1, 1)).ravel()) == [1, 3, 5, 6]
    px = np.array([10, 20, 30, 40, 50, 60, 70, 80, 90, 100, 110, 120], np.uint8)
    assert list(O.Pipeline(2, 2, 3).reduce_h(2.0, O.BILINEAR).run(px.reshape(2, 2, 3)).ravel()) == [25, 35, 45, 85, 95, 105]
    assert list(O.Pipeline(1, 4, 3).reduce_v(2.0, O.BILINEAR).run(px.reshape(4, 1, 3)).ravel()) == [29, 39, 49, 81, 91, 101]
    # reducev.rs:549-570: the 30-value Lanczos3 x1.5 list
    exp = [57, 74, 25, 42, 59, 76, 98, 115, 100, 117, 51, 68, 47, 64, 89, 106, 130, 147, 126,
           143, 77, 94, 79, 96, 118, 135, 152, 169, 120, 137]
    assert list(O.Pipeline(6, 8, 1).reduce_v(1.5).run(_gray8x8_prefix(6, 8)).ravel()) == exp


@pytest.mark.parametrize("kernel", [O.NEAREST, O.BILINEAR, O.BICUBIC, O.CATMULLROM, O.LANCZOS2, O.LANCZOS3])
def test_single_pixel_edge_copy(kernel):
    # reduceh.rs:580-594, reducev.rs:517-530
    assert O.Pipeline(1, 1, 1).reduce_h(2.0, kernel).run(np.array([[[173]]], np.uint8)).ravel()[0] == 173
    assert O.Pipeline(1, 1, 1).reduce_v(2.0, kernel).run(np.array([[[91]]], np.uint8)).ravel()[0] == 91


def test_shrink_known_answers():
    # shrinkh.rs:1479-1506
    assert list(O.Pipeline(4, 1, 1).shrink_h(2).run(np.array([10, 20, 30, 40], np.uint8).reshape(1, 4, 1)).ravel()) == [15, 35]
    rgb = np.array([10, 20, 30, 20, 40, 60, 30, 60, 90, 40, 80, 120], np.uint8).reshape(1, 4, 3)
    assert list(O.Pipeline(4, 1, 3).shrink_h(2).run(rgb).ravel()) == [15, 30, 45, 35, 70, 105]
    f5 = np.array([10, 20, 30, 20, 30, 40, 30, 40, 50, 40, 50, 60, 50, 60, 70, 100, 110, 120, 110, 120,
                   130, 120, 130, 140, 130, 140, 150, 140, 150, 160], np.uint8).reshape(1, 10, 3)
    assert list(O.Pipeline(10, 1, 3).shrink_h(5).run(f5).ravel()) == [30, 40, 50, 120, 130, 140]
    # shrinkh.rs:1792-1806: factor-5 RGB with the 239 tail
    t = np.array([10, 20, 30, 20, 30, 40, 30, 40, 50, 40, 50, 60, 50, 60, 70, 100, 110, 120, 110, 120,
                  130, 120, 130, 140, 130, 140, 150, 140, 150, 160, 200, 210, 220, 210, 220, 230, 220,
                  230, 240, 230, 240, 250, 240, 250, 255], np.uint8).reshape(1, 15, 3)
    assert list(O.Pipeline(15, 1, 3).shrink_h(5).run(t).ravel()) == [30, 40, 50, 120, 130, 140, 220, 230, 239]
    rgba = np.arange(1, 41, dtype=np.uint8).reshape(1, 10, 4)
    assert list(O.Pipeline(10, 1, 4).shrink_h(5).run(rgba).ravel()) == [9, 10, 11, 12, 29, 30, 31, 32]
    # shrinkh.rs:1895-1901: generic dispatch, 2 bands, factor 3
    two = np.array([10, 100, 20, 110, 30, 120, 40, 130, 50, 140, 60, 150], np.uint8).reshape(1, 6, 2)
    assert list(O.Pipeline(6, 1, 2).shrink_h(3).run(two).ravel()) == [20, 110, 50, 140]


def test_multiply_shift_is_not_exact_rounding():
    # f = 3, sum 2 -> (3 * 5592405) >> 24 = 0 although 2/3 rounds to 1: the
    # libvips UCHAR_SHRINK rounding the reference keeps.
    src = np.array([1, 1, 0], np.uint8).reshape(1, 3, 1)
    assert O.Pipeline(3, 1, 1).shrink_h(3).run(src).ravel()[0] == 0


# ------------------------------------------------ tables and invariants

@pytest.mark.parametrize("kernel", [O.NEAREST, O.BILINEAR, O.BICUBIC, O.CATMULLROM, O.LANCZOS2, O.LANCZOS3])
@pytest.mark.parametrize("factor", [1.0, 1.5, 2.0, 2.5, 4.0, 3.7, 1.0778947368421052])
def test_reduce_table_invariants(kernel, factor):
    # reduce_common.rs:332-372
    cf, ci = O.reduce_tables(kernel, factor)
    taps = cf.shape[1]
    assert np.all(np.abs(cf.sum(axis=1) - 1.0) < 1e-12)
    assert np.all(np.abs(ci.astype(np.int64).sum(axis=1) - 4096) <= taps)
    # i16 = trunc(w * 4096): toward zero, never rounded
    assert np.array_equal(ci, np.trunc(cf * 4096.0).astype(np.int16))


def test_tap_counts_appendix():
    for f, taps in [(3.7, 23), (2.5, 17), (2.0, 13), (1.5, 11), (1.0778947368421052, 7), (1.0, 1)]:
        assert O.lib().orc_reduce_taps(O.LANCZOS3, f) == taps
    assert O.lib().orc_reduce_taps(O.LANCZOS3, 400.0) == -1  # reduce_common.rs:414-421


def test_lanczos3_weights_closed_form():
    # Independent closed form: sinc(x) * sinc(x/3) with numpy's sinc.
    for x in [0.1, 0.5, 1.3, 2.2, 2.9]:
        exp = np.sinc(x) * np.sinc(x / 3.0)
        assert abs(O.lib().orc_kernel_weight(O.LANCZOS3, x) - exp) < 1e-14
    assert O.lib().orc_kernel_weight(O.LANCZOS3, 3.0) == 0.0
    # Keys a = -0.5 cubic: (a+2)|x|^3 - (a+3)|x|^2 + 1 and a|x|^3 - 5a|x|^2 + 8a|x| - 4a
    for x in [0.25, 0.75, 1.25, 1.75]:
        a = -0.5
        exp = ((a + 2) * x**3 - (a + 3) * x**2 + 1) if x < 1 else (a * x**3 - 5 * a * x**2 + 8 * a * x - 4 * a)
        assert abs(O.lib().orc_kernel_weight(O.BICUBIC, x) - exp) < 1e-14


def test_factor2_offset_and_start():
    # reduce_common.rs:375-384: offset -0.5, sp(0) = 1, start = -1
    assert abs(O.lib().orc_reduce_offset(O.BILINEAR, 2.0, 8) + 0.5) < 1e-12
    starts, _ = O.reduce_plan(O.BILINEAR, 2.0, 8, 0.0, False, 1)
    assert starts[0] == -1


@pytest.mark.parametrize("kernel", [O.BILINEAR, O.BICUBIC, O.LANCZOS2, O.LANCZOS3])
@pytest.mark.parametrize("factor", [1.0, 1.5, 2.0, 2.5, 3.5, 4.0, 8.0])
def test_uniform_rows_stay_uniform(kernel, factor):
    # reduceh.rs:806-830 (proptests restated over all values)
    for v in [0, 1, 77, 128, 254, 255]:
        row = np.full((1, 64, 1), v, np.uint8)
        assert np.all(O.Pipeline(64, 1, 1).reduce_h(factor, kernel).run(row) == v)


def test_factor1_is_identity():
    src = noise_u8(5, 32, 3, 1)
    for k in [O.BILINEAR, O.LANCZOS3]:
        assert np.array_equal(O.Pipeline(32, 5, 3).reduce_h(1.0, k).run(src), src)


def test_plan_accumulation_differs_from_direct_only_by_drift():
    # u8 ReduceH plans by f64 accumulation (reduceh.rs:112-119); it must agree
    # with the direct position except where the drift crosses a boundary.
    # The quantised position start * 64 + phase may move by at most one 1/64
    # step; at f = 3.7 several columns land on (start + 1, phase 0) instead of
    # (start, phase 64), the same position expressed with a shifted window.
    s1, p1 = O.reduce_plan(O.LANCZOS3, 3.7, 1024, 0.0, True, 277)
    s2, p2 = O.reduce_plan(O.LANCZOS3, 3.7, 1024, 0.0, False, 277)
    q1 = s1 * 64 + p1.astype(np.int64)
    q2 = s2 * 64 + p2.astype(np.int64)
    assert np.all(np.abs(q1 - q2) <= 1)
    assert np.any(s1 != s2)


# ------------------------------------------------------------ Gaussian

def test_gauss_masks_appendix():
    c, s = O.gauss_int_kernel(1.5)
    assert c.tolist() == [8, 16, 20, 16, 8] and s == 68  # gauss_blur.rs:56-65
    c, s = O.gauss_int_kernel(2.0)
    assert c.tolist() == [6, 12, 18, 20, 18, 12, 6] and s == 92
    c, s = O.gauss_int_kernel(0.8)
    assert c.tolist() == [9, 20, 9] and s == 38
    c, s = O.gauss_int_kernel(8.0)
    assert len(c) == 29 and s == 372
    c, s = O.gauss_int_kernel(0.1)
    assert c.tolist() == [1] and s == 1
    # taps = round_ties_even(20 exp(-x^2 / 2 sigma^2)), closed form
    for sigma in [0.8, 1.0, 2.0, 3.3, 8.0]:
        c, s = O.gauss_int_kernel(sigma)
        r = (len(c) - 1) // 2
        sig = float(np.float32(sigma))
        exp = np.rint(20 * np.exp(-(np.arange(-r, r + 1) ** 2) / (2 * sig * sig))).astype(int)
        assert c.tolist() == exp.tolist() and s == exp.sum()


@pytest.mark.parametrize("sigma,bands", [(0.8, 3), (1.5, 1), (2.0, 4), (3.0, 3)])
def test_gauss_u8_matches_bruteforce_2d(sigma, bands):
    # naive_integer_separable_blur_u8 (gauss_blur.rs:2044-2092): clamp-padded
    # H pass to u8, then V pass -- written here with numpy padding.
    src = noise_u8(23, 31, bands, 7)
    c, s = O.gauss_int_kernel(sigma)
    r = (len(c) - 1) // 2
    pad = np.pad(src.astype(np.int64), ((0, 0), (r, r), (0, 0)), mode="edge")
    hz = sum(int(c[t]) * pad[:, t:t + 31, :] for t in range(len(c)))
    hz = np.clip((hz + s // 2) // s, 0, 255)
    pad = np.pad(hz, ((r, r), (0, 0), (0, 0)), mode="edge")
    vt = sum(int(c[t]) * pad[t:t + 23, :, :] for t in range(len(c)))
    exp = np.clip((vt + s // 2) // s, 0, 255).astype(np.uint8)
    out = O.Pipeline(31, 23, bands).gauss_blur(sigma).run(src)
    assert np.array_equal(out, exp)


# ------------------------------------------------------------- alpha

def test_premultiply_exhaustive_against_exact_rationals():
    # premultiply.rs:103-115 in f64 vs exact round-half-away of c*a/255;
    # unpremultiply.rs:96-109 vs exact round-half-away of c*255/a.
    a = np.repeat(np.arange(256, dtype=np.uint8), 256)
    c = np.tile(np.arange(256, dtype=np.uint8), 256)
    img = np.stack([c, c, c, a], axis=-1).reshape(256, 256, 4)
    pm = O.Pipeline(256, 256, 4).premultiply().run(img).reshape(-1, 4)
    um = O.Pipeline(256, 256, 4).unpremultiply().run(img).reshape(-1, 4)

    def rha(fr):  # round half away from zero for a non-negative Fraction
        return int(fr + Fraction(1, 2)) if fr >= 0 else -int(-fr + Fraction(1, 2))

    for i in range(0, 65536, 97):
        ci, ai = int(c[i]), int(a[i])
        assert pm[i, 0] == min(255, rha(Fraction(ci * ai, 255)))
        assert um[i, 0] == (0 if ai == 0 else min(255, rha(Fraction(ci * 255, ai))))
        assert pm[i, 3] == ai and um[i, 3] == ai
    # unpremultiply.rs:239-248: 64 * 255 / 128 = 127.5 -> 128
    px = np.array([64, 0, 128], np.uint8).reshape(1, 1, 3)
    assert O.Pipeline(1, 1, 3).unpremultiply().run(px).ravel().tolist() == [128, 0, 128]


# ------------------------------------------------------------ planners

def test_thumbnail_planner_shapes():
    nodes, dims = O.thumbnail_plan(8192, 8192, 3, O.WIDTH, 400)
    assert [n for n, _ in nodes] == ["ShrinkH", "ShrinkV", "ReduceH", "ReduceV"]
    assert nodes[0][1] == 19 and nodes[2][1] == 20.48 / 19 and dims == (400, 400)
    # thumbnail.rs:588-606: 2048 -> 400 shrinks by 4 with no Affine
    nodes, dims = O.thumbnail_plan(2048, 2048, 3, O.WIDTH, 400)
    assert nodes[0] == ("ShrinkH", 4.0) and nodes[1] == ("ShrinkV", 4.0)
    # thumbnail.rs:629-646: RGBA wraps in premultiply/unpremultiply
    nodes, _ = O.thumbnail_plan(37, 19, 4, O.WIDTH, 11)
    assert nodes[0][0] == "Premultiply" and nodes[-1][0] == "Unpremultiply"
    # Appendix B.3
    nodes, dims = O.thumbnail_plan(6000, 4000, 3, O.FITBOX, 400, 400)
    assert nodes == [("ShrinkH", 15.0), ("ShrinkV", 12.0), ("ReduceV", 4000 / 267 / 12)] and dims == (400, 267)
    # 1-pixel inputs take the Affine-only path
    assert [n for n, _ in O.thumbnail_plan(1, 2048, 3, O.HEIGHT, 0, 400)[0]] == ["Affine"]


def test_integer_shrink_never_undershoots():
    # thumbnail.rs:608-627, restated over the planner's output
    for tw in range(1, 2049, 7):
        try:
            nodes, dims = O.thumbnail_plan(2048, 2048, 1, O.WIDTH, tw)
        except O.OracleError:
            continue  # Affine tail
        assert dims[0] == tw


# ------------------------------------------------------ chain executor

@pytest.mark.parametrize("strip", [1, 3, 16, 1000])
def test_chain_is_tile_height_invariant(strip):
    src = noise_u8(197, 233, 3, 11)
    base = O.Pipeline(233, 197, 3).thumbnail(O.WIDTH, 64).gauss_blur(0.8).run(src, strip_height=16)
    out = O.Pipeline(233, 197, 3).thumbnail(O.WIDTH, 64).gauss_blur(0.8).run(src, strip_height=strip)
    assert np.array_equal(base, out)


def test_single_reduce_equals_clamp_padded_closed_form():
    # One ReduceV on a source is a clamp-padded 1-D filter: compare with numpy
    # edge padding and the table rows (x86 AVX2: direct per-row plan).
    src = noise_u8(50, 7, 2, 3)
    f = 2.3
    _, ci = O.reduce_tables(O.LANCZOS3, f)
    starts, phases = O.reduce_plan(O.LANCZOS3, f, 50, 0.0, False, int(O.lib().orc_reduce_output_len(f, 50)))
    taps = ci.shape[1]
    pad = 64
    padded = np.pad(src.astype(np.int64), ((pad, pad), (0, 0), (0, 0)), mode="edge")
    exp = np.stack([
        np.clip((sum(int(ci[p, t]) * padded[s + t + pad] for t in range(taps)) + 2048) >> 12, 0, 255)
        for s, p in zip(starts, phases)]).astype(np.uint8)
    out = O.Pipeline(7, 50, 2).reduce_v(f).run(src)
    assert np.array_equal(out, exp)


def test_thumbnail_then_blur_differs_from_separate_ops_at_borders():
    # Intermediates are evaluated on Z^2 (execute.rs:413-480): chaining
    # thumbnail -> gauss_blur is NOT thumbnail, then blur of the clamped result.
    src = noise_u8(300, 300, 3, 5)
    chained = O.Pipeline(300, 300, 3).thumbnail(O.WIDTH, 70).gauss_blur(2.0).run(src)
    thumb = O.Pipeline(300, 300, 3).thumbnail(O.WIDTH, 70).run(src)
    separate = O.Pipeline(70, 70, 3).gauss_blur(2.0).run(thumb)
    assert np.array_equal(chained[5:-5, 5:-5], separate[5:-5, 5:-5])
    assert not np.array_equal(chained, separate)


def test_synthetic_bytes_known_answers():
    # splitmix64 with seed 0: the published first outputs of the generator.
    w = O.synthetic_bytes(24, 0).view("<u8")
    assert [int(x) for x in w] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # Ragged lengths are prefixes of the same stream.
    assert np.array_equal(O.synthetic_bytes(13, 77), O.synthetic_bytes(16, 77)[:13])
    # The C generator writes the same stream.
    for n, seed in ((0, 1), (13, 77), (4099, 2**63 + 5), (1 << 16, 0)):
        assert np.array_equal(O.synthetic_bytes_fast(n, seed), O.synthetic_bytes(n, seed))


# ------------------------------------------------------------- ReduceOp

# The reference's ReduceOp fixtures (reduce.rs:535-578): 4x4x3 at factor 2 and
# 5x4x1 at 1.5 x 2.0, both Bilinear.
REDUCE_OP_RGB4 = np.array([
    0, 10, 20, 30, 45, 60, 90, 60, 30, 120, 90, 40, 15, 35, 55, 45, 75, 105, 95, 65,
    35, 125, 95, 45, 25, 50, 75, 55, 85, 115, 105, 70, 40, 135, 100, 50, 35, 60, 85,
    65, 95, 125, 115, 80, 45, 145, 105, 55], np.uint8).reshape(4, 4, 3)
REDUCE_OP_GRAY5x4 = np.array([
    0, 20, 80, 120, 200, 15, 35, 95, 135, 215, 25, 55, 110, 150, 230, 40, 70, 125, 165,
    245], np.uint8).reshape(4, 5, 1)

# Bilinear weights with offset 0 (no bind_input_len), derived by hand from the
# table convention of reduce_common.rs:279-316: output i reads source
# positions p -> weight.  Factor 2: sp = 2i + 0.5, phase 32, a [1,2,1]/4 filter
# centred on 2i.  Factor 1.5: sp = 1.5i + 0.25, phases 16/48 alternate:
# {-1: 1/3, 0: 5/9, 1: 1/9} around floor(sp) for phase 16, {-1: 1/9, 0: 5/9,
# 1: 1/3} for phase 48.
def _bilinear_taps(factor, i):
    F = Fraction
    if factor == 2.0:
        return {2 * i - 1: F(1, 4), 2 * i: F(1, 2), 2 * i + 1: F(1, 4)}
    assert factor == 1.5
    fl = (3 * i) // 2
    w = (F(1, 3), F(5, 9), F(1, 9)) if i % 2 == 0 else (F(1, 9), F(5, 9), F(1, 3))
    return {fl - 1: w[0], fl: w[1], fl + 1: w[2]}


def _round_half_away(q):
    # from_f64 on u8: clamp, then f64::round (half away from zero).
    q = min(max(q, Fraction(0)), Fraction(255))
    return int(q + Fraction(1, 2)) if q >= 0 else -int(-q + Fraction(1, 2))


def _reduce_op_closed_form(img, hf, vf):
    h, w, b = img.shape
    ow, oh = int(np.floor(w / hf + 0.5)), int(np.floor(h / vf + 0.5))
    # H pass over every source row the V pass reads (clamp to edge), rounded
    # to u8 as ReduceOp's scratch is.
    rows = {}
    def hrow(y):
        yy = min(max(y, 0), h - 1)
        if yy not in rows:
            rows[yy] = [[_round_half_away(sum(wt * int(img[yy, min(max(p, 0), w - 1), c])
                                              for p, wt in _bilinear_taps(hf, x).items()))
                         for c in range(b)] for x in range(ow)]
        return rows[yy]
    out = np.zeros((oh, ow, b), np.uint8)
    for y in range(oh):
        for x in range(ow):
            for c in range(b):
                out[y, x, c] = _round_half_away(sum(wt * hrow(p)[x][c] for p, wt in _bilinear_taps(vf, y).items()))
    return out


def test_reduce_op_identity_and_dims():
    # factor_one_is_identity, factor_two_halves_dimensions (reduce.rs:483-496)
    img = np.arange(12, dtype=np.uint8).reshape(3, 4, 1)
    assert np.array_equal(O.Pipeline(4, 3, 1).reduce(1.0, 1.0, O.BILINEAR).run(img), img)
    z = np.zeros((6, 8, 1), np.uint8)
    assert O.Pipeline(8, 6, 1).reduce(2.0, 2.0, O.BILINEAR).run(z).shape == (3, 4, 1)


@pytest.mark.parametrize("img,hf,vf,shape", [
    (REDUCE_OP_RGB4, 2.0, 2.0, (2, 2, 3)),
    (REDUCE_OP_GRAY5x4, 1.5, 2.0, (2, 3, 1)),
])
def test_reduce_op_reference_fixtures_closed_form(img, hf, vf, shape):
    h, w, b = img.shape
    got = O.Pipeline(w, h, b).reduce(hf, vf, O.BILINEAR).run(img)
    assert got.shape == shape
    assert np.array_equal(got, _reduce_op_closed_form(img, hf, vf))


@pytest.mark.parametrize("w,h,b,hf,vf,seed", [
    (2, 2, 1, 1.5, 2.0, 1), (5, 6, 3, 2.0, 1.5, 2), (6, 3, 2, 1.5, 1.5, 3), (37, 23, 3, 2.0, 2.0, 4),
    (40, 41, 1, 1.5, 2.0, 5),
])
def test_reduce_op_bilinear_closed_form_random(w, h, b, hf, vf, seed):
    # The proptest of reduce.rs:603-619 draws 2..6 x 2..6 x 1..3 images at
    # factors {1.5, 2, 3}; the exact-rational closed form pins borders, the
    # offset-0 phase, u8 scratch rounding and half-away rounding.
    img = noise_u8(h, w, b, seed)
    got = O.Pipeline(w, h, b).reduce(hf, vf, O.BILINEAR).run(img)
    assert np.array_equal(got, _reduce_op_closed_form(img, hf, vf))


def test_reduce_op_ties_round_half_away():
    # Two-pixel rows at factor 2 Bilinear: out = (x0 + 2 x0 + x1) / 4 with the
    # left edge clamped; x1 = x0 + 2 makes every H output an exact .5 tie.
    img = np.array([[[10], [12]], [[10], [12]]], np.uint8)  # 2x2
    got = O.Pipeline(2, 2, 1).reduce(2.0, 2.0, O.BILINEAR).run(img)
    assert got.shape == (1, 1, 1) and int(got[0, 0, 0]) == 11  # 10.5 -> 11


@pytest.mark.parametrize("kernel", [O.LANCZOS3, O.BICUBIC, O.BILINEAR])
@pytest.mark.parametrize("fmt", [O.U8, O.U16])
def test_reduce_op_uniform_stays_uniform(kernel, fmt):
    dt = np.uint8 if fmt == O.U8 else np.uint16
    v = 200 if fmt == O.U8 else 51234
    img = np.full((19, 23, 3), v, dt)
    got = O.Pipeline(23, 19, 3, fmt).reduce(1.7, 2.3, kernel).run(img)
    assert got.shape == (8, 14, 3) and (got == v).all()


def test_reduce_op_f32_matches_exact_separable_sum():
    # f32: no rounding between the passes beyond the f32 scratch; against an
    # FP64 separable sum with the oracle's own (pinned) tables, within f32
    # rounding of the scratch and the output.
    r = np.random.default_rng(9)
    img = r.random((17, 21, 2), dtype=np.float32)
    got = O.Pipeline(21, 17, 2, O.F32).reduce(1.5, 2.0, O.BILINEAR).run(img)
    exp = _reduce_op_closed_form_f(img.astype(np.float64), 1.5, 2.0)
    assert got.shape == exp.shape
    assert np.max(np.abs(got - exp)) < 2e-7


def _reduce_op_closed_form_f(img, hf, vf):
    h, w, b = img.shape
    ow, oh = int(np.floor(w / hf + 0.5)), int(np.floor(h / vf + 0.5))
    mid = np.zeros((h, ow, b))
    for x in range(ow):
        for p, wt in _bilinear_taps(hf, x).items():
            mid[:, x] += float(wt) * img[:, min(max(p, 0), w - 1)]
    mid = mid.astype(np.float32).astype(np.float64)
    out = np.zeros((oh, ow, b))
    for y in range(oh):
        for p, wt in _bilinear_taps(vf, y).items():
            out[y] += float(wt) * mid[min(max(p, 0), h - 1)]
    return out


def test_reduce_op_rejections():
    p = O.Pipeline(8, 8, 1)
    with pytest.raises(O.OracleError):
        p.reduce(0.5, 2.0, O.BILINEAR)
    with pytest.raises(O.OracleError):
        O.Pipeline(8, 8, 1).reduce(2.0, float("inf"), O.BILINEAR)
    with pytest.raises(O.OracleError) as e:
        O.Pipeline(8, 8, 1).reduce(2.0, 2.0, 99)
    assert e.value.code == O.EKERNEL


# ------------------------------------------------ thumbnail crop / extract

def test_thumbnail_crop_plan_reference_case():
    # thumbnail.rs:553-575: 400x200 FitBox 120x120 with crop reduces to the
    # covering 240x120 and takes the centred 120x120 at (60, 0).
    nodes, dims = O.thumbnail_plan(400, 200, 3, O.FITBOX, 120, 120, crop=True)
    assert [n for n, _ in nodes][-1] == "ExtractArea"
    assert nodes[-1][1] == (60, 0, 120, 120) and dims == (120, 120)
    # without crop the same box is fitted, and nothing is extracted
    nodes, dims = O.thumbnail_plan(400, 200, 3, O.FITBOX, 120, 120)
    assert "ExtractArea" not in [n for n, _ in nodes] and dims == (120, 60)
    # alpha: the extract sits inside the premultiplied section
    nodes, _ = O.thumbnail_plan(400, 200, 4, O.FITBOX, 120, 120, crop=True)
    assert [n for n, _ in nodes][-2:] == ["ExtractArea", "Unpremultiply"]


def _cover_dims(iw, ih, tw, th):
    # The smallest image of the input's aspect ratio that covers tw x th:
    # scale by the larger of the two ratios (never below the box).
    s = max(tw / iw, th / ih)
    return max(int(np.floor(iw * s + 0.5)), tw), max(int(np.floor(ih * s + 0.5)), th)


@pytest.mark.parametrize("seed", range(8))
def test_thumbnail_crop_is_centre_of_covering_thumbnail(seed):
    r = np.random.default_rng(500 + seed)
    iw, ih = int(r.integers(40, 300)), int(r.integers(40, 300))
    b = int(r.integers(1, 5))
    tw, th = int(r.integers(4, min(iw, ih))), int(r.integers(4, min(iw, ih)))
    src = noise_u8(ih, iw, b, seed)
    gw, gh = _cover_dims(iw, ih, tw, th)
    try:
        got = O.Pipeline(iw, ih, b).thumbnail(O.FITBOX, tw, th, crop=True).run(src)
    except O.OracleError as e:
        assert e.code == O.EUNSUPPORTED  # the plan needs the Affine tail
        return
    full = O.Pipeline(iw, ih, b).thumbnail(O.FORCE_EXACT, gw, gh).run(src)
    x0, y0 = (gw - tw) // 2, (gh - th) // 2
    assert got.shape == (th, tw, b)
    assert np.array_equal(got, full[y0:y0 + th, x0:x0 + tw])


def test_extract_area_is_a_slice_and_checks_bounds():
    src = noise_u8(37, 53, 3, 4)
    got = O.Pipeline(53, 37, 3).extract_area(7, 5, 20, 11).run(src)
    assert np.array_equal(got, src[5:16, 7:27])
    blurred = O.Pipeline(53, 37, 3).gauss_blur(1.0).run(src)
    got = O.Pipeline(53, 37, 3).gauss_blur(1.0).extract_area(0, 30, 53, 7).run(src)
    assert np.array_equal(got, blurred[30:37])
    for args in [(0, 0, 0, 5), (0, 0, 5, 0), (50, 0, 4, 5), (0, 33, 5, 5)]:
        with pytest.raises(O.OracleError):
            O.Pipeline(53, 37, 3).extract_area(*args)


# ------------------------------------------------------ thumbnail Affine tail

def test_affine_reference_cases():
    # affine/tests.rs:428-546, the ExtendMode::Copy row: nearest, tx = -2
    src = np.array([10, 11, 12, 13, 20, 21, 22, 23, 30, 31, 32, 33], np.uint8).reshape(1, 3, 4)
    got = O.Pipeline(3, 1, 4).affine(1.0, 1.0, -2.0, 0.0, 7, 1, O.NEAREST).run(src)
    assert got.ravel().tolist() == [10, 11, 12, 13] * 3 + [20, 21, 22, 23] + [30, 31, 32, 33] * 3
    # affine/tests.rs:840-857: RGBA bilinear premultiplies before interpolating
    src = np.array([255, 0, 0, 0, 0, 0, 255, 255], np.uint8).reshape(1, 2, 4)
    got = O.Pipeline(2, 1, 4).affine(1.0, 1.0, 0.5, 0.0, 1, 1, O.BILINEAR).run(src)
    assert got.ravel().tolist() == [0, 0, 255, 128]
    # affine/tests.rs:225-360: identity transforms return the input
    for b in (1, 3, 4):
        src = noise_u8(23, 31, b, b)
        if b == 4:
            src[..., 3] = np.maximum(src[..., 3], 1)
        for k in (O.NEAREST, O.BILINEAR, O.BICUBIC):
            got = O.Pipeline(31, 23, b).affine(1.0, 1.0, 0.0, 0.0, 31, 23, k).run(src)
            if b == 4 and k != O.NEAREST:
                # premultiply / divide round trips: within one level
                assert np.abs(got.astype(int) - src).max() <= 1
            else:
                assert np.array_equal(got, src), (b, k)


def test_thumbnail_affine_plan_reference_case():
    # thumbnail.rs:663-677: 1x2048 -> Height(400), Lanczos3 tail mapped to Bicubic
    nodes, dims = O.thumbnail_plan(1, 2048, 3, O.HEIGHT, 0, 400)
    assert dims == (1, 400) and [n for n, _ in nodes] == ["Affine"]
    a, d, tx, ty, w, h, k = nodes[0][1]
    assert (w, h, k) == (1, 400, O.BICUBIC) and a == 1.0 and d == 2048 / 400
    assert tx == 0.0 and ty == 0.5 * (1.0 - 2048 / 400)
    # thumbnail.rs:31-37: Nearest keeps a zero translation, Bilinear stays Bilinear
    nodes, _ = O.thumbnail_plan(1, 2048, 3, O.HEIGHT, 0, 400, kernel=O.NEAREST)
    assert nodes[0][1][2:4] == (0.0, 0.0) and nodes[0][1][6] == O.NEAREST
    nodes, _ = O.thumbnail_plan(1, 2048, 3, O.HEIGHT, 0, 400, kernel=O.BILINEAR)
    assert nodes[0][1][6] == O.BILINEAR
    # ForceExact upscale on one axis: reduce the other, then the Affine fix-up
    nodes, dims = O.thumbnail_plan(200, 1000, 4, O.FORCE_EXACT, 300, 300)
    assert [n for n, _ in nodes] == ["Premultiply", "ShrinkV", "ReduceV", "Affine", "Unpremultiply"]
    assert dims == (300, 300) and nodes[3][1][4:6] == (300, 300) and nodes[3][1][0] == 1.0 / (300 / 200)


def _catmull_rom(t):
    t = np.abs(t)
    return np.where(t < 1, 1.5 * t ** 3 - 2.5 * t ** 2 + 1,
                    np.where(t < 2, -0.5 * t ** 3 + 2.5 * t ** 2 - 4 * t + 2, 0.0))


def _affine_model(src, a, d, tx, ty, ow, oh, kernel, tile=(128, 128)):
    """Independent numpy statement of the axis-aligned Affine with the Copy
    extend: source positions start at each scheduler tile's origin and step by
    a (columns) and d (rows); interpolation in float64, before rounding."""
    h, w, b = src.shape
    amax = {np.uint8: 255.0, np.uint16: 65535.0}.get(src.dtype.type, 1.0)

    def coords(n, step, off, t):
        out = np.empty(n)
        for t0 in range(0, n, t):
            v = step * t0 + off
            for i in range(t0, min(t0 + t, n)):
                out[i] = v
                v += step
        return out

    xs, ys = coords(ow, a, tx, tile[0]), coords(oh, d, ty, tile[1])
    img = src.astype(np.float64)
    if b == 4:  # the Affine premultiplies four-band images itself
        img[..., :3] *= img[..., 3:4] / amax

    def px(iy, ix):  # (oh, ow) integer grids -> (oh, ow, b)
        return img[np.clip(iy, 0, h - 1), np.clip(ix, 0, w - 1)]

    Y, X = np.meshgrid(ys, xs, indexing="ij")
    if kernel == O.NEAREST:
        v = px(np.trunc(Y).astype(int), np.trunc(X).astype(int))
    elif kernel == O.BILINEAR:
        x0, y0 = np.trunc(X).astype(int), np.trunc(Y).astype(int)
        fx, fy = (X - x0)[..., None], (Y - y0)[..., None]
        top = px(y0, x0) * (1 - fx) + px(y0, x0 + 1) * fx
        bot = px(y0 + 1, x0) * (1 - fx) + px(y0 + 1, x0 + 1) * fx
        v = top * (1 - fy) + bot * fy
    else:
        cx, cy = np.floor(X).astype(int), np.floor(Y).astype(int)
        acc = np.zeros((oh, ow, b))
        wsum = np.zeros((oh, ow, 1))
        for ky in range(-1, 3):
            wy = _catmull_rom(cy + ky - Y)[..., None]
            for kx in range(-1, 3):
                wgt = _catmull_rom(cx + kx - X)[..., None] * wy
                acc += px(cy + ky, cx + kx) * wgt
                wsum += wgt
        v = acc / wsum
    if b == 4:
        alpha = np.clip(v[..., 3:4], 0, amax)
        na = alpha / amax
        v = np.concatenate([np.where(na > 0, v[..., :3] / np.where(na > 0, na, 1), 0.0), alpha], axis=-1)
    return v


def _assert_rounds_to(got, v, what):
    """got == round-half-away(clamp(v)), except within 1e-6 of a tie."""
    if got.dtype == np.float32:
        np.testing.assert_allclose(got, v, rtol=1e-6, atol=1e-6, err_msg=what)
        return
    hi = 255.0 if got.dtype == np.uint8 else 65535.0
    c = np.clip(v, 0, hi)
    exp = np.floor(c + 0.5)
    near_tie = np.abs(c - np.floor(c) - 0.5) < 1e-6
    ok = (got == exp) | (near_tie & (np.abs(got.astype(np.float64) - exp) <= 1))
    assert ok.all(), f"{what}: {np.count_nonzero(~ok)} of {ok.size} differ"


@pytest.mark.parametrize("kernel", [O.NEAREST, O.BILINEAR, O.BICUBIC])
@pytest.mark.parametrize("bands", [1, 3, 4])
@pytest.mark.parametrize("fmt", [np.uint8, np.uint16, np.float32])
def test_affine_matches_numpy_model(kernel, bands, fmt):
    r = np.random.default_rng(kernel * 100 + bands * 10 + np.dtype(fmt).itemsize)
    h, w = 37, 53
    if fmt == np.float32:
        src = r.random((h, w, bands)).astype(np.float32)
    else:
        src = r.integers(0, np.iinfo(fmt).max + 1, size=(h, w, bands)).astype(fmt)
    if bands == 4:
        src[..., 3] = np.where(r.random((h, w)) < 0.15, 0, src[..., 3])
    for (ow, oh) in [(97, 71), (40, 30), (150, 140)]:
        a, d = w / ow, h / oh
        tx = 0.0 if kernel == O.NEAREST else 0.5 * (1 - a)
        ty = 0.0 if kernel == O.NEAREST else 0.5 * (1 - d)
        for small, tile in [(True, (128, 128)), (False, (ow, 16))]:
            got = O.Pipeline(w, h, bands, {np.uint8: O.U8, np.uint16: O.U16, np.float32: O.F32}[fmt]) \
                .affine(a, d, tx, ty, ow, oh, kernel, small).run(src)
            v = _affine_model(src, a, d, tx, ty, ow, oh, kernel, tile)
            _assert_rounds_to(got, v, f"{kernel} {bands} {np.dtype(fmt)} {ow}x{oh} small={small}")


@pytest.mark.parametrize("kernel", [O.NEAREST, O.BILINEAR, O.BICUBIC])
def test_affine_uniform_stays_uniform(kernel):
    for b in (1, 3, 4):
        src = np.full((19, 23, b), 77, np.uint8)
        got = O.Pipeline(23, 19, b).affine(23 / 61, 19 / 45, 0.3, -0.2, 61, 45, kernel).run(src)
        assert (got == 77).all()


def test_thumbnail_affine_tail_runs():
    # ForceExact upscale of one axis, and a 1-pixel-wide input (Affine only)
    src = noise_u8(1000, 200, 3, 3)
    got = O.Pipeline(200, 1000, 3).thumbnail(O.FORCE_EXACT, 300, 300).run(src)
    assert got.shape == (300, 300, 3)
    col = noise_u8(2048, 1, 3, 4)
    got = O.Pipeline(1, 2048, 3).thumbnail(O.HEIGHT, 0, 400).run(col)
    assert got.shape == (400, 1, 3)
    # the Affine alone is the numpy model on the 1-pixel column
    v = _affine_model(col, 1.0, 2048 / 400, 0.0, 0.5 * (1 - 2048 / 400), 1, 400, O.BICUBIC, (1, 16))
    _assert_rounds_to(got, v, "1x2048 -> 1x400")


# ------------------------------------------------------------ resize upscale

N_REDUCE_H, N_AFFINE, N_ZOOM = 3, 15, 16


def test_resize_upscale_plans_reference_cases():
    # resize.rs:441-457: 0.5 x 2.5 Bilinear on 100x100 -> ReduceH(2), Zoom(1, 2), Affine; 50x250
    p = O.Pipeline(100, 100, 3).resize(0.5, 2.5, O.BILINEAR)
    nodes = p.nodes()
    assert [t for t, _ in nodes] == [N_REDUCE_H, N_ZOOM, N_AFFINE] and nodes[0][1] == 2.0
    assert p.output()[:2] == (50, 250)
    # resize.rs:485-496: Nearest upscale keeps a zero offset (1.5x of 100 -> Affine only)
    p = O.Pipeline(100, 100, 1).resize(1.5, 1.5, O.NEAREST)
    assert [t for t, _ in p.nodes()] == [N_AFFINE] and p.output()[:2] == (150, 150)
    # resize.rs:518-530: output dimensions are round(in * scale), at least 1
    for w, h, hs, vs in [(1, 1, 3.9, 0.05), (256, 7, 0.37, 3.3), (33, 200, 2.0, 1.01)]:
        p = O.Pipeline(w, h, 1).resize(hs, vs, O.LANCZOS3)
        assert p.output()[:2] == (max(1, int(np.floor(w * hs + 0.5))), max(1, int(np.floor(h * vs + 0.5))))


def test_resize_integer_upscale_is_replication():
    # An integer factor is a Zoom alone: each pixel repeated (zoom.rs:364-400)
    src = noise_u8(40, 50, 3, 8)
    got = O.Pipeline(50, 40, 3).resize(2.0, 3.0, O.LANCZOS3).run(src)
    assert np.array_equal(got, np.repeat(np.repeat(src, 3, axis=0), 2, axis=1))


@pytest.mark.parametrize("kernel", [O.LANCZOS2, O.LANCZOS3, O.CATMULLROM])
def test_affine_lanczos_matches_numpy_model(kernel):
    # the separable Affine with the resize kernels: Lanczos2 / CatmullRom over
    # 4 taps, Lanczos3 over 6 (affine/core.rs:648-697,728-749)
    def weight(t):
        t = np.abs(t)
        if kernel == O.CATMULLROM:
            return _catmull_rom(t)
        a = 3.0 if kernel == O.LANCZOS3 else 2.0
        with np.errstate(invalid="ignore", divide="ignore"):
            v = a * np.sin(np.pi * t) * np.sin(np.pi * t / a) / (np.pi * t) ** 2
        return np.where(t < 1e-12, 1.0, np.where(t < a, v, 0.0))

    S = 3 if kernel == O.LANCZOS3 else 2
    src = noise_u8(29, 41, 3, kernel)
    ow, oh = 100, 77
    a, d = 41 / ow, 29 / oh
    tx, ty = 0.5 * (1 - a), 0.5 * (1 - d)
    got = O.Pipeline(41, 29, 3).affine(a, d, tx, ty, ow, oh, kernel).run(src)
    # single 128 x 128 tile: positions step from the origin
    X = np.empty(ow)
    v_ = tx
    for i in range(ow):
        X[i] = v_
        v_ += a
    Y = np.empty(oh)
    v_ = ty
    for i in range(oh):
        Y[i] = v_
        v_ += d
    img = src.astype(np.float64)
    acc = np.zeros((oh, ow, 3))
    wsum = np.zeros((oh, ow, 1))
    cx, cy = np.floor(X).astype(int), np.floor(Y).astype(int)
    for ky in range(1 - S, S + 1):
        wy = weight(cy + ky - Y)[:, None, None]
        for kx in range(1 - S, S + 1):
            wx = weight(cx + kx - X)[None, :, None]
            iy = np.clip(cy + ky, 0, 28)[:, None]
            ix = np.clip(cx + kx, 0, 40)[None, :]
            acc += img[iy, ix] * (wx * wy)
            wsum += wx * wy
    _assert_rounds_to(got, acc / wsum, f"kernel {kernel}")


# ------------------------------------------------------- shrink-on-load hint

def test_shrink_on_load_reference_cases():
    lib = O.lib()
    # thumbnail.rs:577-585: thresholds leave at least 2 for the final resize
    for common, f in [(3.99, 0), (4.0, 2), (7.99, 2), (8.0, 4), (15.99, 4), (16.0, 16)]:
        assert lib.orc_shrink_on_load_factor(common) == f
    # thumbnail.rs:532-536: 4000x3000 -> Width(99) shrinks 16 on load
    assert lib.orc_thumbnail_shrink_on_load(4000, 3000, O.WIDTH, 99, 0, 0) == 16
    # thumbnail.rs:553-564: the crop case asks for none
    assert lib.orc_thumbnail_shrink_on_load(400, 200, O.FITBOX, 120, 120, 1) == 0
    # thumbnail.rs:664-680: ForceExact at the input size asks for none
    assert lib.orc_thumbnail_shrink_on_load(1024, 777, O.FORCE_EXACT, 1024, 777, 0) == 0
    assert lib.orc_thumbnail_shrink_on_load(1, 5000, O.HEIGHT, 0, 10, 0) == 0  # 1-pixel branch


def test_library_shrink_on_load_matches_oracle():
    # host-only planner call of the library (no device work)
    import viprs_amd as v
    lib = O.lib()
    r = np.random.default_rng(77)
    for i in range(400):
        w, h = int(r.integers(1, 20000)), int(r.integers(1, 20000))
        target = int(r.integers(0, 4))
        tw, th = int(r.integers(1, 3000)), int(r.integers(1, 3000))
        crop = bool(i % 3 == 0)
        exp = lib.orc_thumbnail_shrink_on_load(w, h, target, tw, th, int(crop))
        assert v.thumbnail_shrink_on_load_factor(w, h, 3, target, tw, th, crop) == exp, (w, h, target, tw, th, crop)
