"""GPU parity through the C ABI: every result must equal the oracle's
(bit-exact for u8/u16, and for f32 on the reference-exact FP64 kernels)."""
import pathlib

import numpy as np
import pytest

import viprs_amd as v
from oracle import oracle as O
from tests.helpers import (float_image, gradient_noise_u8, grayscale_source, noise_u8, noise_u16,
                           rgb_source, smooth_grayscale_source, structural_source)

pytestmark = pytest.mark.gpu
FIX = pathlib.Path(__file__).resolve().parent / "golden" / "reference_fixtures"
FMT = {np.dtype(np.uint8): v.U8, np.dtype(np.uint16): v.U16, np.dtype(np.float32): v.F32}


def build(lib_pipeline, steps):
    for name, args in steps:
        getattr(lib_pipeline, name)(*args)
    return lib_pipeline


def run_both(ctx, src, steps, device=True):
    """Run the same builder chain on the GPU and in the oracle."""
    h, w, b = src.shape
    fmt = FMT[src.dtype]
    gp = build(v.Pipeline(ctx, w, h, b, fmt), steps)
    op = build(O.Pipeline(w, h, b, fmt), steps)
    exp = op.run(src)
    if device:
        dsrc = v.DeviceBuffer.from_numpy(ctx, src)
        ddst = gp.output_buffer()
        gp.run(dsrc, ddst)
        ctx.synchronize()
        got = ddst.numpy()
    else:
        got = gp.run_host(src)
    return got, exp


def assert_same(got, exp, what=""):
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    if not np.array_equal(got, exp):
        bad = np.argwhere(got != exp)
        y, x, b = bad[0]
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at (y={y}, x={x}, b={b}): "
                             f"gpu={got[y, x, b]} oracle={exp[y, x, b]}")


# ------------------------------------------------------------- fixtures

def test_reference_fixtures_on_gpu(ctx):
    cases = [
        ("reduceh_libvips/rgb_factor2_linear.bin", rgb_source(8, 6), [("reduce_h", (2.0, v.BILINEAR))]),
        ("reducev_libvips/rgb_factor2_linear.bin", rgb_source(6, 8), [("reduce_v", (2.0, v.BILINEAR))]),
        ("reduceh_libvips/grayscale_factor1_5_cubic.bin", grayscale_source(8, 6), [("reduce_h", (1.5, v.BICUBIC))]),
        ("libvips_parity_resample/test_reduceh_factor_1_5_lanczos3.bin", smooth_grayscale_source(8, 6),
         [("reduce_h", (1.5, v.LANCZOS3))]),
        ("libvips_parity_resample/test_reducev_factor_1_5_lanczos3.bin", smooth_grayscale_source(6, 8),
         [("reduce_v", (1.5, v.LANCZOS3))]),
        ("shrinkh/grayscale_factor_2.bin", structural_source(10, 16), [("shrink_h", (2,))]),
        ("shrinkv/grayscale_factor_2.bin", structural_source(10, 16), [("shrink_v", (2,))]),
        ("shrink_libvips/rgb_factor4.bin", rgb_source(16, 16), [("shrink_v", (4,)), ("shrink_h", (4,))]),
        ("reduce_libvips/edge_4x4_factor2_linear.bin", np.arange(16, dtype=np.uint8).reshape(4, 4, 1),
         [("reduce_v", (2.0, v.BILINEAR)), ("reduce_h", (2.0, v.BILINEAR))]),
    ]
    for rel, src, steps in cases:
        h, w, b = src.shape
        got = build(v.Pipeline(ctx, w, h, b), steps).run_host(src)
        assert np.array_equal(got.ravel(), np.fromfile(FIX / rel, np.uint8)), rel


def test_gauss_f32_fixture_on_gpu(ctx):
    exp = np.fromfile(FIX / "gauss_blur/uniform_sigma_1_5.bin", np.float32)
    got = v.Pipeline(ctx, 64, 64, 1, v.F32).gauss_blur(1.5).run_host(np.full((64, 64, 1), 37.5, np.float32))
    assert np.max(np.abs(got.ravel() - exp)) <= 1e-3


def test_float_and_resize_fixtures_on_gpu(ctx):
    # The fixtures that pin the oracle's ConvSep, float Gaussian and resize
    # (tests/test_oracle_pins.py), run through the device path at the
    # reference's tolerance, and equal to the oracle bit for bit.
    from tests.test_oracle_pins import (convsep_source, fractional_ramp, max_ulp, nonuniform_source,
                                        sigma1_mask)

    f32_cases = [
        ("convsep_libvips/gaussian_sigma_1_0.bin", convsep_source(9, 7),
         [("convsep", (list(sigma1_mask()), 1.0, 0.0))]),
        ("gauss_blur_libvips/nonuniform_5x3.bin", nonuniform_source(5, 3), [("gauss_blur", (1.0,))]),
        ("gauss_blur/fractional_ramp_sigma_1_0.bin", fractional_ramp(), [("gauss_blur", (1.0,))]),
        ("gauss_blur_libvips/uniform_field.bin", np.full((8, 8, 1), 37.5, np.float32), [("gauss_blur", (1.0,))]),
    ]
    for rel, src, steps in f32_cases:
        got, exp = run_both(ctx, src, steps)
        assert max_ulp(got, np.fromfile(FIX / rel, "<f4")) <= 1, rel
        assert_same(got, exp, rel)
    u8_cases = [
        ("libvips_parity_resample/test_resize_half_bilinear_rgb.bin", rgb_source(16, 12), 0),
        ("libvips_parity_resample/test_resize_height_one_rounding.bin", smooth_grayscale_source(100, 1), 0),
    ]
    for rel, src, tol in u8_cases:
        got, exp = run_both(ctx, src, [("resize", (0.5, 0.5, v.BILINEAR))])
        assert np.array_equal(got.ravel(), np.fromfile(FIX / rel, np.uint8)), rel
        assert_same(got, exp, rel)
    for rel, src in [("libvips_parity_convolution/test_gaussblur_u8_nonuniform_5x3.bin", grayscale_source(5, 3)),
                     ("libvips_parity_convolution/test_gaussblur_u8_rgb_8x8.bin", rgb_source(8, 8))]:
        got, exp = run_both(ctx, src, [("gauss_blur", (1.0,))])
        assert np.max(np.abs(got.ravel().astype(int) - np.fromfile(FIX / rel, np.uint8).astype(int))) <= 1, rel
        assert_same(got, exp, rel)


# ------------------------------------------------------- configurations

def test_resize_1024x768_lanczos3_factor_3_7(ctx):
    # Plain CPU-runnable resize: reduce_h(3.7) then reduce_v(3.7), Lanczos3.
    src = noise_u8(768, 1024, 3, 20241)
    got, exp = run_both(ctx, src, [("reduce_h", (3.7, v.LANCZOS3)), ("reduce_v", (3.7, v.LANCZOS3))])
    assert got.shape == (208, 277, 3)
    assert_same(got, exp, "reduce 3.7")


@pytest.mark.parametrize("seed", range(12))
def test_random_thumbnails(ctx, seed):
    r = np.random.default_rng(1000 + seed)
    w, h = int(r.integers(20, 1600)), int(r.integers(20, 1600))
    b = int(r.integers(1, 5))
    target = [v.FITBOX, v.WIDTH, v.HEIGHT][seed % 3]
    tw = int(r.integers(8, 500))
    src = noise_u8(h, w, b, seed)
    steps = [("thumbnail", (target, tw, tw, v.LANCZOS3))]
    try:
        O.Pipeline(w, h, b).thumbnail(target, tw, tw)
    except O.OracleError as e:
        assert e.code == O.EUNSUPPORTED
        with pytest.raises(v.VgpuError):
            v.Pipeline(ctx, w, h, b).thumbnail(target, tw, tw)
        return
    got, exp = run_both(ctx, src, steps)
    assert_same(got, exp, f"thumbnail {w}x{h}x{b} -> {tw}")


@pytest.mark.parametrize("kernel", [v.NEAREST, v.BILINEAR, v.BICUBIC, v.LANCZOS2, v.LANCZOS3])
def test_thumbnail_kernels_with_blur(ctx, kernel):
    src = gradient_noise_u8(731, 997, 3, 77)
    got, exp = run_both(ctx, src, [("thumbnail", (v.WIDTH, 256, 0, kernel)), ("gauss_blur", (0.8,))])
    assert_same(got, exp, f"thumbnail+blur kernel {kernel}")


def test_thumbnail_rgba_premultiplied_chain(ctx):
    src = noise_u8(613, 1207, 4, 99)
    src[..., 3] = np.where(np.random.default_rng(3).random((613, 1207)) < 0.1, 0, src[..., 3])
    got, exp = run_both(ctx, src, [("thumbnail", (v.FITBOX, 400, 400, v.LANCZOS3))])
    assert_same(got, exp, "rgba thumbnail")


def test_thumbnail_crop_reference_case(ctx):
    # thumbnail.rs:553-575: 400x200 -> FitBox 120x120 with crop
    src = gradient_noise_u8(200, 400, 3, 5)
    got, exp = run_both(ctx, src, [("thumbnail", (v.FITBOX, 120, 120, v.LANCZOS3, True))])
    assert got.shape == (120, 120, 3)
    assert_same(got, exp, "crop 400x200 -> 120x120")


@pytest.mark.parametrize("seed", range(10))
def test_random_thumbnail_crops(ctx, seed):
    r = np.random.default_rng(2000 + seed)
    w, h = int(r.integers(20, 2400)), int(r.integers(20, 2400))
    b = int(r.integers(1, 5))
    target = [v.FITBOX, v.FORCE_EXACT, v.WIDTH, v.HEIGHT, v.FITBOX][seed % 5]
    tw, th = int(r.integers(4, min(w, h))), int(r.integers(4, min(w, h)))
    src = noise_u8(h, w, b, seed)
    if b == 4:
        src[..., 3] = np.where(r.random((h, w)) < 0.1, 0, src[..., 3])
    steps = [("thumbnail", (target, tw, th, v.LANCZOS3, True))]
    try:
        O.Pipeline(w, h, b).thumbnail(target, tw, th, crop=True)
    except O.OracleError as e:
        assert e.code == O.EUNSUPPORTED
        with pytest.raises(v.VgpuError):
            v.Pipeline(ctx, w, h, b).thumbnail(target, tw, th, crop=True)
        return
    got, exp = run_both(ctx, src, steps)
    assert_same(got, exp, f"crop {w}x{h}x{b} target {target} -> {tw}x{th}")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_extract_area_chains(ctx, dtype):
    h, w, b = 301, 517, 3
    src = {np.uint8: noise_u8, np.uint16: noise_u16, np.float32: float_image}[dtype](h, w, b, 9)
    chains = [
        [("extract_area", (13, 7, 200, 100))],
        [("extract_area", (0, 0, w, h))],
        [("gauss_blur", (1.5,)), ("extract_area", (w - 40, h - 30, 40, 30))],
        [("extract_area", (100, 50, 300, 201)), ("reduce_h", (2.5, v.LANCZOS3)), ("reduce_v", (1.7, v.BICUBIC))],
        [("reduce_h", (1.3, v.LANCZOS3)), ("extract_area", (5, 0, 60, h)), ("shrink_v", (3, 0))],
    ]
    for steps in chains:
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, f"{np.dtype(dtype)} {steps}")
        got, exp = run_both(ctx, src, steps, device=False)
        assert_same(got, exp, f"{np.dtype(dtype)} host {steps}")


def test_extract_area_validation(ctx):
    for args in [(0, 0, 0, 5), (0, 0, 5, 0), (97, 0, 4, 5), (0, 96, 5, 5)]:
        with pytest.raises(v.VgpuError) as e:
            v.Pipeline(ctx, 100, 100, 3).extract_area(*args)
        assert e.value.status == v.api.E_INVALID_ARGUMENT


AFFINE_CASES = [
    # (w, h, bands, dtype, target, tw, th, kernel, crop)
    (1, 2048, 3, np.uint8, v.HEIGHT, 0, 400, v.LANCZOS3, False),     # thumbnail.rs:663-677
    (2048, 1, 4, np.uint8, v.WIDTH, 300, 0, v.BICUBIC, False),
    (1, 300, 4, np.uint8, v.HEIGHT, 0, 120, v.BILINEAR, False),      # SmallTile, alpha
    (1, 1, 3, np.uint8, v.FORCE_EXACT, 5, 7, v.LANCZOS3, False),
    (200, 1000, 3, np.uint8, v.FORCE_EXACT, 300, 300, v.LANCZOS3, False),
    (300, 200, 4, np.uint8, v.FORCE_EXACT, 700, 500, v.LANCZOS3, False),  # several 128 x 128 tiles
    (300, 200, 2, np.uint8, v.FORCE_EXACT, 333, 271, v.NEAREST, False),
    (257, 190, 3, np.uint16, v.FORCE_EXACT, 600, 401, v.BICUBIC, False),
    (257, 190, 1, np.float32, v.FORCE_EXACT, 600, 401, v.BILINEAR, False),
    (700, 90, 3, np.uint8, v.FORCE_EXACT, 500, 333, v.LANCZOS3, False),   # ThinStrip, ReduceH + Affine
    (300, 200, 3, np.uint8, v.FORCE_EXACT, 500, 500, v.LANCZOS3, True),   # crop of an upscale
    (300, 200, 4, np.uint8, v.FITBOX, 100, 150, v.LANCZOS3, True),
    (3000, 2, 3, np.uint8, v.FORCE_EXACT, 12000, 4, v.BICUBIC, False),  # 1-row strips beyond 10000 px
    # u8 ReduceH ahead of a SmallTile Affine: one run per 128-column tile column
    (400, 300, 1, np.uint8, v.FORCE_EXACT, 150, 900, v.LANCZOS3, False),
    (500, 120, 4, np.uint8, v.FORCE_EXACT, 300, 260, v.BICUBIC, False),
    (511, 97, 2, np.uint8, v.FORCE_EXACT, 200, 300, v.BILINEAR, False),
]


@pytest.mark.parametrize("case", range(len(AFFINE_CASES)))
def test_thumbnail_affine_tail(ctx, case):
    w, h, b, dtype, target, tw, th, kernel, crop = AFFINE_CASES[case]
    src = {np.uint8: noise_u8, np.uint16: noise_u16, np.float32: float_image}[dtype](h, w, b, 40 + case)
    if b in (2, 4) and dtype == np.uint8:
        src[..., -1] = np.where(np.random.default_rng(case).random((h, w)) < 0.2, 0, src[..., -1])
    got, exp = run_both(ctx, src, [("thumbnail", (target, tw, th, kernel, crop))])
    assert_same(got, exp, f"affine tail {AFFINE_CASES[case]}")


@pytest.mark.parametrize("seed", range(10))
def test_random_affine_tails(ctx, seed):
    r = np.random.default_rng(3000 + seed)
    w, h = int(r.integers(2, 1200)), int(r.integers(2, 1200))
    b = int(r.integers(1, 5))
    tw, th = int(r.integers(2, 1600)), int(r.integers(2, 1600))
    kernel = [v.LANCZOS3, v.BILINEAR, v.NEAREST, v.BICUBIC][seed % 4]
    crop = seed % 3 == 0
    src = noise_u8(h, w, b, seed)
    got, exp = run_both(ctx, src, [("thumbnail", (v.FORCE_EXACT, tw, th, kernel, crop))])
    assert_same(got, exp, f"ForceExact {w}x{h}x{b} -> {tw}x{th} kernel {kernel} crop {crop}")


def test_affine_tail_limits(ctx):
    # an op after the Affine: declined when the pipeline runs
    p = v.Pipeline(ctx, 1, 700, 3).thumbnail(v.HEIGHT, 0, 300).gauss_blur(1.0)
    src = v.DeviceBuffer.from_numpy(ctx, noise_u8(700, 1, 3, 1))
    with pytest.raises(v.VgpuError) as e:
        p.run(src, p.output_buffer())
    assert e.value.status == v.api.E_UNSUPPORTED


@pytest.mark.parametrize("hs,vs,kernel", [(0.5, 0.5, v.BILINEAR), (0.3, 0.41, v.LANCZOS3),
                                          (0.137, 0.09, v.BICUBIC), (1.0, 0.25, v.LANCZOS2)])
def test_resize_downscale(ctx, hs, vs, kernel):
    src = noise_u8(555, 777, 3, 5)
    got, exp = run_both(ctx, src, [("resize", (hs, vs, kernel))])
    assert_same(got, exp, f"resize {hs} {vs}")


@pytest.mark.parametrize("sigma,bands", [(0.8, 3), (2.0, 4), (1.5, 1), (8.0, 3), (0.1, 2)])
def test_gauss_u8(ctx, sigma, bands):
    src = noise_u8(301, 403, bands, 31)
    got, exp = run_both(ctx, src, [("gauss_blur", (sigma,))])
    assert_same(got, exp, f"gauss {sigma}")


@pytest.mark.parametrize("rows", [0, 16, 32, -1])
@pytest.mark.parametrize("bands", [3, 4])
def test_gauss_u8_forms(ctx, bands, rows):
    # 3/4-band u8 GaussBlurH -> GaussBlurV: VALU tiles of 16 or 32 rows
    # (0 = chosen by the planner; radius <= 6 and scale <= 256), else
    # (rows = -1, or a wider mask) the int8-limb MFMA kernel.  Ragged and tiny sizes, and a Gaussian after a
    # thumbnail (its input is an intermediate with Z^2 halo rows).
    for h, w in ((301, 403), (70, 64), (1, 1), (2, 3), (5, 130), (129, 1)):
        src = noise_u8(h, w, bands, 31 + h + w)
        for sigma in (0.8, 1.5, 2.0, 3.0, 4.0):
            gp = v.Pipeline(ctx, w, h, bands).gauss_blur(sigma).set_option(v.api.OPT_GAUSS_ROWS, rows)
            exp = O.Pipeline(w, h, bands).gauss_blur(sigma).run(src)
            assert_same(gp.run_host(src), exp, f"{w}x{h}x{bands} sigma {sigma} rows {rows}")
            if rows >= 0 and sigma <= 3.0:
                assert gp.stage_kernels() == ["gauss_u8"], gp.stage_kernels()
    src = noise_u8(517, 777, bands, 5)
    steps = [("thumbnail", (v.WIDTH, 300, 0, v.LANCZOS3)), ("gauss_blur", (0.8,))]
    gp = build(v.Pipeline(ctx, 777, 517, bands), steps).set_option(v.api.OPT_GAUSS_ROWS, rows)
    exp = build(O.Pipeline(777, 517, bands), steps).run(src)
    assert_same(gp.run_host(src), exp, f"thumbnail + blur rows {rows}")


@pytest.mark.parametrize("f", [1, 2, 3, 5, 7, 19, 40])
@pytest.mark.parametrize("bands", [1, 2, 3, 4, 5])
def test_shrink_factors_and_bands(ctx, f, bands):
    src = noise_u8(97, 211, bands, f * 10 + bands)
    got, exp = run_both(ctx, src, [("shrink_h", (f,)), ("shrink_v", (f,))])
    assert_same(got, exp, f"shrink {f} bands {bands}")
    got, exp = run_both(ctx, src, [("shrink_v", (f,)), ("shrink_h", (f, True))])
    assert_same(got, exp, f"shrink v/h ceil {f} bands {bands}")


@pytest.mark.parametrize("hs,vs", [(2, 2), (3, 2), (4, 4), (1, 5), (7, 1)])
@pytest.mark.parametrize("bands", [1, 3, 4])
def test_builder_shrink_both_axes(ctx, hs, vs, bands):
    # PipelineBuilder::shrink(h, v): Shrink<F> (shrink.rs:37-293) = ShrinkH then
    # ShrinkV in one op.  For u8 2x2 the reference's direct form is
    # ((floor((a+b+1)/2) + floor((c+d+1)/2) + 1) >> 1) (shrink.rs:225-293).
    src = noise_u8(101, 203, bands, hs * 10 + vs + bands)
    got, exp = run_both(ctx, src, [("shrink", (hs, vs))])
    assert_same(got, exp, f"shrink({hs}, {vs}) bands {bands}")
    if (hs, vs) == (2, 2):
        s = src.astype(np.int32)
        a, b_, c, d = s[0:100:2, 0:202:2], s[0:100:2, 1:202:2], s[1:100:2, 0:202:2], s[1:100:2, 1:202:2]
        direct = ((((a + b_ + 1) >> 1) + ((c + d + 1) >> 1) + 1) >> 1).astype(np.uint8)
        assert np.array_equal(got[:50, :101], direct)
    for src16 in (noise_u16(64, 96, bands, 3),):
        got, exp = run_both(ctx, src16, [("shrink", (hs, vs))])
        assert_same(got, exp, f"u16 shrink({hs}, {vs}) bands {bands}")


def test_u16_paths(ctx):
    src = noise_u16(211, 307, 3, 8)
    for steps in ([("reduce_h", (2.7, v.LANCZOS3)), ("reduce_v", (1.9, v.BICUBIC))],
                  [("shrink_h", (3,)), ("shrink_v", (2,))],
                  [("thumbnail", (v.WIDTH, 100, 0, v.LANCZOS3))],
                  [("gauss_blur", (1.2,))]):
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, f"u16 {steps}")


def test_f32_reference_exact(ctx):
    # The f32 kernels carry the reference's FP64 accumulation in its
    # operation order: bit-identical results, for 1 band (AVX2 grouping) and
    # 3 bands (scalar FMA chain), including the ReduceV vector/tail split.
    for b in (1, 3):
        src = float_image(203, 257, b, 12 + b)
        for steps in ([("reduce_h", (2.5, v.LANCZOS3)), ("reduce_v", (2.5, v.LANCZOS3))],
                      [("gauss_blur", (8.0,))],
                      [("shrink_h", (3,)), ("shrink_v", (3,))],
                      [("convsep", ([1.0, 2.0, 3.0, 2.0, 1.0], 9.0, 0.5))]):
            got, exp = run_both(ctx, src, steps)
            assert_same(got, exp, f"f32 bands {b} {steps}")


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("bands", [1, 2, 3, 4])
def test_f32_reduce_forms_exact(ctx, form, bands):
    # Both forms of the fused f32 reduce -- blocks walking down strips of
    # output columns (0) and one block per output tile (1) -- are
    # bit-identical to the oracle in the reference's order: ragged strips and
    # row runs, windows clamped at the top and bottom, one phase (2.0), two
    # (2.5) and many, H taps held in registers (<= 24) and in the LDS table
    # (31 at 5.0); the plain FP32 mode gives the same bits in both forms.
    from viprs_amd.api import OPT_REDUCE_FORM
    for (w, h), steps in (((613, 389), [("reduce_h", (2.5, v.LANCZOS3)), ("reduce_v", (2.5, v.LANCZOS3))]),
                          ((613, 389), [("reduce_h", (3.7, v.LANCZOS3)), ("reduce_v", (1.3, v.BICUBIC))]),
                          ((47, 1201), [("reduce_h", (1.1, v.BILINEAR)), ("reduce_v", (6.2, v.LANCZOS2))]),
                          ((1733, 29), [("reduce_h", (2.0, v.LANCZOS3)), ("reduce_v", (2.0, v.LANCZOS3))]),
                          ((1211, 301), [("reduce_h", (5.0, v.LANCZOS3)), ("reduce_v", (2.2, v.LANCZOS3))]),
                          ((9, 7), [("reduce_h", (1.5, v.LANCZOS3)), ("reduce_v", (1.5, v.LANCZOS3))])):
        src = float_image(w, h, bands, 70 + bands + w)
        exp = build(O.Pipeline(w, h, bands, v.F32), steps).run(src)
        got = build(v.Pipeline(ctx, w, h, bands, v.F32), steps).set_option(OPT_REDUCE_FORM, form).run_host(src)
        assert_same(got, exp, f"form {form} bands {bands} {w}x{h} {steps}")
        f32 = [build(v.Pipeline(ctx, w, h, bands, v.F32), steps).set_float_mode(v.Pipeline.FLOAT_FP32)
               .set_option(OPT_REDUCE_FORM, f).run_host(src) for f in (0, 1)]
        assert_same(f32[0], f32[1], f"fp32 forms bands {bands} {w}x{h} {steps}")
        assert np.abs(f32[0].astype(np.float64) - exp).max() <= 2e-6


@pytest.mark.parametrize("fmt", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("bands", [1, 2, 3, 4, 5])
def test_reduce_stage_forms(ctx, fmt, bands):
    # Lone ReduceH and lone ReduceV run through the fused reduce kernel with
    # an identity axis for the missing pass; V-then-H order gives two such
    # stages; five bands take the unfused per-pass kernels.
    if fmt == np.float32:
        src = float_image(157, 233, bands, 40 + bands)
    elif fmt == np.uint16:
        src = noise_u16(157, 233, bands, 40 + bands)
    else:
        src = noise_u8(157, 233, bands, 40 + bands)
    for steps in ([("reduce_h", (2.3, v.LANCZOS3))],
                  [("reduce_v", (3.1, v.LANCZOS3))],
                  [("reduce_v", (1.6, v.BICUBIC)), ("reduce_h", (4.4, v.LANCZOS2))],
                  [("reduce_h", (1.4, v.BILINEAR)), ("reduce_v", (5.9, v.LANCZOS3))]):
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, f"{np.dtype(fmt)} bands {bands} {steps}")


def reduce_op_source(fmt, h, w, b, seed):
    if fmt == np.float32:
        return float_image(h, w, b, seed)
    if fmt == np.uint16:
        return noise_u16(h, w, b, seed)
    return noise_u8(h, w, b, seed)


@pytest.mark.parametrize("fmt", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("bands", [1, 2, 3, 4, 5])
def test_reduce_op(ctx, fmt, bands):
    # PipelineBuilder::reduce: one fused f64 stage for <= 4 bands, the two
    # per-pass kernels for 5.
    src = reduce_op_source(fmt, 141, 203, bands, 60 + bands)
    for hf, vf, k in ((1.5, 2.5, v.LANCZOS3), (2.0, 2.0, v.BILINEAR), (1.0, 3.3, v.BICUBIC),
                      (4.7, 1.0, v.LANCZOS2), (3.0, 3.0, v.NEAREST), (7.9, 6.1, v.CATMULLROM)):
        steps = [("reduce", (hf, vf, k))]
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, f"{np.dtype(fmt)} bands {bands} {steps}")
    h, w, b = src.shape
    p = v.Pipeline(ctx, w, h, b, FMT[src.dtype]).reduce(1.5, 2.5, v.LANCZOS3)
    assert p.stage_kernels() == (["reduce_hv"] if bands <= 4 else ["reduce_h", "reduce_v"])


def test_reduce_op_reference_integration_case(ctx):
    # spatial_integration.rs:199-232: 777x333x3, sample (i * 31 + 11) % 256,
    # reduce(1.5, 2.5, Lanczos3).
    src = ((np.arange(777 * 333 * 3) * 31 + 11) % 256).astype(np.uint8).reshape(333, 777, 3)
    got, exp = run_both(ctx, src, [("reduce", (1.5, 2.5, v.LANCZOS3))])
    assert got.shape == (133, 518, 3)
    assert_same(got, exp, "integration case")


def test_reduce_op_in_chains_and_fixtures(ctx):
    rgb4 = np.array([0, 10, 20, 30, 45, 60, 90, 60, 30, 120, 90, 40, 15, 35, 55, 45, 75, 105, 95, 65,
                     35, 125, 95, 45, 25, 50, 75, 55, 85, 115, 105, 70, 40, 135, 100, 50, 35, 60, 85,
                     65, 95, 125, 115, 80, 45, 145, 105, 55], np.uint8).reshape(4, 4, 3)
    got, exp = run_both(ctx, rgb4, [("reduce", (2.0, 2.0, v.BILINEAR))])
    assert_same(got, exp, "reduce.rs 4x4x3 fixture")
    src = noise_u8(300, 401, 3, 77)
    for steps in ([("shrink", (2, 2)), ("reduce", (1.7, 1.3, v.LANCZOS3)), ("gauss_blur", (0.8,))],
                  [("reduce", (2.2, 2.2, v.LANCZOS3)), ("reduce_h", (1.5, v.BICUBIC))],
                  [("gauss_blur", (1.5,)), ("reduce", (3.0, 1.5, v.BILINEAR))],
                  [("reduce", (1.5, 1.5, v.LANCZOS3)), ("reduce", (1.5, 1.5, v.LANCZOS3))]):
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, str(steps))
    got, exp = run_both(ctx, noise_u8(1, 1, 4, 3), [("reduce", (2.0, 2.0, v.LANCZOS3))])
    assert_same(got, exp, "1x1")


def test_reduce_op_validation(ctx):
    with pytest.raises(v.VgpuError) as e:
        v.Pipeline(ctx, 8, 8, 1).reduce(0.5, 2.0, v.BILINEAR)
    assert e.value.status == v.api.E_INVALID_ARGUMENT
    with pytest.raises(v.VgpuError) as e:
        v.Pipeline(ctx, 8, 8, 1).reduce(0.5, 2.0, v.LBB)  # factors are checked first
    assert e.value.status == v.api.E_INVALID_ARGUMENT
    with pytest.raises(v.VgpuError) as e:
        v.Pipeline(ctx, 8, 8, 1).reduce(2.0, 2.0, v.NOHALO)
    assert e.value.status == v.api.E_INVALID_KERNEL
    with pytest.raises(v.VgpuError) as e:
        v.Pipeline(ctx, 8, 8, 1).reduce(2.0, 400.0, v.LANCZOS3)
    assert e.value.status == v.api.E_INVALID_ARGUMENT


@pytest.mark.parametrize("bands", [2, 3, 4])
def test_alpha_fusion_forms(ctx, bands):
    # Premultiply/Unpremultiply fuse into the shrink and reduce stages for 2
    # and 4 bands; 3 bands keep them as separate pixel stages.
    src = noise_u8(203, 311, bands, 70 + bands)
    src[..., -1] = np.where(np.random.default_rng(bands).random((203, 311)) < 0.15, 0, src[..., -1])
    for steps in ([("premultiply", ()), ("shrink_h", (3,)), ("shrink_v", (2,)), ("unpremultiply", ())],
                  [("premultiply", ()), ("shrink_v", (4,)), ("unpremultiply", ())],
                  [("premultiply", ()), ("reduce_h", (2.2, v.LANCZOS3)), ("reduce_v", (2.2, v.LANCZOS3)),
                   ("unpremultiply", ())],
                  [("premultiply", ()), ("reduce_v", (1.7, v.LANCZOS3)), ("unpremultiply", ())],
                  [("premultiply", ()), ("shrink_h", (2,)), ("shrink_v", (2,)), ("reduce_h", (1.3, v.LANCZOS3)),
                   ("reduce_v", (1.3, v.LANCZOS3)), ("unpremultiply", ())]):
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, f"bands {bands} {steps}")


@pytest.mark.parametrize("bands", [2, 4])
@pytest.mark.parametrize("f", [4, 7, 19, 40])
def test_premultiplied_shrink_large_factors(ctx, bands, f):
    # Wide box windows take the lean shrink form; premultiplied 4-band windows
    # inside the span sum on packed 16-bit lanes, the rest per pixel.  Ragged
    # sizes put a chunk cut by the end of the buffer into the last row.
    src = noise_u8(3 * f + 5, 64 * f + 7, bands, 300 + f + bands)
    src[..., -1] = np.where(np.random.default_rng(f).random(src.shape[:2]) < 0.15, 0, src[..., -1])
    src[: f, : f, -1] = 255
    for steps in ([("premultiply", ()), ("shrink_h", (f,)), ("shrink_v", (f,)), ("unpremultiply", ())],
                  [("premultiply", ()), ("shrink_h", (f, True)), ("shrink_v", (max(f // 3, 1), True)),
                   ("unpremultiply", ())],
                  [("premultiply", ()), ("shrink_h", (f,)), ("shrink_v", (f,))]):
        got, exp = run_both(ctx, src, steps)
        assert_same(got, exp, f"bands {bands} f {f} {steps}")
    big = noise_u8(1531, 2049, bands, 7 + f)
    big[..., -1] = np.where(np.random.default_rng(f + 1).random(big.shape[:2]) < 0.2, 0, big[..., -1])
    got, exp = run_both(ctx, big, [("thumbnail", (v.FITBOX, 2049 // f, 2049 // f, v.LANCZOS3))])
    assert_same(got, exp, f"rgba thumbnail f {f}")


U8_PAIR_CASES = [
    [("reduce_h", (3.7, v.LANCZOS3)), ("reduce_v", (3.7, v.LANCZOS3))],
    [("reduce_h", (1.0779, v.LANCZOS3)), ("reduce_v", (1.9, v.LANCZOS3))],
    [("reduce_v", (2.6, v.BICUBIC)), ("reduce_h", (1.2, v.LANCZOS2))],
    [("reduce_h", (5.5, v.BILINEAR))],
    [("reduce_v", (4.2, v.LANCZOS3))],
    [("gauss_blur", (0.8,))],
    [("gauss_blur", (2.0,))],
    [("gauss_blur", (4.5,))],
    [("thumbnail", (v.FITBOX, 150, 150, v.LANCZOS3)), ("gauss_blur", (0.8,))],
]


@pytest.mark.parametrize("form", ["tile", "wave", "plain"])
@pytest.mark.parametrize("bands", [1, 2, 3, 4])
def test_u8_pairs_mfma_and_plain(ctx, form, bands):
    # u8 ReduceH -> ReduceV and GaussBlurH -> GaussBlurV run on the int8-limb
    # MFMA kernel by default (block tiles staged in LDS, or one wavefront per
    # strip task) and on the plain integer kernels when it is switched off;
    # every form must equal the oracle bit for bit, edges included.
    src = noise_u8(211, 389, bands, 60 + bands)
    if bands in (2, 4):
        src[..., -1] = np.where(np.random.default_rng(bands).random((211, 389)) < 0.1, 0, src[..., -1])
    cases = list(U8_PAIR_CASES)
    if bands in (2, 4):
        cases.append([("premultiply", ()), ("reduce_h", (1.7, v.LANCZOS3)), ("reduce_v", (1.7, v.LANCZOS3)),
                      ("unpremultiply", ())])
    for steps in cases:
        h, w, b = src.shape
        gp = build(v.Pipeline(ctx, w, h, b), steps).set_integer_mfma(form != "plain")
        if form == "wave":
            gp.set_option(v.api.OPT_SEP_FORM, 1)
        exp = build(O.Pipeline(w, h, b), steps).run(src)
        assert_same(gp.run_host(src), exp, f"form={form} bands {bands} {steps}")


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("bands", [1, 2, 3, 4])
def test_box_shrink_on_mfma(ctx, bands, form):
    # The box pre-shrink as an MFMA contraction (opt-in): 0/1 weights with edge
    # counts folded, the reference's per-factor roundings, premultiply fused.
    src = noise_u8(173, 311, bands, 90 + bands)
    cases = [[("shrink_h", (f,)), ("shrink_v", (f,))] for f in (2, 3, 5, 7, 19)]
    cases += [[("shrink_h", (4,))], [("shrink_v", (6,))], [("shrink_v", (3,)), ("shrink_h", (5, True))],
              [("thumbnail", (v.WIDTH, 40, 0, v.LANCZOS3))]]
    if bands in (2, 4):
        cases.append([("premultiply", ()), ("shrink_h", (3,)), ("shrink_v", (4,)), ("unpremultiply", ())])
    for steps in cases:
        h, w, b = src.shape
        gp = build(v.Pipeline(ctx, w, h, b), steps).set_integer_mfma(True, shrink=True)
        gp.set_option(v.api.OPT_SEP_FORM, form)
        exp = build(O.Pipeline(w, h, b), steps).run(src)
        assert_same(gp.run_host(src), exp, f"mfma shrink bands {bands} {steps}")


@pytest.mark.parametrize("w,h,b", [(1, 1, 3), (2, 1, 1), (1, 5, 4), (7, 3, 2), (19, 19, 3), (20, 1000, 1)])
def test_tiny_and_ragged(ctx, w, h, b):
    src = noise_u8(h, w, b, w * 31 + h)
    for steps in ([("reduce_h", (1.7, v.LANCZOS3))], [("reduce_v", (2.2, v.BICUBIC))],
                  [("gauss_blur", (2.0,))], [("shrink_h", (2,)), ("shrink_v", (3, True))]):
        try:
            O.Pipeline(w, h, b)
            got, exp = run_both(ctx, src, steps, device=False)
        except O.OracleError:
            continue
        assert_same(got, exp, f"{w}x{h}x{b} {steps}")


# ------------------------------------------------------------ op level

def test_op_process_region_matches_oracle_reduce_h_offsets(ctx):
    src = noise_u8(9, 300, 3, 4)
    op = v.Op.reduce_h(ctx, 2.3, v.LANCZOS3, 3, 300)
    for x0 in (-5, 0, 17):
        out_r = v.Region(x0, 0, 100, 9)
        in_r = op.required_input_region(out_r)
        # Materialise the input tile with clamp-to-edge like MemorySource.
        xs = np.clip(np.arange(in_r.x, in_r.x + in_r.width), 0, 299)
        tile = np.ascontiguousarray(src[:, xs, :])
        dst = v.DeviceBuffer(ctx, (9, 100, 3), np.uint8)
        op.process_region(v.DeviceBuffer.from_numpy(ctx, tile), dst, in_r, out_r)
        ctx.synchronize()
        got = dst.numpy()
        if x0 == 0:
            exp = O.Pipeline(300, 9, 3).reduce_h(2.3).run(src)[:, :100]
            assert np.array_equal(got, exp)
        # Closed form of one ReduceH tile: accumulated u8 plan from the tile
        # origin, taps clamped into the tile (reduceh.rs:101-163).
        starts, phases = O.reduce_plan(O.LANCZOS3, 2.3, 300, float(x0), True, 100)
        _, ci = O.reduce_tables(O.LANCZOS3, 2.3)
        cols = np.clip(starts[:, None] + np.arange(ci.shape[1])[None, :] - in_r.x, 0, in_r.width - 1)
        acc = np.einsum("xt,yxtb->yxb", ci[phases].astype(np.int64), tile[:, cols, :].astype(np.int64))
        assert np.array_equal(got, np.clip((acc + 2048) >> 12, 0, 255).astype(np.uint8))


@pytest.mark.parametrize("fmt,bands", [(np.uint8, 3), (np.uint16, 1), (np.float32, 4), (np.uint8, 5)])
def test_op_reduce_bridge_tiles_match_whole_image(ctx, fmt, bands):
    # ReduceBridge::dyn_process_region on strips and an interior tile, each
    # input tile read from the source with clamp-to-edge as MemorySource
    # does: every tile equals the oracle's whole-image ReduceOp over it.
    src = reduce_op_source(fmt, 120, 170, bands, 90 + bands)
    h, w, b = src.shape
    hf, vf, k = 1.5, 2.5, v.LANCZOS3
    exp = O.Pipeline(w, h, b, FMT[src.dtype]).reduce(hf, vf, k).run(src)
    op = v.Op.reduce(ctx, hf, vf, k, b, FMT[src.dtype])
    ow, oh = op.output_size(w, h)
    assert (oh, ow) == exp.shape[:2]
    for out_r in (v.Region(0, 0, ow, oh), v.Region(0, 0, ow, 16), v.Region(0, 32, ow, 16),
                  v.Region(7, 40, 60, 8), v.Region(0, oh - 3, ow, 3)):
        in_r = op.required_input_region(out_r)
        ys = np.clip(np.arange(in_r.y, in_r.y + in_r.height), 0, h - 1)
        xs = np.clip(np.arange(in_r.x, in_r.x + in_r.width), 0, w - 1)
        tile = np.ascontiguousarray(src[ys][:, xs, :])
        dst = v.DeviceBuffer(ctx, (out_r.height, out_r.width, b), src.dtype)
        op.process_region(v.DeviceBuffer.from_numpy(ctx, tile), dst, in_r, out_r)
        ctx.synchronize()
        want = exp[out_r.y:out_r.y + out_r.height, out_r.x:out_r.x + out_r.width]
        assert_same(dst.numpy(), want, f"tile {out_r.x},{out_r.y},{out_r.width}x{out_r.height}")
    # The input tile's rows must be the scratch rows.
    out_r = v.Region(0, 0, ow, 16)
    in_r = op.required_input_region(out_r)
    bad = v.Region(in_r.x, in_r.y + 1, in_r.width, in_r.height)
    tile = v.DeviceBuffer(ctx, (in_r.height, in_r.width, b), src.dtype)
    with pytest.raises(v.VgpuError):
        op.process_region(tile, v.DeviceBuffer(ctx, (16, ow, b), src.dtype), bad, out_r)


def test_op_premultiply_exhaustive(ctx):
    a = np.repeat(np.arange(256, dtype=np.uint8), 256)
    c = np.tile(np.arange(256, dtype=np.uint8), 256)
    img = np.stack([c, 255 - c, c // 2, a], axis=-1).reshape(256, 256, 4)
    for kind in ("premultiply", "unpremultiply"):
        op = getattr(v.Op, kind)(ctx, 4)
        r = v.Region(0, 0, 256, 256)
        dst = v.DeviceBuffer(ctx, (256, 256, 4), np.uint8)
        op.process_region(v.DeviceBuffer.from_numpy(ctx, img), dst, r, r)
        ctx.synchronize()
        exp = getattr(O.Pipeline(256, 256, 4), kind)().run(img)
        assert np.array_equal(dst.numpy(), exp), kind


def test_op_rejects_bad_lengths_and_regions(ctx):
    op = v.Op.gauss_blur_h(ctx, 2.0, 3)
    out_r = v.Region(0, 0, 10, 2)
    in_r = op.required_input_region(out_r)
    assert (in_r.x, in_r.width) == (-3, 16)
    src = v.DeviceBuffer(ctx, (2, 16, 3), np.uint8)
    with pytest.raises(v.VgpuError) as e:
        op.process_region(src, v.DeviceBuffer(ctx, (2, 9, 3), np.uint8), in_r, out_r)
    assert e.value.status == v.api.E_BUFFER_SIZE
    with pytest.raises(v.VgpuError):
        op.process_region(src, v.DeviceBuffer(ctx, (2, 10, 3), np.uint8), v.Region(0, 0, 16, 2), out_r)


def test_builder_validation_errors(ctx):
    p = v.Pipeline(ctx, 100, 100, 3)
    with pytest.raises(v.VgpuError) as e:
        p.reduce_h(0.5, v.LANCZOS3)
    assert e.value.status == v.api.E_INVALID_ARGUMENT
    with pytest.raises(v.VgpuError) as e:
        p.reduce_h(2.0, v.LBB)
    assert e.value.status == v.api.E_INVALID_KERNEL
    with pytest.raises(v.VgpuError):
        p.reduce_v(400.0, v.LANCZOS3)  # over the 2000-tap limit
    with pytest.raises(v.VgpuError):
        v.Pipeline(ctx, 10, 10, 0)


RESIZE_UP = [
    # (w, h, bands, dtype, hscale, vscale, kernel)
    (100, 80, 3, np.uint8, 2.0, 3.0, v.LANCZOS3),      # Zoom alone
    (100, 80, 3, np.uint8, 2.5, 1.3, v.LANCZOS3),      # Zoom + Affine, 6 taps
    (64, 64, 1, np.uint8, 1.7, 1.7, v.NEAREST),
    (77, 51, 4, np.uint8, 3.3, 2.2, v.BILINEAR),       # alpha sampled premultiplied
    (90, 70, 3, np.uint8, 1.9, 2.6, v.BICUBIC),
    (90, 70, 2, np.uint8, 1.45, 1.45, v.LANCZOS2),
    (90, 70, 3, np.uint8, 2.2, 1.6, v.CATMULLROM),
    (300, 200, 3, np.uint8, 1.0, 2.4, v.LANCZOS3),      # one axis only
    (300, 200, 3, np.uint8, 1.3, 0.5, v.LANCZOS3),      # ReduceV then the Affine
    (120, 90, 3, np.uint16, 2.3, 1.7, v.LANCZOS3),
    (120, 90, 1, np.float32, 2.3, 1.7, v.BICUBIC),
    (120, 90, 4, np.float32, 1.6, 1.6, v.LANCZOS3),
]


@pytest.mark.parametrize("case", range(len(RESIZE_UP)))
def test_resize_upscale(ctx, case):
    w, h, b, dtype, hs, vs, kernel = RESIZE_UP[case]
    src = {np.uint8: noise_u8, np.uint16: noise_u16, np.float32: float_image}[dtype](h, w, b, 60 + case)
    if b == 4:
        src[..., 3] = np.where(np.random.default_rng(case).random((h, w)) < 0.2, 0, src[..., 3])
    got, exp = run_both(ctx, src, [("resize", (hs, vs, kernel))])
    assert_same(got, exp, f"resize up {RESIZE_UP[case]}")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_resize_mixed_down_and_up(ctx, dtype):
    # resize.rs:441-457: ReduceH(2), Zoom(1, 2), Affine on 100x100 -> 50x250;
    # u8 runs one 128-column tile column at a time
    src = (noise_u8 if dtype == np.uint8 else noise_u16)(100, 100, 3, 1)
    got, exp = run_both(ctx, src, [("resize", (0.5, 2.5, v.BILINEAR))])
    assert got.shape == (250, 50, 3)
    assert_same(got, exp, f"mixed resize {np.dtype(dtype)}")
    src = (noise_u8 if dtype == np.uint8 else noise_u16)(300, 450, 4, 2)
    got, exp = run_both(ctx, src, [("resize", (0.61, 1.9, v.LANCZOS3))])
    assert_same(got, exp, f"mixed resize 450x300 {np.dtype(dtype)}")
    got, exp = run_both(ctx, src, [("resize", (0.61, 1.9, v.LANCZOS3))], device=False)
    assert_same(got, exp, f"mixed resize host {np.dtype(dtype)}")


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_op_convsep_pair_matches_pipeline(ctx, dtype):
    # ConvSepH then ConvSepV as operations over the regions the backward walk
    # gives them (convsep.rs:53-207), equal to the oracle's builder convsep
    h, w, b = 45, 61, 3
    src = noise_u8(h, w, b, 12) if dtype == np.uint8 else float_image(h, w, b, 12)
    coeffs, scale, offset = [1.0, -2.0, 5.0, -2.0, 1.0], 3.0, 0.25
    r = len(coeffs) // 2
    exp = O.Pipeline(w, h, b, FMT[np.dtype(dtype)]).convsep(coeffs, scale, offset).run(src)
    opH = v.Op.convsep_h(ctx, coeffs, scale, offset, b, FMT[np.dtype(dtype)])
    opV = v.Op.convsep_v(ctx, coeffs, scale, 0.0, b, v.F32)
    out_r = v.Region(0, 0, w, h)
    mid_r = opV.required_input_region(out_r)
    in_r = opH.required_input_region(mid_r)
    assert (in_r.x, in_r.y, in_r.width, in_r.height) == (-r, -r, w + 2 * r, h + 2 * r)
    # the source read of that region clamps to the image (sources/memory.rs:149-235)
    tile = np.pad(src, ((r, r), (r, r), (0, 0)), mode="edge")
    d_in = v.DeviceBuffer.from_numpy(ctx, tile)
    d_mid = v.DeviceBuffer(ctx, (mid_r.height, mid_r.width, b), np.float32)
    d_out = v.DeviceBuffer(ctx, (h, w, b), np.float32)
    opH.process_region(d_in, d_mid, in_r, mid_r)
    opV.process_region(d_mid, d_out, mid_r, out_r)
    ctx.synchronize()
    assert_same(d_out.numpy(), exp, f"op convsep {np.dtype(dtype)}")
