"""Pins for the oracle's float, u16 and convolution functions (CPU only).

Two kinds of check:

* the reference's committed fixture files for ConvSep, the float Gaussian,
  the u8 Gaussian and resize, at the tolerance the reference's own tests
  hold them to (tests/functional/golden/convolution.rs:273-288,396-420;
  tests/functional/golden/libvips/resample.rs:26-50;
  tests/functional/parity/convolution.rs:48-240; parity/resample.rs:1-90;
  parity/colour.rs:1-52; golden/legacy/transforms.rs:140-226;
  golden/legacy/resize.rs:233-265);
* closed forms independent of any summation order: every float output is
  compared with the EXACT real value of its weighted sum (rational
  arithmetic), widened by a rigorous bound on FP64 accumulation error.
  Wherever that interval rounds to a single f32, the oracle must produce
  exactly that f32, whatever order it adds in; a dropped tap, a wrong index
  or a wrong clamp moves the sum far outside the interval.  Integer formats
  are compared with exact integer arithmetic.
"""
from fractions import Fraction
import pathlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import (grayscale_source, noise_u16, rgb_source, smooth_grayscale_source)

FIX = pathlib.Path(__file__).resolve().parent / "golden" / "reference_fixtures"
U = Fraction(1, 2 ** 53)  # FP64 unit roundoff


def _ordered(a):
    b = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    return np.where(b < 0, -(b & 0x7FFFFFFF), b)


def max_ulp(got, exp):
    return int(np.max(np.abs(_ordered(got).ravel() - _ordered(exp).ravel())))


def fixture_f32(rel):
    return np.fromfile(FIX / rel, "<f4")


def fixture_u8(rel):
    return np.fromfile(FIX / rel, np.uint8)


# --------------------------------------------------------------- sources

def convsep_source(w, h):
    # golden/convolution.rs:128-134: f32 arithmetic in this order
    y, x = np.mgrid[0:h, 0:w]
    xf, yf, xy = x.astype(np.float32), y.astype(np.float32), (x * y).astype(np.float32)
    return (xf * np.float32(1.75) - yf * np.float32(2.25) + xy * np.float32(0.5)).astype(np.float32)[..., None]


def nonuniform_source(w, h):
    # golden/libvips/resample.rs:35-37; parity/support.rs:180-184
    y, x = np.mgrid[0:h, 0:w]
    xf, yf, xy = x.astype(np.float32), y.astype(np.float32), (x * y).astype(np.float32)
    return (xf * np.float32(1.75) - yf * np.float32(2.25) + xy).astype(np.float32)[..., None]


def fractional_ramp():
    # scripts/generate_golden_fixtures.sh:34-45: x*0.5 - y*0.75 - 3.125 on 8x8
    y, x = np.mgrid[0:8, 0:8]
    return (x * 0.5 - y * 0.75 - 3.125).astype(np.float32)[..., None]


def sigma1_mask():
    # gaussian_kernel_1d(1.0) (gauss_kernel.rs:13-54,57-65,115-117): integer
    # taps round(20 exp(-x^2/2)) = [12, 20, 12], normalised by their sum 44.
    return np.array([12.0 / 44.0, 20.0 / 44.0, 12.0 / 44.0])


# ------------------------------------------------------------- fixtures

def test_convsep_fixture_gaussian_sigma_1():
    # convsep_libvips_matches_gaussblur (golden/convolution.rs:396-420):
    # ConvSep::new(gaussian_kernel_1d(1.0)) on 9x7, held at <= 1 ULP.
    src = convsep_source(9, 7)
    out = O.Pipeline(9, 7, 1, O.F32).convsep(sigma1_mask(), 1.0, 0.0).run(src)
    exp = fixture_f32("convsep_libvips/gaussian_sigma_1_0.bin")
    assert out.size == exp.size == 63
    assert max_ulp(out, exp) <= 1


@pytest.mark.parametrize("rel,src,tol_ulp,tol_abs", [
    ("gauss_blur_libvips/nonuniform_5x3.bin", lambda: nonuniform_source(5, 3), 1, None),
    ("libvips_parity_convolution/test_gaussblur_nonuniform_5x3.bin", lambda: nonuniform_source(5, 3), None, 1e-6),
    ("gauss_blur/fractional_ramp_sigma_1_0.bin", fractional_ramp, 1, None),
    ("gauss_blur_libvips/uniform_field.bin", lambda: np.full((8, 8, 1), 37.5, np.float32), 1, None),
    ("libvips_parity_convolution/test_gaussblur_uniform_field.bin",
     lambda: np.full((8, 8, 1), 37.5, np.float32), None, 1e-6),
])
def test_gauss_f32_fixtures(rel, src, tol_ulp, tol_abs):
    # GaussBlur::new(1.0) H then V on F32 (golden/libvips/resample.rs:26-50,
    # parity/convolution.rs:48-90,193-240); fractional ramp: the libvips
    # gaussblur output of scripts/generate_golden_fixtures.sh:160.
    img = src()
    h, w, _ = img.shape
    out = O.Pipeline(w, h, 1, O.F32).gauss_blur(1.0).run(img)
    exp = fixture_f32(rel)
    assert out.size == exp.size
    if tol_ulp is not None:
        assert max_ulp(out, exp) <= tol_ulp
    else:
        assert np.max(np.abs(out.ravel() - exp)) <= tol_abs


@pytest.mark.parametrize("rel,src", [
    ("libvips_parity_convolution/test_gaussblur_u8_nonuniform_5x3.bin", lambda: grayscale_source(5, 3)),
    ("libvips_parity_convolution/test_gaussblur_u8_rgb_8x8.bin", lambda: rgb_source(8, 8)),
])
def test_gauss_u8_fixtures(rel, src):
    # parity/convolution.rs:91-190: GaussBlurH/V<U8>(1.0), held at +-1.
    img = src()
    h, w, b = img.shape
    out = O.Pipeline(w, h, b).gauss_blur(1.0).run(img)
    exp = fixture_u8(rel)
    assert out.size == exp.size
    assert np.max(np.abs(out.ravel().astype(int) - exp.astype(int))) <= 1


@pytest.mark.parametrize("rel,src,dims", [
    ("libvips_parity_resample/test_resize_half_bilinear_rgb.bin", lambda: rgb_source(16, 12), (8, 6)),
    ("libvips_parity_resample/test_resize_height_one_rounding.bin", lambda: smooth_grayscale_source(100, 1), (50, 1)),
])
def test_resize_fixtures_exact(rel, src, dims):
    # parity/resample.rs:1-90: resize(0.5, 0.5, Bilinear), held at tolerance 0.
    img = src()
    h, w, b = img.shape
    p = O.Pipeline(w, h, b).resize(0.5, 0.5, O.BILINEAR)
    assert p.output()[:2] == dims
    assert np.array_equal(p.run(img).ravel(), fixture_u8(rel))


@pytest.mark.parametrize("rel,tol", [("thumbnail_libvips/grayscale_512_to_width128.bin", 16),
                                     ("thumbnail_libvips/grayscale_512_to_width128_cache.bin", 16)])
def test_thumbnail_512_fixtures(rel, tol):
    # golden/legacy/transforms.rs:140-226 (the cached run gives the same pixels)
    out = O.Pipeline(512, 512, 1).thumbnail(O.WIDTH, 128).run(smooth_grayscale_source(512, 512))
    assert out.shape == (128, 128, 1)
    assert np.max(np.abs(out.ravel().astype(int) - fixture_u8(rel).astype(int))) <= tol


def test_thumbnail_fit_box_fixture():
    # parity/colour.rs:1-52: 96x64 FitBox(40, 24) -> 36x24, held at +-20.
    p = O.Pipeline(96, 64, 1).thumbnail(O.FITBOX, 40, 24)
    assert p.output()[:2] == (36, 24)
    out = p.run(grayscale_source(96, 64))
    exp = fixture_u8("libvips_parity_resample/test_thumbnail_fit_box_40x24.bin")
    assert np.max(np.abs(out.ravel().astype(int) - exp.astype(int))) <= 20


def test_reduce_op_lanczos3_fixture():
    # golden/legacy/resize.rs:233-265: reduce(2.0, 2.0, Lanczos3) of the 8x8
    # smooth source, held at +-4 against libvips.
    out = O.Pipeline(8, 8, 1).reduce(2.0, 2.0, O.LANCZOS3).run(smooth_grayscale_source(8, 8))
    exp = fixture_u8("reduce_libvips/smooth_grayscale_factor2_lanczos3.bin")
    assert out.size == exp.size == 16
    assert np.max(np.abs(out.ravel().astype(int) - exp.astype(int))) <= 4


def test_every_fixture_file_is_used():
    # Every committed reference fixture is read by some CPU test.
    here = pathlib.Path(__file__).resolve().parent
    text = "".join(p.read_text() for p in here.glob("test_oracle_*.py"))
    unused = [str(f.relative_to(FIX)) for f in sorted(FIX.rglob("*.bin"))
              if f.name not in text and f.stem not in text]
    assert unused == []


# ---------------------------------------------- exact-value closed forms

def _exact_f32(val, bound):
    """The f32 every FP64 evaluation within `bound` of the exact `val` rounds
    to, or None when the interval straddles an f32 rounding boundary."""
    if bound == 0:
        # an FP64 evaluation proven exact: val is a double, rounded once (RNE)
        assert Fraction(float(val)) == val
        return np.float32(float(val))
    lo = np.float32(float(val - bound))
    hi = np.float32(float(val + bound))
    return lo if lo == hi else None


def _conv_exact(rows, weights, scale=Fraction(1), offset=Fraction(0)):
    """Exact weighted sums of `rows` (list of windows of Fractions or None)
    with rigorous FP64 error bounds: n products, n additions, then the
    division by scale and the offset add (each one rounding)."""
    n = len(weights)
    out_val, out_bound = [], []
    for win in rows:
        if any(v is None for v in win):
            out_val.append(None)
            out_bound.append(None)
            continue
        terms = [w * v for w, v in zip(weights, win)]
        s = sum(terms, Fraction(0))
        mag = sum((abs(t) for t in terms), Fraction(0))
        acc_bound = (2 * n + 2) * U * mag
        val = s / scale + offset
        bound = acc_bound / abs(scale) + 3 * U * (mag / abs(scale) + abs(offset))
        out_val.append(val)
        out_bound.append(bound)
    return out_val, out_bound


def _to_f32_or_none(vals, bounds):
    return [None if v is None else _exact_f32(v, b) for v, b in zip(vals, bounds)]


def _check_f32(got, expected, what, min_determined=0.95):
    # Samples whose exact value lies within the error bound of an f32
    # rounding boundary are left open.  With masks of small-denominator
    # rational weights (taps / 44) the exact sum of f32 inputs lands ON an
    # f32 midpoint for about 1 in 2 * denominator samples; those depend on
    # the FP64 rounding of the weights and the summation order.
    got = np.asarray(got, np.float32).ravel()
    det = [i for i, e in enumerate(expected) if e is not None]
    assert len(det) >= min_determined * len(expected), (what, len(det), len(expected))
    bad = [i for i in det if got[i] != expected[i]]
    assert not bad, f"{what}: {len(bad)} of {len(det)} determined samples differ, first {bad[0]}: " \
                    f"{got[bad[0]]!r} vs {expected[bad[0]]!r}"


def _fr(x):
    return None if x is None else Fraction(float(x))


def _sep_pass_exact(img, weights, axis, scale=Fraction(1), offset=Fraction(0)):
    """One clamp-padded 1-D pass along `axis` (1 = H, 0 = V) of an odd mask,
    as a (h, w, b) array of exact f32 results or None."""
    h, w, b = img.shape
    n = len(weights)
    r = n // 2
    wts = [Fraction(float(c)) for c in weights]
    rows = []
    for y in range(h):
        for x in range(w):
            for c in range(b):
                if axis == 1:
                    win = [img[y, min(max(x + t - r, 0), w - 1), c] for t in range(n)]
                else:
                    win = [img[min(max(y + t - r, 0), h - 1), x, c] for t in range(n)]
                rows.append([_fr(v) for v in win])
    vals, bounds = _conv_exact(rows, wts, scale, offset)
    out = np.empty((h, w, b), object)
    out.ravel()[:] = _to_f32_or_none(vals, bounds)
    return out


@pytest.mark.parametrize("sigma,bands,dtype", [(1.0, 1, np.float32), (2.0, 3, np.float32),
                                                (0.8, 4, np.float32), (1.5, 2, np.uint16)])
def test_gauss_float_exact_closed_form(sigma, bands, dtype):
    # GaussBlurH -> GaussBlurV float path (gauss_blur.rs:410-464): each pass
    # sums f64(sample) * (tap / scale) over a clamp-padded window and stores
    # f32.  Masks: the integer taps of the tap rule, pinned in
    # test_gauss_masks_appendix, divided by their sum.
    rng = np.random.default_rng(int(sigma * 10) + bands)
    h, w = 13, 17
    if dtype == np.uint16:
        img = rng.integers(0, 65536, size=(h, w, bands)).astype(np.uint16)
    else:
        img = (rng.normal(0.5, 0.4, size=(h, w, bands)) * np.exp2(rng.integers(-3, 4, size=(h, w, bands)))
               ).astype(np.float32)
    taps, scale = O.gauss_int_kernel(sigma)
    weights = [float(t) / float(scale) for t in taps]
    hz = _sep_pass_exact(img.astype(np.float32) if dtype == np.float32 else img, weights, axis=1)
    vt = _sep_pass_exact(hz, weights, axis=0)
    fmt = O.U16 if dtype == np.uint16 else O.F32
    out = O.Pipeline(w, h, bands, fmt).gauss_blur(sigma).run(img)
    _check_f32(out, list(vt.ravel()), f"gauss sigma={sigma}")


@pytest.mark.parametrize("n,scale,offset,bands,dtype", [(3, 1.0, 0.0, 1, np.float32),
                                                        (5, 3.0, 0.5, 3, np.float32),
                                                        (7, -2.5, 10.0, 2, np.uint8),
                                                        (1, 1.0, -1.0, 4, np.uint16)])
def test_convsep_exact_closed_form(n, scale, offset, bands, dtype):
    # ConvSepH: fma chain over the mask, then acc / scale + offset as f32;
    # ConvSepV: the same with the V mask's (scale, 0) (convsep.rs:87-107,
    # 184-204, 218-233; common.rs:150-153).
    rng = np.random.default_rng(n * 7 + bands)
    h, w = 11, 14
    coeffs = rng.normal(0.0, 1.0, size=n)
    if dtype == np.float32:
        img = rng.normal(0.0, 3.0, size=(h, w, bands)).astype(np.float32)
    elif dtype == np.uint8:
        img = rng.integers(0, 256, size=(h, w, bands)).astype(np.uint8)
    else:
        img = rng.integers(0, 65536, size=(h, w, bands)).astype(np.uint16)
    fs, fo = Fraction(scale), Fraction(offset)
    hz = _sep_pass_exact(img, coeffs, axis=1, scale=fs, offset=fo)
    vt = _sep_pass_exact(hz, coeffs, axis=0, scale=fs, offset=Fraction(0))
    fmt = {np.float32: O.F32, np.uint8: O.U8, np.uint16: O.U16}[dtype]
    out = O.Pipeline(w, h, bands, fmt).convsep(coeffs, scale, offset).run(img)
    assert out.dtype == np.float32
    _check_f32(out, list(vt.ravel()), f"convsep n={n}")


def _reduce_windows(kernel, factor, length, accumulate=False):
    cf, ci = O.reduce_tables(kernel, factor)
    n_out = int(O.lib().orc_reduce_output_len(factor, length))
    starts, phases = O.reduce_plan(kernel, factor, length, 0.0, accumulate, n_out)
    return cf, ci, starts, phases


@pytest.mark.parametrize("kernel,factor", [(O.LANCZOS3, 2.5), (O.LANCZOS3, 1.0778947368421052),
                                           (O.BICUBIC, 1.5), (O.BILINEAR, 3.7), (O.LANCZOS2, 2.0)])
@pytest.mark.parametrize("axis,bands", [("h", 1), ("h", 3), ("v", 1), ("v", 4)])
def test_reduce_f32_exact_closed_form(kernel, factor, axis, bands):
    # ReduceH / ReduceV on f32 (reduce_simd.rs:47-144 scalar, 2621-2694 and
    # 2824-2888 AVX2): f64 weights of the (start, phase) table row times the
    # clamped samples, summed in FP64 in some order, stored as f32.  The
    # exact value decides the f32 outside a narrow band, whatever the order.
    rng = np.random.default_rng(int(factor * 100) + bands + (7 if axis == "v" else 0))
    h, w = (9, 41) if axis == "h" else (41, 9)
    img = rng.normal(0.5, 0.3, size=(h, w, bands)).astype(np.float32)
    length = w if axis == "h" else h
    cf, _, starts, phases = _reduce_windows(kernel, factor, length)
    taps = cf.shape[1]
    rows = []
    for y in range(h if axis == "h" else len(starts)):
        for x in range(len(starts) if axis == "h" else w):
            for c in range(bands):
                i = x if axis == "h" else y
                pos = [min(max(int(starts[i]) + t, 0), length - 1) for t in range(taps)]
                win = [img[y, p, c] if axis == "h" else img[p, x, c] for p in pos]
                rows.append(([Fraction(float(cf[phases[i], t])) for t in range(taps)], [_fr(v) for v in win]))
    exp = []
    for wts, win in rows:
        vals, bounds = _conv_exact([win], wts)
        exp.append(_exact_f32(vals[0], bounds[0]))
    p = O.Pipeline(w, h, bands, O.F32)
    (p.reduce_h if axis == "h" else p.reduce_v)(factor, kernel)
    out = p.run(img)
    _check_f32(out, exp, f"reduce_{axis} f32 k={kernel} f={factor}")


@pytest.mark.parametrize("kernel,factor", [(O.LANCZOS3, 2.5), (O.LANCZOS3, 3.7), (O.BICUBIC, 1.5),
                                           (O.BILINEAR, 2.0)])
@pytest.mark.parametrize("axis,bands", [("h", 1), ("h", 3), ("v", 2)])
def test_reduce_u16_exact_integer_closed_form(kernel, factor, axis, bands):
    # u16 ReduceH / ReduceV (reduce_simd.rs:47-144 and the AVX2 u16 paths;
    # sample_conv.rs:21-37): sum of i16 table weight * sample in i64, then
    # clamp((acc + 2048) >> 12, 0, 65535); positions direct per output.
    h, w = (6, 53) if axis == "h" else (53, 6)
    img = noise_u16(h, w, bands, int(factor * 10) + bands)
    length = w if axis == "h" else h
    _, ci, starts, phases = _reduce_windows(kernel, factor, length)
    taps = ci.shape[1]
    src = img.astype(np.int64)
    idx = np.clip(starts[:, None] + np.arange(taps)[None, :], 0, length - 1)  # (n_out, taps)
    wts = ci[phases].astype(np.int64)                                       # (n_out, taps)
    if axis == "h":
        acc = np.einsum("ot,yotc->yoc", wts, src[:, idx, :])
    else:
        acc = np.einsum("ot,otxc->oxc", wts, src[idx, :, :])
    exp = np.clip((acc + 2048) >> 12, 0, 65535).astype(np.uint16)
    p = O.Pipeline(w, h, bands, O.U16)
    (p.reduce_h if axis == "h" else p.reduce_v)(factor, kernel)
    assert np.array_equal(p.run(img), exp)


def _round_half_away_u16(v):
    # from_f64_clamped for u16: clamp, then f64::round (half away from zero)
    v = np.clip(v, 0.0, 65535.0)
    fl = np.floor(v)
    return np.where(v - fl >= 0.5, fl + 1, fl).astype(np.uint16)


@pytest.mark.parametrize("f", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("axis", ["h", "v"])
def test_shrink_u16_closed_form(f, axis):
    # ShrinkH / ShrinkV for non-u8 formats (shrinkh.rs:271-300,
    # shrinkv.rs:168-187): f64 sum of the box, times 1/f, from_f64_clamped.
    bands = 3
    h, w = (5, 7 * f + 3) if axis == "h" else (7 * f + 3, 5)
    img = noise_u16(h, w, bands, 40 + f)
    inv = 1.0 / f
    if axis == "h":
        ow = w // f
        s = img[:, :ow * f, :].astype(np.float64).reshape(h, ow, f, bands).sum(axis=2)
    else:
        oh = h // f
        s = img[:oh * f].astype(np.float64).reshape(oh, f, w, bands).sum(axis=1)
    exp = _round_half_away_u16(s * inv)  # integer sums are exact in f64
    p = O.Pipeline(w, h, bands, O.U16)
    (p.shrink_h if axis == "h" else p.shrink_v)(f)
    assert np.array_equal(p.run(img), exp)


@pytest.mark.parametrize("f", [2, 3, 6])
@pytest.mark.parametrize("axis", ["h", "v"])
def test_shrink_f32_exact_closed_form(f, axis):
    # As above for f32: the sum of the box in FP64 (f roundings at most), one
    # multiply by the rounded 1/f, then `as f32`.
    rng = np.random.default_rng(f + (5 if axis == "v" else 0))
    bands = 2
    h, w = (4, 5 * f) if axis == "h" else (5 * f, 4)
    img = rng.normal(0.0, 10.0, size=(h, w, bands)).astype(np.float32)
    inv = Fraction(1.0 / f)
    exp = []
    oh, ow = (h, w // f) if axis == "h" else (h // f, w)
    for y in range(oh):
        for x in range(ow):
            for c in range(bands):
                box = ([img[y, x * f + k, c] for k in range(f)] if axis == "h"
                       else [img[y * f + k, x, c] for k in range(f)])
                s, exact = Fraction(0), True
                for v in box:  # in box order, as the reference adds
                    s += Fraction(float(v))
                    exact = exact and Fraction(float(s)) == s
                mag = sum((abs(Fraction(float(v))) for v in box), Fraction(0))
                if exact and inv.denominator & (inv.denominator - 1) == 0 and inv.numerator == 1:
                    # every partial sum is a double and 1/f a power of two:
                    # the FP64 result is exact, ties included
                    exp.append(_exact_f32(s * inv, 0))
                else:
                    exp.append(_exact_f32(s * inv, (f + 2) * U * mag * abs(inv) + U * abs(s * inv)))
    p = O.Pipeline(w, h, bands, O.F32)
    (p.shrink_h if axis == "h" else p.shrink_v)(f)
    _check_f32(p.run(img), exp, f"shrink_{axis} f32 f={f}")


def test_exact_band_catches_a_dropped_tap():
    # The closed form must be able to fail: dropping one tap from the mask
    # (a plausible indexing mistake) disagrees with the oracle.
    img = fractional_ramp()
    w_ok = list(sigma1_mask())
    w_bad = [w_ok[0], w_ok[1], 0.0]
    out = O.Pipeline(8, 8, 1, O.F32).convsep(np.array(w_ok), 1.0, 0.0).run(img)
    hz = _sep_pass_exact(img, w_bad, axis=1)
    vt = _sep_pass_exact(hz, w_bad, axis=0)
    with pytest.raises(AssertionError):
        _check_f32(out, list(vt.ravel()), "dropped tap")


# --- PNG box pre-shrinks (decoder_source/shrink.rs:47-117; png/decode_full.rs:320-452)

def _box_exact(src, f, ceil):
    """Exact rational mean of every window, brute force: u8 / u16 rounded
    half up (the means are non-negative), f32 the correctly rounded double of
    the mean cast to f32 (what sum / n gives when every partial sum is a
    double).  Windows are cut at the image edge; floor mode keeps at least
    one output pixel per axis."""
    h, w, b = src.shape
    dw = -(-w // f) if ceil else max(w // f, 1)
    dh = -(-h // f) if ceil else max(h // f, 1)
    out = np.zeros((dh, dw, b), src.dtype)
    for dy in range(dh):
        for dx in range(dw):
            win = src[dy * f:min(dy * f + f, h), dx * f:min(dx * f + f, w)]
            n = win.shape[0] * win.shape[1]
            for c in range(b):
                m = sum((Fraction(float(v)) for v in win[:, :, c].ravel()), Fraction(0)) / n
                if src.dtype == np.float32:
                    out[dy, dx, c] = np.float32(float(m))
                else:
                    out[dy, dx, c] = int(m + Fraction(1, 2))  # floor(m + 1/2)
    return out


def test_box_shrink_png_tracking_decoder():
    # decoder_source/tests.rs:389-453: an 8x8 one-band image of 128 shrinks
    # to 2x2 at factor 4.
    src = np.full((8, 8, 1), 128, np.uint8)
    got = O.box_shrink(src, 4)
    assert got.shape == (2, 2, 1) and (got == 128).all()
    p = O.Pipeline(8, 8, 1).box_shrink(4)
    assert p.output()[:2] == (2, 2)


def _png_tail_pixels(w, h):
    # png/tests.rs:606-647: the partial-tail test image
    y, x = np.mgrid[0:h, 0:w].astype(np.uint32)
    return np.stack([(x * 13 + y) & 255, (y * 7 + x * 3) & 255, x ^ ((y * 11) & 255)], -1).astype(np.uint8)


def _png_ga_pixels(w, h):
    # png/tests.rs:650-680: grey + alpha pairs
    y, x = np.mgrid[0:h, 0:w].astype(np.int64)
    return np.stack([(x * 17 + y * 9) & 255, (255 - ((x * 21 + y * 5) & 255)) & 255], -1).astype(np.uint8)


def test_box_shrink_png_decode_fixtures():
    # decode_path_with_shrink_keeps_partial_tail_blocks (17x18x3, f=16: 2x2
    # output, partial last column and row) and ..._preserves_grayscale_alpha_pairs
    # (4x4x2, f=2), both in the row decoder's ceil form.
    tail = _png_tail_pixels(17, 18)
    got = O.box_shrink(tail, 16, ceil=True)
    assert got.shape == (2, 2, 3)
    assert np.array_equal(got, _box_exact(tail, 16, True))
    ga = _png_ga_pixels(4, 4)
    assert np.array_equal(O.box_shrink(ga, 2, ceil=True), _box_exact(ga, 2, True))


@pytest.mark.parametrize("f", [2, 3, 4, 8])
def test_box_shrink_u8_ramp_closed_form(f):
    # v = x + 2y + 3: a window's mean is v at its centre, x0 + (f-1)/2 and
    # y0 + (f-1)/2, so an even f lands on a half (3/2 per axis step of 1/2):
    # rounding half up gives ceil, a truncating divide would give floor.
    w, h = 10 * f + f // 2, 7 * f + 1
    y, x = np.mgrid[0:h, 0:w]
    src = (x + 2 * y + 3).astype(np.uint8)[..., None]
    got = O.box_shrink(src, f)
    dy, dx = np.mgrid[0:h // f, 0:w // f]
    centre = Fraction(f - 1, 2)
    exp = np.vectorize(lambda a, b: int(a * f + centre + 2 * (b * f + centre) + 3 + Fraction(1, 2)))(dx, dy)
    assert got.shape == (h // f, w // f, 1)
    assert np.array_equal(got[..., 0], exp)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("f,ceil", [(2, False), (3, False), (5, False), (4, True), (3, True)])
def test_box_shrink_brute_force(dtype, f, ceil):
    if ceil and dtype != np.uint8:
        pytest.skip("the row decoder's form is u8 only")
    r = np.random.default_rng(f * 7 + int(ceil))
    h, w, b = 13, 17, 3
    if dtype == np.float32:
        src = (r.integers(-512, 2048, (h, w, b)) / 1024).astype(np.float32)
    else:
        src = r.integers(0, np.iinfo(dtype).max + 1, (h, w, b), dtype=np.int64).astype(dtype)
    assert np.array_equal(O.box_shrink(src, f, ceil), _box_exact(src, f, ceil))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_box_shrink_window_cut_when_smaller_than_factor(dtype):
    # shrink.rs:58-59,70-73: a 5x3 image at f = 4 gives one pixel, the mean
    # of all 15.
    src = (np.arange(30).reshape(3, 5, 2) * 7).astype(dtype)
    got = O.box_shrink(src, 4)
    assert got.shape == (1, 1, 2)
    assert np.array_equal(got, _box_exact(src, 4, False))


def test_box_shrink_catches_a_truncating_divide():
    # A plausible mistake (sum / n without the + n/2) disagrees with the
    # closed form on the ramp.
    src = (np.arange(16, dtype=np.uint8) * 3).reshape(4, 4, 1)
    good = _box_exact(src, 2, False)
    bad = src.reshape(2, 2, 2, 2).astype(np.uint32).sum((1, 3)) // 4
    assert not np.array_equal(good[..., 0], bad)
    assert np.array_equal(O.box_shrink(src, 2), good)


@pytest.mark.parametrize("w,h,target,tw,th,crop", [
    (8192, 8192, O.WIDTH, 400, 0, False),    # hint 16: 512 x 512, then reduce 1.28
    (4000, 3000, O.FITBOX, 300, 300, False),  # hint 4
    (1200, 900, O.WIDTH, 250, 0, False),     # hint 2
    (900, 700, O.WIDTH, 300, 0, False),      # no hint: the plain thumbnail
    (3000, 2000, O.FITBOX, 200, 200, True),  # crop
])
def test_thumbnail_shrink_on_load_composes(w, h, target, tw, th, crop):
    # builder/resample.rs:878-906: a box pre-shrink by the hint, then the
    # thumbnail planned over the shrunk image.
    bands = 3
    f = O.lib().orc_thumbnail_shrink_on_load(w, h, target, tw, th, int(crop))
    p = O.Pipeline(w, h, bands).thumbnail_shrink_on_load(target, tw, th, crop=crop)
    if f:
        sw, sh = max(w // f, 1), max(h // f, 1)
    else:
        sw, sh = w, h
    q = O.Pipeline(sw, sh, bands).thumbnail(target, tw, th, crop=crop)
    assert p.output() == q.output()
    assert p.nodes() == q.nodes()
    if w * h <= 4_000_000:
        src = O.synthetic_bytes_fast(w * h * bands, w + h).reshape(h, w, bands)
        shrunk = O.box_shrink(src, f) if f else src
        assert np.array_equal(p.run(src), q.run(shrunk))
