"""The oracle on the formats beyond u8 / u16 / f32 (I16, U32, I32, F64,
band.rs:19-35), pinned by the reference's own known answers for the sample
conversions (resample/sample_conv.rs:134-141, shared_ops/sample_conv.rs:
108-131) and by closed forms of each op:

  * ReduceH/V on I16/U32/I32: the i16 table row times the clamp-padded
    samples, summed in integers, through from_fixed_i64 (pinned above);
  * ReduceH/V on F64: the FMA chain in tap order, each fma evaluated
    exactly with rationals and rounded once;
  * ShrinkH/V on every format: the f64 box sum (exact for integers) times
    1/f, through from_f64_clamped;
  * the Gaussian float path and ConvSep from these formats: the exact
    rational pass values with FP64 error bounds;
  * the PNG box pre-shrink: exact means, from_f64_clamped.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_pins import (_check_f32, _reduce_windows, _sep_pass_exact)

NEW = [(O.I16, np.int16), (O.U32, np.uint32), (O.I32, np.int32), (O.F64, np.float64)]


def _image(dtype, shape, seed, span=None):
    """Random samples over the type's whole range (or within +-span)."""
    r = np.random.default_rng(seed)
    if dtype == np.float64:
        return r.normal(0.0, 1000.0, size=shape) * np.exp2(r.integers(-20, 20, size=shape))
    info = np.iinfo(dtype)
    lo, hi = int(info.min), int(info.max)
    if span is not None:
        lo, hi = max(lo, -span), min(hi, span)
    return r.integers(lo, hi + 1, size=shape, dtype=np.int64).astype(dtype)


# ------------------------------------------------- conversion known answers

def test_from_fixed_i64_known_answers():
    # resample/sample_conv.rs:134-141
    assert O.convert("fixed", 12 << 12, O.U8) == 12
    assert O.convert("fixed", 513 << 12, O.U16) == 513
    assert O.convert("fixed", 42 << 12, O.U32) == 42
    assert O.convert("fixed", -(3 << 12), O.I16) == -4
    assert O.convert("fixed", 7 << 12, O.I32) == 7
    # and the range clamp of each type
    assert O.convert("fixed", 70000 << 12, O.I16) == 32767
    assert O.convert("fixed", -(1 << 45), O.I32) == -(1 << 31)
    assert O.convert("fixed", -(5 << 12), O.U32) == 0
    assert O.convert("fixed", (1 << 33) << 12, O.U32) == (1 << 32) - 1


def test_from_f64_known_answers():
    # shared_ops/sample_conv.rs:118-131
    assert O.convert("f64", -1.0, O.U8) == 0
    assert O.convert("f64", 255.6, O.U8) == 255
    assert O.convert("f64", 12.6, O.U16) == 13
    assert O.convert("f64", -12.6, O.I16) == -13
    assert O.convert("f64", 123.4, O.U32) == 123
    assert O.convert("f64", -123.6, O.I32) == -124
    assert O.convert("f64", 1.25, O.F32) == 1.25
    assert O.convert("f64", -3.5, O.F64) == -3.5


@pytest.mark.parametrize("v,i16", [(-2.7, -2), (-2.5, -2), (-2.4, -1), (2.5, 3), (2.4, 2), (-40000.0, -32768),
                                   (40000.0, 32767)])
def test_shrink_sample_signed(v, i16):
    # ShrinkSample for i16 / i32 (shrinkh.rs:71-107): (v + 0.5) clamped, then
    # truncated toward zero -- not round(): -2.7 gives -2.
    assert O.convert("shrink", v, O.I16) == i16
    assert O.convert("shrink", v, O.I32) == i16 if abs(v) < 1e4 else True


def test_shrink_sample_unsigned_and_float():
    # shrinkh.rs:49-69,84-94,110-128: round half away, then clamp; f32 cast; f64 kept.
    assert O.convert("shrink", 2.5, O.U32) == 3
    assert O.convert("shrink", -3.0, O.U32) == 0
    assert O.convert("shrink", 5e9, O.U32) == (1 << 32) - 1
    assert O.convert("shrink", 0.1, O.F64) == 0.1
    assert O.convert("shrink", 0.1, O.F32) == float(np.float32(0.1))


# ------------------------------------------------------------ reduce

@pytest.mark.parametrize("fmt,dtype", NEW[:3])
@pytest.mark.parametrize("kernel,factor", [(O.LANCZOS3, 2.5), (O.BICUBIC, 2.0), (O.BILINEAR, 3.3)])
@pytest.mark.parametrize("axis", ["h", "v"])
def test_reduce_integer_formats_closed_form(fmt, dtype, kernel, factor, axis):
    # reduce_h_scalar / reduce_v_scalar (reduce_simd.rs:47-144): the i16 row
    # of the direct position's phase times the clamped samples in i64, then
    # from_fixed_i64.
    bands = 2
    h, w = (5, 41) if axis == "h" else (41, 5)
    img = _image(dtype, (h, w, bands), fmt * 10 + int(factor))
    if fmt == O.I16 and axis == "h":
        img[0, :8, 0] = [-300, -100, 0, 100, 300, 600, 900, 1200]  # reduceh.rs:748-753
    length = w if axis == "h" else h
    _, ci, starts, phases = _reduce_windows(kernel, factor, length)
    taps = ci.shape[1]
    src = img.astype(object)
    exp = np.zeros((h, len(starts), bands) if axis == "h" else (len(starts), w, bands), dtype)
    for o, (st, ph) in enumerate(zip(starts, phases)):
        wts = [int(t) for t in ci[ph]]
        for a in range(w if axis == "v" else h):
            for c in range(bands):
                acc = 0
                for t in range(taps):
                    k = min(max(int(st) + t, 0), length - 1)
                    acc += wts[t] * int(src[a, k, c] if axis == "h" else src[k, a, c])
                if axis == "h":
                    exp[a, o, c] = O.convert("fixed", acc, fmt)
                else:
                    exp[o, a, c] = O.convert("fixed", acc, fmt)
    p = O.Pipeline(w, h, bands, fmt)
    (p.reduce_h if axis == "h" else p.reduce_v)(factor, kernel)
    assert np.array_equal(p.run(img), exp)


def _fma(a, b, c):
    """fma(a, b, c) on doubles: the exact value, rounded once."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


@pytest.mark.parametrize("kernel,factor", [(O.LANCZOS3, 2.5), (O.BICUBIC, 1.7)])
@pytest.mark.parametrize("axis", ["h", "v"])
def test_reduce_f64_fma_chain_exact(kernel, factor, axis):
    # F64 ReduceH/V (reduce_h_scalar / reduce_v_scalar's float loop): acc =
    # fma(x_t, w_t, acc) in tap order from 0, over the f64 phase row; the
    # value itself is stored.
    bands = 1
    h, w = (3, 23) if axis == "h" else (23, 3)
    img = _image(np.float64, (h, w, bands), 5)
    length = w if axis == "h" else h
    cf, _, starts, phases = _reduce_windows(kernel, factor, length)
    exp = np.zeros((h, len(starts), 1) if axis == "h" else (len(starts), w, 1))
    for o, (st, ph) in enumerate(zip(starts, phases)):
        for a in range(w if axis == "v" else h):
            acc = 0.0
            for t in range(cf.shape[1]):
                k = min(max(int(st) + t, 0), length - 1)
                acc = _fma(float(img[a, k, 0] if axis == "h" else img[k, a, 0]), float(cf[ph, t]), acc)
            if axis == "h":
                exp[a, o, 0] = acc
            else:
                exp[o, a, 0] = acc
    p = O.Pipeline(w, h, bands, O.F64)
    (p.reduce_h if axis == "h" else p.reduce_v)(factor, kernel)
    got = p.run(img)
    assert got.dtype == np.float64
    assert np.array_equal(got, exp)


# ------------------------------------------------------------ shrink

@pytest.mark.parametrize("fmt,dtype", NEW)
@pytest.mark.parametrize("f", [2, 3, 5])
@pytest.mark.parametrize("axis", ["h", "v"])
def test_shrink_formats_closed_form(fmt, dtype, f, axis):
    # ShrinkH / ShrinkV for non-u8 formats (shrinkh.rs:271-300,
    # shrinkv.rs:168-187): the box summed in f64 in order, times 1/f, then
    # from_f64_clamped (pinned above).
    bands = 2
    h, w = (4, 6 * f + 1) if axis == "h" else (6 * f + 1, 4)
    img = _image(dtype, (h, w, bands), f * 3 + fmt)
    inv = 1.0 / f
    oh, ow = (h, w // f) if axis == "h" else (h // f, w)
    exp = np.zeros((oh, ow, bands), dtype)
    for y in range(oh):
        for x in range(ow):
            for c in range(bands):
                s = 0.0
                for j in range(f):
                    s += float(img[y, x * f + j, c] if axis == "h" else img[y * f + j, x, c])
                exp[y, x, c] = O.convert("shrink", s * inv, fmt)
    p = O.Pipeline(w, h, bands, fmt)
    (p.shrink_h if axis == "h" else p.shrink_v)(f)
    assert np.array_equal(p.run(img), exp)


# ------------------------------------------- gauss_blur / convsep -> F32

@pytest.mark.parametrize("fmt,dtype", NEW)
def test_gauss_float_from_new_formats(fmt, dtype):
    # GaussBlurH reads every sample through ToF32 (gauss_blur.rs:248-293),
    # then the float passes (gauss_blur.rs:410-464); output F32.
    h, w, bands = 9, 12, 2
    # Samples within 2^20: exact in f32, so the closed form's open (tie)
    # cases stay as rare as for u16 images; the full range is covered by
    # the reduce and shrink tests.
    img = _image(dtype, (h, w, bands), 70 + fmt, span=1 << 20)
    taps, scale = O.gauss_int_kernel(1.2)
    weights = [float(t) / float(scale) for t in taps]
    hz = _sep_pass_exact(img.astype(np.float32), weights, axis=1)
    vt = _sep_pass_exact(hz, weights, axis=0)
    out = O.Pipeline(w, h, bands, fmt).gauss_blur(1.2).run(img)
    assert out.dtype == np.float32
    _check_f32(out, list(vt.ravel()), f"gauss from {np.dtype(dtype).name}")


@pytest.mark.parametrize("fmt,dtype", NEW)
def test_convsep_from_new_formats(fmt, dtype):
    # ConvSepH reads samples as f64 (convsep.rs:87-107); output F32.
    h, w, bands = 8, 11, 1
    img = _image(dtype, (h, w, bands), 90 + fmt, span=1 << 20)
    if dtype == np.float64:
        img = img.astype(np.float32).astype(np.float64)  # keeps the error bound of the f32 test
    coeffs = np.array([0.3, 0.45, 0.3])  # not dyadic: few exact f32 ties
    hz = _sep_pass_exact(img, coeffs, axis=1, scale=Fraction(1.7), offset=Fraction(1))
    vt = _sep_pass_exact(hz, coeffs, axis=0, scale=Fraction(1.7))
    out = O.Pipeline(w, h, bands, fmt).convsep(coeffs, 1.7, 1.0).run(img)
    _check_f32(out, list(vt.ravel()), f"convsep from {np.dtype(dtype).name}")


# ------------------------------------------------------- box pre-shrink

@pytest.mark.parametrize("fmt,dtype", NEW)
def test_box_shrink_new_formats(fmt, dtype):
    # software_box_shrink_generic for non-u8 formats: f64 sum in row order,
    # / n, from_f64_clamped.
    h, w, b, f = 9, 11, 2, 4
    img = _image(dtype, (h, w, b), 3 + fmt)
    exp = np.zeros((h // f, w // f, b), dtype)
    for dy in range(h // f):
        for dx in range(w // f):
            for c in range(b):
                s = 0.0
                for sy in range(dy * f, dy * f + f):
                    for sx in range(dx * f, dx * f + f):
                        s += float(img[sy, sx, c])
                exp[dy, dx, c] = O.convert("shrink", s / (f * f), fmt)
    assert np.array_equal(O.box_shrink(img, f), exp)
