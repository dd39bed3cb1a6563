"""ctypes binding for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

U8, U16, F32, I16, U32, I32, F64 = 0, 1, 2, 3, 4, 5, 6
NEAREST, BILINEAR, BICUBIC, LANCZOS2, LANCZOS3, CATMULLROM, LBB, VSQBS, NOHALO = range(9)
FITBOX, FORCE_EXACT, WIDTH, HEIGHT = range(4)
OK, EINVAL, EKERNEL, EFORMAT, EUNSUPPORTED = 0, -1, -2, -3, -4

DTYPES = {U8: np.uint8, U16: np.uint16, F32: np.float32, I16: np.int16, U32: np.uint32, I32: np.int32,
          F64: np.float64}
FMT_OF = {np.dtype(t): f for f, t in DTYPES.items()}

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, u32, i32, dbl = C.c_void_p, C.c_uint32, C.c_int, C.c_double
        L.orc_kernel_weight.restype = dbl
        L.orc_kernel_weight.argtypes = [i32, dbl]
        L.orc_reduce_taps.argtypes = [i32, dbl]
        L.orc_reduce_tables.argtypes = [i32, dbl, vp, vp]
        L.orc_reduce_offset.restype = dbl
        L.orc_reduce_offset.argtypes = [i32, dbl, u32]
        L.orc_reduce_output_len.restype = u32
        L.orc_reduce_output_len.argtypes = [dbl, u32]
        L.orc_reduce_plan.argtypes = [i32, dbl, u32, dbl, i32, u32, vp, vp]
        L.orc_gauss_int_kernel.argtypes = [C.c_float, vp, i32, vp]
        L.orc_pipeline_new.restype = vp
        L.orc_pipeline_new.argtypes = [u32, u32, u32, i32]
        L.orc_pipeline_free.argtypes = [vp]
        L.orc_pipeline_reduce_h.argtypes = [vp, dbl, i32]
        L.orc_pipeline_reduce_v.argtypes = [vp, dbl, i32]
        L.orc_pipeline_reduce.argtypes = [vp, dbl, dbl, i32]
        L.orc_pipeline_shrink_h.argtypes = [vp, u32, i32]
        L.orc_pipeline_shrink_v.argtypes = [vp, u32, i32]
        L.orc_pipeline_shrink.argtypes = [vp, u32, u32]
        L.orc_pipeline_gauss_blur.argtypes = [vp, C.c_float]
        L.orc_pipeline_convsep.argtypes = [vp, vp, u32, dbl, dbl]
        L.orc_pipeline_premultiply.argtypes = [vp]
        L.orc_pipeline_unpremultiply.argtypes = [vp]
        L.orc_pipeline_thumbnail.argtypes = [vp, i32, u32, u32, i32]
        L.orc_pipeline_thumbnail_crop.argtypes = [vp, i32, u32, u32, i32, i32]
        L.orc_box_shrink.argtypes = [vp, u32, u32, u32, i32, u32, i32, vp, vp, vp]
        L.orc_pipeline_box_shrink.argtypes = [vp, u32, i32]
        L.orc_pipeline_thumbnail_shrink_on_load.argtypes = [vp, i32, u32, u32, i32, i32]
        L.orc_pipeline_extract_area.argtypes = [vp, u32, u32, u32, u32]
        L.orc_pipeline_affine.argtypes = [vp, dbl, dbl, dbl, dbl, u32, u32, i32, i32]
        L.orc_shrink_on_load_factor.argtypes = [dbl]
        L.orc_shrink_on_load_factor.restype = u32
        L.orc_thumbnail_shrink_on_load.argtypes = [u32, u32, i32, u32, u32, i32]
        L.orc_thumbnail_shrink_on_load.restype = u32
        L.orc_thumbnail_plan_crop.argtypes = [u32, u32, u32, i32, u32, u32, i32, i32, vp, vp, i32, vp, vp, vp, vp]
        L.orc_pipeline_resize.argtypes = [vp, dbl, dbl, i32]
        L.orc_pipeline_output.argtypes = [vp, vp, vp, vp, vp]
        L.orc_pipeline_node_count.argtypes = [vp]
        L.orc_pipeline_node_info.argtypes = [vp, i32, vp, vp]
        L.orc_pipeline_run.argtypes = [vp, vp, vp, i32, u32]
        L.orc_thumbnail_plan.argtypes = [u32, u32, u32, i32, u32, u32, i32, vp, vp, i32, vp, vp]
        L.orc_f64_reduce_hv.argtypes = [vp, u32, u32, u32, dbl, i32, vp, vp, vp, i32]
        L.orc_f64_gauss.argtypes = [vp, u32, u32, u32, C.c_float, vp, i32]
        L.orc_synthetic_bytes.argtypes = [vp, C.c_size_t, C.c_uint64]
        L.orc_synthetic_bytes.restype = None
        L.orc_interp_2d.argtypes = [i32, vp, dbl, dbl]
        L.orc_interp_2d.restype = dbl
        L.orc_jpeg_encode.argtypes = [vp, u32, u32, u32, i32, i32, vp, C.c_size_t, C.POINTER(C.c_size_t)]
        for name, arg in (("orc_convert_fixed", C.c_int64), ("orc_convert_from_f64", dbl),
                          ("orc_convert_shrunk", dbl)):
            getattr(L, name).argtypes = [arg, i32, vp]
            getattr(L, name).restype = None
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with status {code}")
        self.code = code


def _check(rc: int, what: str) -> None:
    if rc != OK:
        raise OracleError(rc, what)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def reduce_tables(kernel: int, factor: float):
    taps = lib().orc_reduce_taps(kernel, factor)
    if taps < 0:
        raise OracleError(EINVAL, "reduce_taps")
    cf = np.zeros(65 * taps, np.float64)
    ci = np.zeros(65 * taps, np.int16)
    rc = lib().orc_reduce_tables(kernel, factor, _ptr(cf), _ptr(ci))
    if rc < 0:
        raise OracleError(rc, "reduce_tables")
    return cf.reshape(65, taps), ci.reshape(65, taps)


def reduce_plan(kernel, factor, input_len, origin, accumulate, n):
    starts = np.zeros(n, np.int64)
    phases = np.zeros(n, np.uint8)
    lib().orc_reduce_plan(kernel, factor, input_len, float(origin), int(accumulate), n,
                          _ptr(starts), _ptr(phases))
    return starts, phases


def gauss_int_kernel(sigma: float):
    c = np.zeros(512, np.int16)
    scale = np.zeros(1, np.int32)
    n = lib().orc_gauss_int_kernel(sigma, _ptr(c), 512, _ptr(scale))
    if n < 0:
        raise OracleError(n, "gauss_int_kernel")
    return c[:n].copy(), int(scale[0])


class Pipeline:
    """PipelineBuilder restatement over an in-memory source image."""

    def __init__(self, width: int, height: int, bands: int, fmt: int = U8):
        self._p = lib().orc_pipeline_new(width, height, bands, fmt)
        if not self._p:
            raise OracleError(EINVAL, "pipeline_new")
        self.src_shape = (height, width, bands)
        self.src_fmt = fmt

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_pipeline_free(self._p)
            self._p = None

    def reduce_h(self, factor, kernel=LANCZOS3):
        _check(lib().orc_pipeline_reduce_h(self._p, factor, kernel), "reduce_h")
        return self

    def reduce_v(self, factor, kernel=LANCZOS3):
        _check(lib().orc_pipeline_reduce_v(self._p, factor, kernel), "reduce_v")
        return self

    def reduce(self, hfactor, vfactor, kernel=LANCZOS3):
        """PipelineBuilder::reduce: the fused f64 ReduceOp (reduce.rs:104-166)."""
        _check(lib().orc_pipeline_reduce(self._p, hfactor, vfactor, kernel), "reduce")
        return self

    def shrink_h(self, factor, ceil=False):
        _check(lib().orc_pipeline_shrink_h(self._p, factor, int(ceil)), "shrink_h")
        return self

    def shrink_v(self, factor, ceil=False):
        _check(lib().orc_pipeline_shrink_v(self._p, factor, int(ceil)), "shrink_v")
        return self

    def shrink(self, hshrink, vshrink):
        _check(lib().orc_pipeline_shrink(self._p, hshrink, vshrink), "shrink")
        return self

    def gauss_blur(self, sigma):
        _check(lib().orc_pipeline_gauss_blur(self._p, sigma), "gauss_blur")
        return self

    def convsep(self, coeffs, scale=1.0, offset=0.0):
        c = np.ascontiguousarray(coeffs, np.float64)
        _check(lib().orc_pipeline_convsep(self._p, _ptr(c), len(c), scale, offset), "convsep")
        return self

    def premultiply(self):
        _check(lib().orc_pipeline_premultiply(self._p), "premultiply")
        return self

    def unpremultiply(self):
        _check(lib().orc_pipeline_unpremultiply(self._p), "unpremultiply")
        return self

    def thumbnail(self, target, width, height=0, kernel=LANCZOS3, crop=False):
        _check(lib().orc_pipeline_thumbnail_crop(self._p, target, width, height, kernel, int(crop)), "thumbnail")
        return self

    def box_shrink(self, factor, ceil=False):
        """The source replaced by its box-shrunk copy (the PNG pre-shrinks,
        decoder_source/shrink.rs:47-117 and png/decode_full.rs:320-452)."""
        _check(lib().orc_pipeline_box_shrink(self._p, factor, int(ceil)), "box_shrink")
        return self

    def thumbnail_shrink_on_load(self, target, width, height=0, kernel=LANCZOS3, crop=False):
        """thumbnail() over a PNG source that box-shrinks on load
        (builder/resample.rs:871-906)."""
        _check(lib().orc_pipeline_thumbnail_shrink_on_load(self._p, target, width, height, kernel, int(crop)),
               "thumbnail_shrink_on_load")
        return self

    def extract_area(self, x, y, width, height):
        """PipelineBuilder::extract_area (builder/geometry.rs:40-80)."""
        _check(lib().orc_pipeline_extract_area(self._p, x, y, width, height), "extract_area")
        return self

    def affine(self, a, d, tx, ty, width, height, kernel=BICUBIC, small_tile=True):
        """The thumbnail's Affine node (matrix [a, 0, 0, d], ExtendMode::Copy)."""
        _check(lib().orc_pipeline_affine(self._p, a, d, tx, ty, width, height, kernel, int(small_tile)), "affine")
        return self

    def resize(self, hscale, vscale, kernel=LANCZOS3):
        _check(lib().orc_pipeline_resize(self._p, hscale, vscale, kernel), "resize")
        return self

    def output(self):
        w, h, b = (C.c_uint32(), C.c_uint32(), C.c_uint32())
        f = C.c_int()
        lib().orc_pipeline_output(self._p, C.addressof(w), C.addressof(h), C.addressof(b),
                                  C.addressof(f))
        return w.value, h.value, b.value, f.value

    def nodes(self):
        out = []
        t, p = C.c_int(), C.c_double()
        for i in range(lib().orc_pipeline_node_count(self._p)):
            lib().orc_pipeline_node_info(self._p, i, C.addressof(t), C.addressof(p))
            out.append((t.value, p.value))
        return out

    def run(self, src: np.ndarray, threads: int = 0, strip_height: int = 0) -> np.ndarray:
        h, w, b = self.src_shape
        src = np.ascontiguousarray(src, DTYPES[self.src_fmt]).reshape(h, w, b)
        ow, oh, ob, of = self.output()
        dst = np.zeros((oh, ow, ob), DTYPES[of])
        if threads <= 0:
            threads = os.cpu_count() or 1
        _check(lib().orc_pipeline_run(self._p, _ptr(src), _ptr(dst), threads, strip_height), "run")
        return dst


def interp_2d(kernel: int, neighbourhood, x_phase: float, y_phase: float) -> float:
    """VSQBS or LBB over a 4 x 4 neighbourhood (rows first)."""
    n = np.ascontiguousarray(neighbourhood, np.float64).reshape(16)
    return lib().orc_interp_2d(kernel, _ptr(n), x_phase, y_phase)


def convert(kind: str, value, fmt: int):
    """One sample conversion: kind "fixed" (from_fixed_i64 of an int), "f64"
    (FromF64::from_f64) or "shrink" (ShrinkSample::from_f64_clamped)."""
    out = np.zeros(1, DTYPES[fmt])
    fn = {"fixed": lib().orc_convert_fixed, "f64": lib().orc_convert_from_f64,
          "shrink": lib().orc_convert_shrunk}[kind]
    fn(value, fmt, _ptr(out))
    return out[0].item()


def box_shrink(src: np.ndarray, factor: int, ceil: bool = False) -> np.ndarray:
    """The PNG box pre-shrink of one image (orc_box_shrink)."""
    h, w, b = src.shape
    fmt = FMT_OF[src.dtype]
    src = np.ascontiguousarray(src)
    ow, oh = C.c_uint32(), C.c_uint32()
    _check(lib().orc_box_shrink(None, w, h, b, fmt, factor, int(ceil), None, C.addressof(ow), C.addressof(oh)),
           "box_shrink")
    dst = np.zeros((oh.value, ow.value, b), src.dtype)
    _check(lib().orc_box_shrink(_ptr(src), w, h, b, fmt, factor, int(ceil), _ptr(dst), None, None), "box_shrink")
    return dst


def jpeg_encode(pixels: np.ndarray, quality: int = 75, subsampling: int = 0) -> bytes:
    """Baseline JFIF JPEG of (h, w, 1|3|4) u8 pixels (jpeg.c): subsampling 0
    auto (4:2:0 below quality 90), 1 off, 2 4:2:0."""
    a = np.ascontiguousarray(pixels, np.uint8)
    if a.ndim == 2:
        a = a[..., None]
    h, w, b = a.shape
    n = C.c_size_t()
    _check(lib().orc_jpeg_encode(_ptr(a), w, h, b, quality, subsampling, None, 0, C.byref(n)), "jpeg_encode")
    out = C.create_string_buffer(n.value)
    _check(lib().orc_jpeg_encode(_ptr(a), w, h, b, quality, subsampling, out, n.value, C.byref(n)), "jpeg_encode")
    return out.raw[:n.value]


def thumbnail_plan(in_w, in_h, bands, target, tw, th=0, kernel=LANCZOS3, crop=False):
    """Node list [(name, param)] and output size; an ExtractArea node's
    param is its (x, y, width, height), an Affine node's its
    (a, d, tx, ty, width, height, kernel) for the matrix [a, 0, 0, d]."""
    types = np.zeros(16, np.int32)
    params = np.zeros(16, np.float64)
    extract = np.zeros(4, np.uint32)
    affine = np.zeros(7, np.float64)
    ow, oh = C.c_uint32(), C.c_uint32()
    n = lib().orc_thumbnail_plan_crop(in_w, in_h, bands, target, tw, th, kernel, int(crop), _ptr(types),
                                      _ptr(params), 16, C.addressof(ow), C.addressof(oh), _ptr(extract),
                                      _ptr(affine))
    if n < 0:
        raise OracleError(n, "thumbnail_plan")
    names = {1: "Premultiply", 2: "ShrinkH", 3: "ShrinkV", 4: "ReduceH", 5: "ReduceV",
             6: "Unpremultiply", 7: "ExtractArea", 8: "Affine"}
    out = []
    for t, p in zip(types[:n], params[:n]):
        if int(t) == 7:
            out.append((names[7], tuple(int(v) for v in extract)))
        elif int(t) == 8:
            a = [float(v) for v in affine]
            out.append((names[8], (a[0], a[1], a[2], a[3], int(a[4]), int(a[5]), int(a[6]))))
        else:
            out.append((names[int(t)], float(p)))
    return out, (ow.value, oh.value)


def f64_reduce_hv(src: np.ndarray, factor: float, kernel: int = LANCZOS3, threads: int = 0):
    h, w, b = src.shape
    src = np.ascontiguousarray(src, np.float32)
    ow = int(lib().orc_reduce_output_len(factor, w))
    oh = int(lib().orc_reduce_output_len(factor, h))
    dst = np.zeros((oh, ow, b), np.float64)
    a, c = C.c_uint32(), C.c_uint32()
    _check(lib().orc_f64_reduce_hv(_ptr(src), w, h, b, factor, kernel, _ptr(dst), C.addressof(a),
                                   C.addressof(c), threads or (os.cpu_count() or 1)), "f64_reduce_hv")
    return dst


def f64_gauss(src: np.ndarray, sigma: float, threads: int = 0):
    h, w, b = src.shape
    src = np.ascontiguousarray(src, np.float32)
    dst = np.zeros((h, w, b), np.float64)
    _check(lib().orc_f64_gauss(_ptr(src), w, h, b, sigma, _ptr(dst), threads or (os.cpu_count() or 1)),
           "f64_gauss")
    return dst


def synthetic_bytes_fast(nbytes: int, seed: int) -> np.ndarray:
    """synthetic_bytes() computed by the C oracle (for multi-GB inputs)."""
    out = np.empty(nbytes, np.uint8)
    lib().orc_synthetic_bytes(_ptr(out), nbytes, seed & (2**64 - 1))
    return out


def synthetic_bytes(nbytes: int, seed: int) -> np.ndarray:
    """Seeded synthetic bytes, the host restatement of the generator the
    benchmarks fill device inputs with: 8-byte word i is
    splitmix64(seed + (i + 1) * 0x9E3779B97F4A7C15), little-endian
    (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
    OOPSLA 2014)."""
    words = (nbytes + 7) // 8
    with np.errstate(over="ignore"):
        i = np.arange(1, words + 1, dtype=np.uint64)
        z = np.uint64(seed & (2**64 - 1)) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z.astype("<u8").view(np.uint8)[:nbytes].copy()
