"""Python host mirror of the reference's builder / operation API over the C ABI.

Names follow the reference: ``Pipeline`` mirrors ``PipelineBuilder``
(crates/viprs-runtime/src/pipeline/builder/resample.rs, builder/colour.rs),
``Op`` mirrors the ``DynOperation`` bridges (crates/viprs-core/src/op/
dynamic_op.rs).  Build-time validation errors surface as ``VgpuError`` with
the status code, the way the reference returns ``BuildError``.

Device buffers are ``DeviceBuffer`` objects from the context's pool (or any
object exposing ``data_ptr()`` and ``nbytes``); ``run_host`` takes and returns
numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional

import numpy as np

from . import _native
from ._native import Region

U8, U16, F32, I16, U32, I32, F64 = 0, 1, 2, 3, 4, 5, 6
NEAREST, BILINEAR, BICUBIC, LANCZOS2, LANCZOS3, CATMULLROM, LBB, VSQBS, NOHALO = range(9)
FITBOX, FORCE_EXACT, WIDTH, HEIGHT = range(4)

OK = 0
E_INVALID_ARGUMENT = -1
E_INVALID_KERNEL = -2
E_UNSUPPORTED_FORMAT = -3
E_UNSUPPORTED = -4
E_BUFFER_SIZE = -5
E_DEVICE = -6
E_NO_MEMORY = -7
E_CANCELLED = -8
E_ADMISSION = -9
E_NOT_READY = -10
E_SCRATCH_LIMIT = -11
NOALLOC_SCRATCH_BYTES = 64 << 20
OPT_SEP_ROWS_PER_TASK, OPT_SHRINK_RUN, OPT_GAUSS_ROWS, OPT_BATCH_REDUCE_CHAIN, OPT_SEP_FORM = 2, 3, 4, 5, 6
OPT_REDUCE_FORM = 7

NP_DTYPE = {U8: np.uint8, U16: np.uint16, F32: np.float32, I16: np.int16, U32: np.uint32, I32: np.int32,
            F64: np.float64}
SAMPLE_BYTES = {U8: 1, U16: 2, F32: 4}


class VgpuError(RuntimeError):
    def __init__(self, status: int, what: str):
        msg = _native.load().vgpu_status_string(status).decode()
        super().__init__(f"{what}: {msg} ({status})")
        self.status = status


def _check(status: int, what: str) -> None:
    if status != OK:
        raise VgpuError(status, what)


class Context:
    """Device executor context: HIP device, stream, buffer pool."""

    def __init__(self, device: int = 0):
        self._lib = _native.load()
        h = C.c_void_p()
        _check(self._lib.vgpu_context_create(device, C.byref(h)), "vgpu_context_create")
        self._h = h
        self.device = device
        # Pipelines, ops, batches and buffers made on this context: released
        # before the context itself, whatever order Python collects them in.
        self._children = weakref.WeakSet()

    def _adopt(self, obj) -> None:
        self._children.add(obj)

    @property
    def handle(self):
        return self._h

    def set_stream(self, hip_stream: Optional[int]) -> None:
        _check(self._lib.vgpu_context_set_stream(self._h, C.c_void_p(hip_stream or 0)), "set_stream")

    @property
    def stream(self) -> int:
        return self._lib.vgpu_context_stream(self._h) or 0

    def synchronize(self) -> None:
        _check(self._lib.vgpu_synchronize(self._h), "vgpu_synchronize")

    def measure_stream(self, nbytes: int, iterations: int = 20):
        """(read GB/s, copy GB/s) of this device: a nontemporal streaming
        read of nbytes and a device copy of nbytes (read + written bytes)."""
        r, c = C.c_float(), C.c_float()
        _check(self._lib.vgpu_measure_stream(self.handle, nbytes, iterations, C.byref(r), C.byref(c)),
               "vgpu_measure_stream")
        nb = nbytes // 16 * 16
        return nb / (r.value / 1e3) / 1e9, 2 * nb / (c.value / 1e3) / 1e9

    def pool_bytes(self) -> int:
        return int(self._lib.vgpu_context_pool_bytes(self._h))

    def trim(self) -> None:
        _check(self._lib.vgpu_context_trim(self._h), "vgpu_context_trim")

    def close(self) -> None:
        if getattr(self, "_h", None):
            for child in list(getattr(self, "_children", ())):
                child._release()
            self._lib.vgpu_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def thumbnail_shrink_on_load_factor(width: int, height: int, bands: int, target: int, target_width: int,
                                    target_height: int = 0, crop: bool = False) -> int:
    """Thumbnail::into_pipeline_nodes(...).shrink_factor (thumbnail.rs:326-330,
    514-524): the factor (2, 4 or 16; 0 for none) a decoder would shrink on
    load.  Host planning only."""
    f = C.c_uint32()
    _check(_native.load().vgpu_thumbnail_shrink_on_load_factor(width, height, bands, target, target_width,
                                                                target_height, int(bool(crop)), C.byref(f)),
           "thumbnail_shrink_on_load_factor")
    return f.value


def _ptr(t) -> int:
    return t.data_ptr()


def _nbytes(t) -> int:
    n = getattr(t, "nbytes", None)
    if isinstance(n, int):
        return n
    return t.numel() * t.element_size()


class DeviceBuffer:
    """Device memory from the context's buffer pool, shaped like an image."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(ctx._lib.vgpu_buffer_alloc(ctx.handle, self.nbytes, C.byref(p)), "vgpu_buffer_alloc")
        self._p = p
        ctx._adopt(self)

    @classmethod
    def from_numpy(cls, ctx: Context, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(ctx, arr.shape, arr.dtype)
        buf.upload(arr)
        return buf

    def data_ptr(self) -> int:
        return self._p.value or 0

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr, self.dtype)
        if arr.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(self.ctx._lib.vgpu_memcpy_h2d(self.ctx.handle, self._p, C.c_void_p(arr.ctypes.data),
                                             self.nbytes), "vgpu_memcpy_h2d")

    def fill_random(self, seed: int) -> None:
        """Seeded splitmix64 bytes, made on the device (enqueued on the context stream)."""
        _check(self.ctx._lib.vgpu_fill_random_u8(self.ctx.handle, self._p, self.nbytes,
                                                 C.c_uint64(seed & (2**64 - 1))), "vgpu_fill_random_u8")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(self.ctx._lib.vgpu_memcpy_d2h(self.ctx.handle, C.c_void_p(out.ctypes.data), self._p,
                                             self.nbytes), "vgpu_memcpy_d2h")
        return out

    def free(self) -> None:
        if getattr(self, "_p", None) and self.ctx._h:
            self.ctx._lib.vgpu_buffer_free(self.ctx.handle, self._p)
        self._p = None

    _release = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Pipeline:
    """PipelineBuilder over an in-memory source (MemorySource), run on the GPU."""

    def __init__(self, ctx: Context, width: int, height: int, bands: int, fmt: int = U8):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        _check(self._lib.vgpu_pipeline_new(ctx.handle, width, height, bands, fmt, C.byref(h)),
               "vgpu_pipeline_new")
        self._h = h
        ctx._adopt(self)
        self.src_shape = (height, width, bands)
        self.src_fmt = fmt

    @classmethod
    def _borrow(cls, ctx: Context, handle, width: int, height: int, bands: int, fmt: int = U8) -> "Pipeline":
        """A Pipeline object over a pipeline the library owns (process()'s
        build callback): its calls build nodes, it never destroys it."""
        p = cls.__new__(cls)
        p.ctx, p._lib, p._h = ctx, ctx._lib, C.c_void_p(handle)
        p.src_shape, p.src_fmt, p._borrowed = (height, width, bands), fmt, True
        return p

    def _release(self) -> None:
        if getattr(self, "_h", None) and not getattr(self, "_borrowed", False):
            self._lib.vgpu_pipeline_destroy(self._h)
        self._h = None

    def __del__(self):
        self._release()

    # -- builder methods (each returns self, like PipelineBuilder's chaining)
    def reduce_h(self, factor: float, kernel: int = LANCZOS3):
        _check(self._lib.vgpu_pipeline_reduce_h(self._h, factor, kernel), "reduce_h")
        return self

    def reduce_v(self, factor: float, kernel: int = LANCZOS3):
        _check(self._lib.vgpu_pipeline_reduce_v(self._h, factor, kernel), "reduce_v")
        return self

    def reduce(self, h_factor: float, v_factor: float, kernel: int = LANCZOS3):
        """PipelineBuilder::reduce (builder/resample.rs:349-382): the fused
        f64 ReduceOp, filters at offset 0."""
        _check(self._lib.vgpu_pipeline_reduce(self._h, h_factor, v_factor, kernel), "reduce")
        return self

    def shrink_h(self, factor: int, ceil: bool = False):
        _check(self._lib.vgpu_pipeline_shrink_h(self._h, factor, int(ceil)), "shrink_h")
        return self

    def shrink_v(self, factor: int, ceil: bool = False):
        _check(self._lib.vgpu_pipeline_shrink_v(self._h, factor, int(ceil)), "shrink_v")
        return self

    def shrink(self, hshrink: int, vshrink: int):
        _check(self._lib.vgpu_pipeline_shrink(self._h, hshrink, vshrink), "shrink")
        return self

    def gauss_blur(self, sigma: float):
        _check(self._lib.vgpu_pipeline_gauss_blur(self._h, sigma), "gauss_blur")
        return self

    def convsep(self, coeffs, scale: float = 1.0, offset: float = 0.0):
        arr = (C.c_double * len(coeffs))(*[float(c) for c in coeffs])
        _check(self._lib.vgpu_pipeline_convsep(self._h, arr, len(coeffs), scale, offset), "convsep")
        return self

    def premultiply(self):
        _check(self._lib.vgpu_pipeline_premultiply(self._h), "premultiply")
        return self

    def unpremultiply(self):
        _check(self._lib.vgpu_pipeline_unpremultiply(self._h), "unpremultiply")
        return self

    def box_shrink(self, factor: int, ceil: bool = False):
        """The PNG source's box pre-shrink as the first step: ceil=False is
        software_box_shrink_generic (decoder_source/shrink.rs:47-117), ceil=True
        the u8 row decoder's form (png/decode_full.rs:320-452)."""
        _check(self._lib.vgpu_pipeline_box_shrink(self._h, factor, int(bool(ceil))), "box_shrink")
        return self

    def thumbnail_shrink_on_load(self, target: int, width: int, height: int = 0, kernel: int = LANCZOS3,
                                 crop: bool = False):
        """thumbnail() over a PNG source that box-shrinks its decoded image by
        the plan's shrink-on-load hint (builder/resample.rs:871-906)."""
        _check(self._lib.vgpu_pipeline_thumbnail_shrink_on_load(self._h, target, width, height, kernel,
                                                                 int(bool(crop))), "thumbnail_shrink_on_load")
        return self

    def thumbnail(self, target: int, width: int, height: int = 0, kernel: int = LANCZOS3, crop: bool = False):
        """Thumbnail::new(target, kernel), with_crop() when crop is true
        (thumbnail.rs:185-190): cover the box, then a centred ExtractArea."""
        _check(
            self._lib.vgpu_pipeline_thumbnail_crop(self._h, target, width, height, kernel, int(bool(crop))),
            "thumbnail",
        )
        return self

    def extract_area(self, x: int, y: int, width: int, height: int):
        """PipelineBuilder::extract_area (builder/geometry.rs:40-80)."""
        _check(self._lib.vgpu_pipeline_extract_area(self._h, x, y, width, height), "extract_area")
        return self

    def resize(self, hscale: float, vscale: float, kernel: int = LANCZOS3):
        _check(self._lib.vgpu_pipeline_resize(self._h, hscale, vscale, kernel), "resize")
        return self

    # -- compiled pipeline
    def output_info(self):
        w, h, b, f = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_int()
        _check(self._lib.vgpu_pipeline_output_info(self._h, C.byref(w), C.byref(h), C.byref(b),
                                                   C.byref(f)), "output_info")
        return w.value, h.value, b.value, f.value

    @property
    def output_shape(self):
        w, h, b, _ = self.output_info()
        return (h, w, b)

    def set_integer_mfma(self, enable, shrink: bool = False):
        """u8 reduce / Gaussian pairs on the int8-limb MFMA kernel (default) or
        the plain integer kernels; shrink=True puts the box pre-shrink on it
        too.  All choices are bit-exact."""
        mode = (1 if enable else 0) | (2 if enable and shrink else 0)
        _check(self._lib.vgpu_pipeline_set_integer_mfma(self._h, mode), "set_integer_mfma")
        return self

    FLOAT_EXACT, FLOAT_FP32, FLOAT_FP16X3 = 0, 1, 2

    def set_float_mode(self, mode: int):
        """f32 pairs: FLOAT_EXACT (the reference's FP64 order, bit-identical;
        default), FLOAT_FP32 (plain FP32) or FLOAT_FP16X3 (FP16 hi/lo split on
        MFMA, FP32 accumulation)."""
        _check(self._lib.vgpu_pipeline_set_float_mode(self._h, int(mode)), "set_float_mode")
        return self

    def set_option(self, option: int, value: int):
        """Kernel tuning knob (OPT_*); results do not change."""
        _check(self._lib.vgpu_pipeline_set_option(self._h, option, value), "set_option")
        return self

    def stage_count(self) -> int:
        n = self._lib.vgpu_pipeline_stage_count(self._h)
        if n < 0:
            raise VgpuError(n, "stage_count")
        return n

    def stage_kernels(self) -> list:
        """Names of the kernels the fused stages launch, in order."""
        out = []
        for i in range(self.stage_count()):
            buf = C.create_string_buffer(64)
            _check(self._lib.vgpu_pipeline_stage_kernel(self._h, i, buf, 64), "stage_kernel")
            out.append(buf.value.decode())
        return out

    def run(self, src, dst) -> None:
        """Enqueue on the context stream; src / dst are device tensors."""
        _check(self._lib.vgpu_pipeline_run(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                           C.c_void_p(_ptr(dst)), _nbytes(dst)), "run")

    def run_cancellable(self, src, dst, token: "CancelToken") -> None:
        """TileScheduler::run_cancellable: raises VgpuError(E_CANCELLED) when
        the token is set before the run or between its stage launches."""
        _check(self._lib.vgpu_pipeline_run_cancellable(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                                       C.c_void_p(_ptr(dst)), _nbytes(dst), token._h),
               "run_cancellable")

    def run_noalloc(self, src, dst) -> None:
        """As run(), with the plan tables and intermediates in the library's
        fixed scratch array instead of pooled allocations
        (vgpu_pipeline_run_noalloc)."""
        _check(self._lib.vgpu_pipeline_run_noalloc(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                                   C.c_void_p(_ptr(dst)), _nbytes(dst)), "run_noalloc")

    def output_buffer(self) -> "DeviceBuffer":
        w, h, b, f = self.output_info()
        return DeviceBuffer(self.ctx, (h, w, b), NP_DTYPE[f])

    def benchmark(self, src, dst, warmup: int, iterations: int):
        """(mean ms per step over back-to-back steps, per-step ms array)."""
        total = C.c_float()
        steps = (C.c_float * iterations)()
        _check(self._lib.vgpu_pipeline_benchmark(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                                 C.c_void_p(_ptr(dst)), _nbytes(dst), warmup,
                                                 iterations, C.byref(total), steps), "benchmark")
        return total.value / iterations, np.array(steps[:], np.float64)

    def run_profiled(self, src, dst, iterations: int = 1):
        """Mean device milliseconds per fused stage (HIP events)."""
        buf = (C.c_float * 64)()
        n = self._lib.vgpu_pipeline_run_profiled(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                                 C.c_void_p(_ptr(dst)), _nbytes(dst),
                                                 iterations, buf, 64)
        if n < 0:
            raise VgpuError(n, "run_profiled")
        return [buf[i] for i in range(n)]

    def time_stage(self, src, dst, stage: int, iterations: int = 50) -> float:
        """Mean device milliseconds of fused stage `stage` over back-to-back
        launches between two HIP events."""
        ms = C.c_float()
        _check(self._lib.vgpu_pipeline_time_stage(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                                  C.c_void_p(_ptr(dst)), _nbytes(dst), stage, iterations,
                                                  C.byref(ms)), "time_stage")
        return ms.value

    def time_stage_sources(self, srcs, dst, stage: int, iterations: int = 50) -> float:
        """time_stage with launch i reading srcs[i % len(srcs)]."""
        arr = (C.c_void_p * len(srcs))(*[_ptr(x) for x in srcs])
        ms = C.c_float()
        _check(self._lib.vgpu_pipeline_time_stage_sources(self._h, arr, len(srcs), _nbytes(srcs[0]),
                                                          C.c_void_p(_ptr(dst)), _nbytes(dst), stage,
                                                          iterations, C.byref(ms)), "time_stage_sources")
        return ms.value

    def benchmark_sources(self, srcs, dst, warmup: int, iterations: int):
        """benchmark() with step i reading srcs[i % len(srcs)]."""
        arr = (C.c_void_p * len(srcs))(*[_ptr(x) for x in srcs])
        total = C.c_float()
        steps = (C.c_float * iterations)()
        _check(self._lib.vgpu_pipeline_benchmark_sources(self._h, arr, len(srcs), _nbytes(srcs[0]),
                                                         C.c_void_p(_ptr(dst)), _nbytes(dst), warmup,
                                                         iterations, C.byref(total), steps), "benchmark_sources")
        return total.value / iterations, np.array(steps[:], np.float64)

    def run_host(self, src: np.ndarray) -> np.ndarray:
        src = np.ascontiguousarray(src, NP_DTYPE[self.src_fmt])
        w, h, b, f = self.output_info()
        dst = np.empty((h, w, b), NP_DTYPE[f])
        _check(self._lib.vgpu_pipeline_run_host(self._h, C.c_void_p(src.ctypes.data), src.nbytes,
                                                C.c_void_p(dst.ctypes.data), dst.nbytes), "run_host")
        return dst


class Batch:
    """Batched thumbnails: independent u8 images (1..4 bands), one target.

    Every output equals what ``Pipeline(...).thumbnail(...)`` gives the image
    alone; the batch runs all images as a few persistent launches (one per plan
    level and kernel family) captured once into a HIP graph.
    """

    def __init__(self, ctx: Context, sizes, target: int, width: int, height: int = 0,
                 kernel: int = LANCZOS3):
        self._lib = _native.load()
        self.ctx = ctx
        self._h = None
        n = len(sizes)
        ws = (C.c_uint32 * max(n, 1))(*[int(s[0]) for s in sizes])
        hs = (C.c_uint32 * max(n, 1))(*[int(s[1]) for s in sizes])
        bs = (C.c_uint32 * max(n, 1))(*[int(s[2]) for s in sizes])
        h = C.c_void_p()
        _check(self._lib.vgpu_batch_thumbnail_new(ctx.handle, n, ws, hs, bs, target, width, height,
                                                  kernel, C.byref(h)), "vgpu_batch_thumbnail_new")
        self._h = h
        ctx._adopt(self)
        self.sizes = [tuple(int(v) for v in s) for s in sizes]

    def _release(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()

    def __len__(self) -> int:
        return len(self.sizes)

    def output_shape(self, i: int):
        w, h, b = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(self._lib.vgpu_batch_output_info(self._h, i, C.byref(w), C.byref(h), C.byref(b)),
               "vgpu_batch_output_info")
        return (h.value, w.value, b.value)

    def output_buffers(self):
        return [DeviceBuffer(self.ctx, self.output_shape(i), np.uint8) for i in range(len(self))]

    def info(self) -> dict:
        st = _native.BatchStats()
        _check(self._lib.vgpu_batch_info(self._h, C.byref(st)), "vgpu_batch_info")
        return st.as_dict()

    def configure(self, grid_cap: int = 0, use_graph: bool = True, reduce_first_tile: int = 0,
                  reduce_grid_cap: Optional[int] = None) -> None:
        """Persistent grid sizes (shrink launches; reduce / MFMA separable
        launches, by default the same), HIP graph on or off, reduce tile."""
        rg = grid_cap if reduce_grid_cap is None else reduce_grid_cap
        _check(self._lib.vgpu_batch_configure(self._h, grid_cap, rg, int(use_graph), reduce_first_tile),
               "vgpu_batch_configure")

    def set_integer_mfma(self, enable, shrink: bool = False) -> None:
        mode = (1 if enable else 0) | (2 if enable and shrink else 0)
        _check(self._lib.vgpu_batch_set_integer_mfma(self._h, mode), "vgpu_batch_set_integer_mfma")

    def set_option(self, option: int, value: int) -> None:
        """Kernel tuning knob (OPT_*) for every image; results do not change."""
        _check(self._lib.vgpu_batch_set_option(self._h, option, value), "vgpu_batch_set_option")

    def _tables(self, srcs, dsts):
        n = len(self)
        if len(srcs) != n or len(dsts) != n:
            raise ValueError(f"batch of {n} images needs {n} sources and outputs")
        sp = (C.c_void_p * max(n, 1))(*[_ptr(s) for s in srcs])
        sb = (C.c_size_t * max(n, 1))(*[_nbytes(s) for s in srcs])
        dp = (C.c_void_p * max(n, 1))(*[_ptr(d) for d in dsts])
        db = (C.c_size_t * max(n, 1))(*[_nbytes(d) for d in dsts])
        return sp, sb, dp, db

    def run(self, srcs, dsts) -> None:
        """Enqueue the batch on the context stream (device buffers)."""
        _check(self._lib.vgpu_batch_run(self._h, *self._tables(srcs, dsts)), "vgpu_batch_run")

    def benchmark(self, srcs, dsts, warmup: int, iterations: int):
        total = C.c_float()
        steps = (C.c_float * iterations)()
        _check(self._lib.vgpu_batch_benchmark(self._h, *self._tables(srcs, dsts), warmup, iterations,
                                              C.byref(total), steps), "vgpu_batch_benchmark")
        return total.value / iterations, np.array(steps[:], np.float64)

    def run_profiled(self, srcs, dsts, iterations: int = 1):
        """[(mean ms, algorithmic bytes, family)] per batched launch (HIP
        events); family "shrink" or "reduce"."""
        ms = (C.c_float * 64)()
        nb = (C.c_uint64 * 64)()
        fam = (C.c_int * 64)()
        n = self._lib.vgpu_batch_run_profiled(self._h, *self._tables(srcs, dsts), iterations, ms, nb, fam, 64)
        if n < 0:
            raise VgpuError(n, "vgpu_batch_run_profiled")
        names = {0: "shrink", 1: "reduce", 2: "sep_mfma", 3: "sep_tile"}
        return [(ms[i], int(nb[i]), names[fam[i]]) for i in range(n)]

    def tune(self, srcs, dsts, iterations: int = 3) -> float:
        best = C.c_float()
        _check(self._lib.vgpu_batch_tune(self._h, *self._tables(srcs, dsts), iterations, C.byref(best)),
               "vgpu_batch_tune")
        return best.value


class Op:
    """One DynOperation bridge executed on the GPU (dyn_process_region)."""

    def __init__(self, ctx: Context, handle: C.c_void_p, bands: int, in_fmt: int):
        self.ctx = ctx
        self._lib = ctx._lib
        self._h = handle
        ctx._adopt(self)
        self.bands = bands
        self.in_fmt = in_fmt
        f = C.c_int()
        _check(self._lib.vgpu_op_output_format(self._h, C.byref(f)), "output_format")
        self.out_fmt = f.value

    def _release(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgpu_op_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()

    @classmethod
    def _make(cls, ctx, fn, args, bands, fmt, what):
        h = C.c_void_p()
        _check(getattr(ctx._lib, fn)(ctx.handle, *args, C.byref(h)), what)
        return cls(ctx, h, bands, fmt)

    @classmethod
    def reduce_h(cls, ctx, factor, kernel, bands, input_w, fmt=U8):
        return cls._make(ctx, "vgpu_op_reduce_h_new", (factor, kernel, bands, input_w, fmt), bands, fmt,
                         "ReduceHBridge::new")

    @classmethod
    def reduce_v(cls, ctx, factor, kernel, bands, input_h, fmt=U8):
        return cls._make(ctx, "vgpu_op_reduce_v_new", (factor, kernel, bands, input_h, fmt), bands, fmt,
                         "ReduceVBridge::new")

    @classmethod
    def reduce(cls, ctx, h_factor, v_factor, kernel, bands, fmt=U8):
        return cls._make(ctx, "vgpu_op_reduce_new", (h_factor, v_factor, kernel, bands, fmt), bands, fmt,
                         "ReduceBridge::new")

    @classmethod
    def shrink_h(cls, ctx, factor, bands, source_width=0, fmt=U8, ceil=False):
        return cls._make(ctx, "vgpu_op_shrink_h_new", (factor, int(ceil), bands, source_width, fmt), bands,
                         fmt, "ShrinkHBridge::new")

    @classmethod
    def shrink_v(cls, ctx, factor, bands, fmt=U8, ceil=False):
        return cls._make(ctx, "vgpu_op_shrink_v_new", (factor, int(ceil), bands, fmt), bands, fmt,
                         "ShrinkVBridge::new")

    @classmethod
    def gauss_blur_h(cls, ctx, sigma, bands, fmt=U8):
        return cls._make(ctx, "vgpu_op_gauss_blur_h_new", (sigma, bands, fmt), bands, fmt, "GaussBlurH::new")

    @classmethod
    def gauss_blur_v(cls, ctx, sigma, bands, fmt=U8):
        return cls._make(ctx, "vgpu_op_gauss_blur_v_new", (sigma, bands, fmt), bands, fmt, "GaussBlurV::new")

    @classmethod
    def convsep_h(cls, ctx, coeffs, scale, offset, bands, fmt=F32):
        arr = (C.c_double * len(coeffs))(*[float(c) for c in coeffs])
        return cls._make(ctx, "vgpu_op_convsep_h_new", (arr, len(coeffs), scale, offset, bands, fmt), bands, fmt,
                         "ConvSepH::with_mask")

    @classmethod
    def convsep_v(cls, ctx, coeffs, scale, offset, bands, fmt=F32):
        arr = (C.c_double * len(coeffs))(*[float(c) for c in coeffs])
        return cls._make(ctx, "vgpu_op_convsep_v_new", (arr, len(coeffs), scale, offset, bands, fmt), bands, fmt,
                         "ConvSepV::with_mask")

    @classmethod
    def premultiply(cls, ctx, bands, fmt=U8):
        return cls._make(ctx, "vgpu_op_premultiply_new", (bands, fmt), bands, fmt, "Premultiply::new")

    @classmethod
    def unpremultiply(cls, ctx, bands, fmt=U8):
        return cls._make(ctx, "vgpu_op_unpremultiply_new", (bands, fmt), bands, fmt, "Unpremultiply::new")

    def required_input_region(self, output: Region) -> Region:
        r = Region()
        _check(self._lib.vgpu_op_required_input_region(self._h, output, C.byref(r)),
               "required_input_region")
        return r

    def output_size(self, in_w: int, in_h: int):
        w, h = C.c_uint32(), C.c_uint32()
        _check(self._lib.vgpu_op_output_size(self._h, in_w, in_h, C.byref(w), C.byref(h)), "output_size")
        return w.value, h.value

    def process_region(self, src, dst, input_region: Region, output_region: Region) -> None:
        _check(self._lib.vgpu_op_process_region(self._h, C.c_void_p(_ptr(src)), _nbytes(src),
                                                C.c_void_p(_ptr(dst)), _nbytes(dst),
                                                input_region, output_region), "process_region")

    def process_region_host(self, src: np.ndarray, dst: np.ndarray, input_region: Region,
                            output_region: Region) -> None:
        """dyn_process_region on host tiles (vgpu_op_process_region_host):
        synchronous, `dst` is written when it returns."""
        assert src.flags.c_contiguous and dst.flags.c_contiguous
        _check(self._lib.vgpu_op_process_region_host(self._h, C.c_void_p(src.ctypes.data), src.nbytes,
                                                     C.c_void_p(dst.ctypes.data), dst.nbytes,
                                                     input_region, output_region), "process_region_host")


class CancelToken:
    """CancellationToken (vgpu_cancel_token)."""

    def __init__(self):
        self._lib = _native.load()
        h = C.c_void_p()
        _check(self._lib.vgpu_cancel_token_new(C.byref(h)), "vgpu_cancel_token_new")
        self._h = h

    def cancel(self) -> None:
        self._lib.vgpu_cancel_token_cancel(self._h)

    def is_cancelled(self) -> bool:
        return bool(self._lib.vgpu_cancel_token_is_cancelled(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.vgpu_cancel_token_destroy(self._h)
            self._h = None


class Server:
    """Thumbnail request loop: requests are admitted by device-memory footprint
    into an arena of ``hbm_budget`` bytes and ``max_inflight`` slots as soon as
    both free up, and run next to each other on ``streams`` HIP streams."""

    def __init__(self, ctx: Context, hbm_budget: int, max_inflight: int = 64, streams: int = 4,
                 planner_threads: int = 4):
        self._lib = _native.load()
        self.ctx = ctx
        self._h = None
        cfg = _native.ServerConfig(int(hbm_budget), int(max_inflight), int(streams), int(planner_threads))
        h = C.c_void_p()
        _check(self._lib.vgpu_server_create(ctx.handle, C.byref(cfg), C.byref(h)), "vgpu_server_create")
        self._h = h
        self._keep = {}  # request id -> source kept alive until the request finishes
        ctx._adopt(self)

    @staticmethod
    def request(src, width: int, height: int, bands: int, fmt: int = U8, *, thumbnail: bool = True,
                target: int = WIDTH, target_w: int = 0, target_h: int = 0, kernel: int = LANCZOS3,
                blur_sigma: float = 0.0, on_device: Optional[bool] = None, pinned: bool = False,
                arrival_ns: int = 0) -> "_native.Request":
        """A vgpu_request.  `src` is a DeviceBuffer, a numpy array (pageable
        host memory, copied at submit) or an integer address with `on_device`
        / `pinned` saying what it is."""
        if isinstance(src, DeviceBuffer):
            addr, dev = src.data_ptr(), 1
        elif isinstance(src, np.ndarray):
            addr, dev = src.ctypes.data, 0
        else:
            addr, dev = int(src), int(bool(on_device))
        return _native.Request(addr, width, height, bands, fmt, dev, int(pinned), int(thumbnail), target,
                               target_w, target_h, kernel, float(blur_sigma), int(arrival_ns))

    def submit(self, req: "_native.Request", keep=None, auto_release: bool = False) -> int:
        rid = C.c_uint64()
        _check(self._lib.vgpu_server_submit(self._h, C.byref(req), int(auto_release), C.byref(rid)),
               "vgpu_server_submit")
        if keep is not None:
            self._keep[rid.value] = keep
        return rid.value

    def wait(self, rid: int, timeout_us: int = -1) -> dict:
        res = _native.RequestResult()
        _check(self._lib.vgpu_server_wait(self._h, rid, timeout_us, C.byref(res)), "vgpu_server_wait")
        self._keep.pop(rid, None)
        return res.as_dict()

    def fetch(self, rid: int) -> np.ndarray:
        r = self.wait(rid)
        _check(r["status"], "request")
        out = np.empty((r["height"], r["width"], r["bands"]), NP_DTYPE[r["format"]])
        _check(self._lib.vgpu_server_fetch(self._h, rid, C.c_void_p(out.ctypes.data), out.nbytes),
               "vgpu_server_fetch")
        return out

    def release(self, rid: int) -> None:
        _check(self._lib.vgpu_server_release(self._h, rid), "vgpu_server_release")

    def cancel(self, rid: int) -> None:
        """The request finishes with E_CANCELLED (vgpu_server_cancel)."""
        _check(self._lib.vgpu_server_cancel(self._h, rid), "vgpu_server_cancel")

    def stats(self) -> dict:
        st = _native.ServerCounters()
        _check(self._lib.vgpu_server_stats(self._h, C.byref(st)), "vgpu_server_stats")
        return st.as_dict()

    def replay(self, reqs, arrival_offsets_ns, breakdown: bool = False):
        """Open-loop trace: request i submitted at start + arrival_offsets_ns[i];
        returns (latency_ns, status, wall_ns) and, with breakdown, also the
        arrival -> plan built and arrival -> launch times (ns)."""
        n = len(reqs)
        arr = (_native.Request * max(n, 1))(*reqs)
        offs = (C.c_int64 * max(n, 1))(*[int(x) for x in arrival_offsets_ns])
        lat = (C.c_int64 * max(n, 1))()
        sts = (C.c_int32 * max(n, 1))()
        pln = (C.c_int64 * max(n, 1))()
        que = (C.c_int64 * max(n, 1))()
        wall = C.c_int64()
        _check(self._lib.vgpu_server_replay(self._h, arr, n, offs, lat, sts, C.byref(wall), pln, que),
               "vgpu_server_replay")
        out = (np.array(lat[:n], np.int64), np.array(sts[:n], np.int32), int(wall.value))
        if breakdown:
            out += (np.array(pln[:n], np.int64), np.array(que[:n], np.int64))
        return out

    def _release(self) -> None:
        if getattr(self, "_h", None):
            self._lib.vgpu_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def close(self) -> None:
        self._release()

    def __del__(self):
        self._release()


# ------------------------------------------------------------ encode / process

JPEG_SUBSAMPLE_AUTO, JPEG_SUBSAMPLE_OFF, JPEG_SUBSAMPLE_420 = 0, 1, 2
ENCODE_JPEG, ENCODE_PNG = 0, 1


def _bytes_call(fn, *args):
    """Two-pass call of an entry point that writes a byte stream: size, then data."""
    n = C.c_size_t()
    rc = fn(*args, None, 0, C.byref(n))
    if rc not in (OK, E_BUFFER_SIZE):
        _check(rc, fn.__name__)
    buf = C.create_string_buffer(max(n.value, 1))
    _check(fn(*args, buf, n.value, C.byref(n)), fn.__name__)
    return buf.raw[:n.value]


def jpeg_encode(ctx: Context, pixels: "DeviceBuffer", quality: int = 75,
                subsampling: int = JPEG_SUBSAMPLE_AUTO) -> bytes:
    """Baseline JPEG of a u8 (h, w, 1|3) device image (vgpu_jpeg_encode)."""
    h, w, b = pixels.shape
    return _bytes_call(ctx._lib.vgpu_jpeg_encode, ctx.handle, C.c_void_p(pixels.data_ptr()), w, h, b, quality,
                       subsampling)


def png_decode(data: bytes) -> np.ndarray:
    """PNG bytes -> (h, w, bands) uint8 (vgpu_png_decode)."""
    L = _native.load()
    w, h, b = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(L.vgpu_png_decode(data, len(data), C.byref(w), C.byref(h), C.byref(b), None, 0), "vgpu_png_decode")
    out = np.empty((h.value, w.value, b.value), np.uint8)
    _check(L.vgpu_png_decode(data, len(data), C.byref(w), C.byref(h), C.byref(b), out.ctypes.data, out.nbytes),
           "vgpu_png_decode")
    return out


def png_encode(pixels: np.ndarray, compression: int = 6) -> bytes:
    """(h, w, bands) uint8 -> PNG bytes (vgpu_png_encode)."""
    a = np.ascontiguousarray(pixels, np.uint8)
    if a.ndim == 2:
        a = a[..., None]
    h, w, b = a.shape
    return _bytes_call(_native.load().vgpu_png_encode, a.ctypes.data, w, h, b, compression)


def process(ctx: Context, data: bytes, build=None, fmt: int = ENCODE_JPEG, quality: int = 75,
            subsampling: int = JPEG_SUBSAMPLE_AUTO, compression: int = 6, token: "CancelToken" = None):
    """process() (process.rs:147-260): decode `data`, let `build(pipeline)`
    add the operations, run them on the device, encode.  Returns (bytes,
    result dict)."""
    L = ctx._lib
    errors = []

    def cb(handle, _user):
        try:
            w, h, b = C.c_uint32(), C.c_uint32(), C.c_uint32()
            f = C.c_int()
            L.vgpu_pipeline_output_info(C.c_void_p(handle), C.byref(w), C.byref(h), C.byref(b), C.byref(f))
            if build is not None:
                build(Pipeline._borrow(ctx, handle, w.value, h.value, b.value, f.value))
            return OK
        except VgpuError as e:
            return e.status
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)
            return E_INVALID_ARGUMENT

    fn = _native.BUILD_FN(cb)
    enc = _native.EncodeOptions(fmt, quality, subsampling, compression)
    res = _native.ProcessResult()
    tok = token._h if token is not None else None
    n = C.c_size_t()
    rc = L.vgpu_process(ctx.handle, data, len(data), fn, None, C.byref(enc), tok, None, 0, C.byref(n), C.byref(res))
    if errors:
        raise errors[0]
    if rc not in (OK, E_BUFFER_SIZE):
        _check(rc, "vgpu_process")
    buf = C.create_string_buffer(max(n.value, 1))
    rc = L.vgpu_process(ctx.handle, data, len(data), fn, None, C.byref(enc), tok, buf, n.value, C.byref(n),
                        C.byref(res))
    if errors:
        raise errors[0]
    _check(rc, "vgpu_process")
    return buf.raw[:n.value], res.as_dict()
