"""ctypes binding of libviprs_amd.so (the C ABI declared in include/viprs_amd.h).

There is no CPU fallback: if the gfx950 library is missing, importing the
binding raises, and every entry point needs a HIP device.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import re

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libviprs_amd.so"
HEADER = HERE.parent / "include" / "viprs_amd.h"


class Region(C.Structure):
    """Region (crates/viprs-core/src/image/region.rs:18-27)."""

    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32)]

    def __repr__(self) -> str:
        return f"Region({self.x}, {self.y}, {self.width}, {self.height})"

    def as_tuple(self):
        return (self.x, self.y, self.width, self.height)


class BatchStats(C.Structure):
    """vgpu_batch_stats."""

    _fields_ = [("images", C.c_uint32), ("launches", C.c_uint32), ("shrink_tasks", C.c_uint32),
                ("reduce_tasks", C.c_uint32), ("fallback_images", C.c_uint32), ("grid_cap", C.c_int32),
                ("use_graph", C.c_int32), ("reduce_first_tile", C.c_int32), ("reduce_grid_cap", C.c_int32),
                ("input_bytes", C.c_uint64),
                ("output_bytes", C.c_uint64), ("intermediate_bytes", C.c_uint64), ("reduce_chain", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    """vgpu_server_config."""

    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("streams", C.c_uint32),
                ("planner_threads", C.c_uint32)]


class Request(C.Structure):
    """vgpu_request."""

    _fields_ = [("src", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("bands", C.c_uint32),
                ("format", C.c_int), ("src_on_device", C.c_int), ("src_pinned", C.c_int),
                ("thumbnail", C.c_int), ("target", C.c_int), ("target_w", C.c_uint32),
                ("target_h", C.c_uint32), ("kernel", C.c_int), ("blur_sigma", C.c_float),
                ("arrival_ns", C.c_int64)]


class RequestResult(C.Structure):
    """vgpu_request_result."""

    _fields_ = [("status", C.c_int), ("width", C.c_uint32), ("height", C.c_uint32), ("bands", C.c_uint32),
                ("format", C.c_int), ("out_bytes", C.c_uint64), ("arena_bytes", C.c_uint64),
                ("plan_ns", C.c_int64), ("queue_ns", C.c_int64), ("run_ns", C.c_int64),
                ("latency_ns", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerCounters(C.Structure):
    """vgpu_server_counters."""

    _fields_ = [(k, C.c_uint64) for k in ("completed", "rejected", "inflight", "peak_inflight", "queued",
                                           "arena_bytes", "arena_used", "arena_peak")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EncodeOptions(C.Structure):
    """vgpu_encode_options."""

    _fields_ = [("format", C.c_int), ("quality", C.c_int), ("subsampling", C.c_int), ("compression", C.c_int)]


class ProcessResult(C.Structure):
    """vgpu_process_result."""

    _fields_ = [("input_format", C.c_char * 8), ("input_width", C.c_uint32), ("input_height", C.c_uint32),
                ("output_format", C.c_char * 8), ("bytes_written", C.c_uint64)]

    def as_dict(self):
        return {"input_format": self.input_format.decode(), "input_dimensions": (self.input_width, self.input_height),
                "output_format": self.output_format.decode(), "bytes_written": self.bytes_written}


BUILD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)

_lib = None


def declared_symbols() -> list[str]:
    """Every function name the public header declares."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(vgpu_[a-z0-9_]+)\s*\(", text)))


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m viprs_amd.build` "
            "(the GPU path has no CPU fallback)")
    L = C.CDLL(str(LIB_PATH))
    vp, u32, i32, dbl, sz, f32 = C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_size_t, C.c_float
    P = C.POINTER
    sig = {
        "vgpu_status_string": (C.c_char_p, [i32]),
        "vgpu_context_create": (i32, [i32, P(vp)]),
        "vgpu_context_destroy": (None, [vp]),
        "vgpu_context_set_stream": (i32, [vp, vp]),
        "vgpu_context_stream": (vp, [vp]),
        "vgpu_synchronize": (i32, [vp]),
        "vgpu_context_pool_bytes": (sz, [vp]),
        "vgpu_context_trim": (i32, [vp]),
        "vgpu_buffer_alloc": (i32, [vp, sz, P(vp)]),
        "vgpu_buffer_free": (i32, [vp, vp]),
        "vgpu_memcpy_h2d": (i32, [vp, vp, vp, sz]),
        "vgpu_memcpy_d2h": (i32, [vp, vp, vp, sz]),
        "vgpu_pipeline_new": (i32, [vp, u32, u32, u32, i32, P(vp)]),
        "vgpu_pipeline_destroy": (None, [vp]),
        "vgpu_pipeline_reduce_h": (i32, [vp, dbl, i32]),
        "vgpu_pipeline_reduce_v": (i32, [vp, dbl, i32]),
        "vgpu_pipeline_reduce": (i32, [vp, dbl, dbl, i32]),
        "vgpu_pipeline_shrink_h": (i32, [vp, u32, i32]),
        "vgpu_pipeline_shrink_v": (i32, [vp, u32, i32]),
        "vgpu_pipeline_shrink": (i32, [vp, u32, u32]),
        "vgpu_pipeline_gauss_blur": (i32, [vp, f32]),
        "vgpu_pipeline_convsep": (i32, [vp, P(dbl), u32, dbl, dbl]),
        "vgpu_pipeline_premultiply": (i32, [vp]),
        "vgpu_pipeline_unpremultiply": (i32, [vp]),
        "vgpu_pipeline_thumbnail": (i32, [vp, i32, u32, u32, i32]),
        "vgpu_pipeline_thumbnail_crop": (i32, [vp, i32, u32, u32, i32, i32]),
        "vgpu_pipeline_box_shrink": (i32, [vp, u32, i32]),
        "vgpu_pipeline_thumbnail_shrink_on_load": (i32, [vp, i32, u32, u32, i32, i32]),
        "vgpu_thumbnail_shrink_on_load_factor": (i32, [u32, u32, u32, i32, u32, u32, i32, P(u32)]),
        "vgpu_pipeline_extract_area": (i32, [vp, u32, u32, u32, u32]),
        "vgpu_pipeline_resize": (i32, [vp, dbl, dbl, i32]),
        "vgpu_pipeline_output_info": (i32, [vp, P(u32), P(u32), P(u32), P(i32)]),
        "vgpu_pipeline_stage_count": (i32, [vp]),
        "vgpu_pipeline_set_option": (i32, [vp, i32, i32]),
        "vgpu_batch_set_option": (i32, [vp, i32, i32]),
        "vgpu_pipeline_stage_kernel": (i32, [vp, i32, C.c_char_p, sz]),
        "vgpu_pipeline_run": (i32, [vp, vp, sz, vp, sz]),
        "vgpu_pipeline_run_host": (i32, [vp, vp, sz, vp, sz]),
        "vgpu_pipeline_run_noalloc": (i32, [vp, vp, sz, vp, sz]),
        "vgpu_pipeline_time_stage_sources": (i32, [vp, P(vp), i32, sz, vp, sz, i32, i32, P(f32)]),
        "vgpu_pipeline_benchmark_sources": (i32, [vp, P(vp), i32, sz, vp, sz, i32, i32, P(f32), P(f32)]),
        "vgpu_pipeline_run_cancellable": (i32, [vp, vp, sz, vp, sz, vp]),
        "vgpu_cancel_token_new": (i32, [vp]),
        "vgpu_cancel_token_destroy": (None, [vp]),
        "vgpu_cancel_token_cancel": (None, [vp]),
        "vgpu_cancel_token_is_cancelled": (i32, [vp]),
        "vgpu_pipeline_run_profiled": (i32, [vp, vp, sz, vp, sz, i32, P(f32), i32]),
        "vgpu_pipeline_time_stage": (i32, [vp, vp, sz, vp, sz, i32, i32, P(f32)]),
        "vgpu_pipeline_benchmark": (i32, [vp, vp, sz, vp, sz, i32, i32, P(f32), P(f32)]),
        "vgpu_op_reduce_h_new": (i32, [vp, dbl, i32, u32, u32, i32, P(vp)]),
        "vgpu_op_reduce_v_new": (i32, [vp, dbl, i32, u32, u32, i32, P(vp)]),
        "vgpu_op_reduce_new": (i32, [vp, dbl, dbl, i32, u32, i32, P(vp)]),
        "vgpu_op_shrink_h_new": (i32, [vp, u32, i32, u32, u32, i32, P(vp)]),
        "vgpu_op_shrink_v_new": (i32, [vp, u32, i32, u32, i32, P(vp)]),
        "vgpu_op_gauss_blur_h_new": (i32, [vp, f32, u32, i32, P(vp)]),
        "vgpu_op_convsep_h_new": (i32, [vp, P(dbl), u32, dbl, dbl, u32, i32, P(vp)]),
        "vgpu_op_convsep_v_new": (i32, [vp, P(dbl), u32, dbl, dbl, u32, i32, P(vp)]),
        "vgpu_op_gauss_blur_v_new": (i32, [vp, f32, u32, i32, P(vp)]),
        "vgpu_op_premultiply_new": (i32, [vp, u32, i32, P(vp)]),
        "vgpu_op_unpremultiply_new": (i32, [vp, u32, i32, P(vp)]),
        "vgpu_op_destroy": (None, [vp]),
        "vgpu_op_required_input_region": (i32, [vp, Region, P(Region)]),
        "vgpu_op_output_size": (i32, [vp, u32, u32, P(u32), P(u32)]),
        "vgpu_op_output_format": (i32, [vp, P(i32)]),
        "vgpu_op_process_region": (i32, [vp, vp, sz, vp, sz, Region, Region]),
        "vgpu_op_process_region_host": (i32, [vp, vp, sz, vp, sz, Region, Region]),
        "vgpu_pipeline_set_integer_mfma": (i32, [vp, i32]),
        "vgpu_pipeline_set_float_mode": (i32, [vp, i32]),
        "vgpu_batch_set_integer_mfma": (i32, [vp, i32]),
        "vgpu_fill_random_u8": (i32, [vp, vp, sz, C.c_uint64]),
        "vgpu_measure_stream": (i32, [vp, sz, i32, P(C.c_float), P(C.c_float)]),
        "vgpu_batch_thumbnail_new": (i32, [vp, u32, P(u32), P(u32), P(u32), i32, u32, u32, i32, P(vp)]),
        "vgpu_batch_destroy": (None, [vp]),
        "vgpu_batch_size": (i32, [vp]),
        "vgpu_batch_output_info": (i32, [vp, u32, P(u32), P(u32), P(u32)]),
        "vgpu_batch_info": (i32, [vp, P(BatchStats)]),
        "vgpu_batch_configure": (i32, [vp, i32, i32, i32, i32]),
        "vgpu_batch_run": (i32, [vp, P(vp), P(sz), P(vp), P(sz)]),
        "vgpu_batch_run_profiled": (i32, [vp, P(vp), P(sz), P(vp), P(sz), i32, P(f32), P(C.c_uint64), P(i32), i32]),
        "vgpu_batch_benchmark": (i32, [vp, P(vp), P(sz), P(vp), P(sz), i32, i32, P(f32), P(f32)]),
        "vgpu_batch_tune": (i32, [vp, P(vp), P(sz), P(vp), P(sz), i32, P(f32)]),
        "vgpu_server_create": (i32, [vp, P(ServerConfig), P(vp)]),
        "vgpu_server_destroy": (None, [vp]),
        "vgpu_server_submit": (i32, [vp, P(Request), i32, P(C.c_uint64)]),
        "vgpu_server_wait": (i32, [vp, C.c_uint64, C.c_int64, P(RequestResult)]),
        "vgpu_server_output": (i32, [vp, C.c_uint64, P(vp), P(sz)]),
        "vgpu_server_fetch": (i32, [vp, C.c_uint64, vp, sz]),
        "vgpu_server_release": (i32, [vp, C.c_uint64]),
        "vgpu_server_cancel": (i32, [vp, C.c_uint64]),
        "vgpu_server_stats": (i32, [vp, P(ServerCounters)]),
        "vgpu_server_replay": (i32, [vp, P(Request), u32, P(C.c_int64), P(C.c_int64), P(C.c_int32),
                                     P(C.c_int64), P(C.c_int64), P(C.c_int64)]),
        "vgpu_jpeg_encode": (i32, [vp, vp, u32, u32, u32, i32, i32, vp, sz, P(sz)]),
        "vgpu_png_decode": (i32, [vp, sz, P(u32), P(u32), P(u32), vp, sz]),
        "vgpu_png_encode": (i32, [vp, u32, u32, u32, i32, vp, sz, P(sz)]),
        "vgpu_process": (i32, [vp, vp, sz, BUILD_FN, vp, P(EncodeOptions), vp, vp, sz, P(sz), P(ProcessResult)]),
        "vgpu_host_alloc": (i32, [sz, P(vp)]),
        "vgpu_host_free": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
