"""C-ABI library checks that need no GPU (CPU only)."""
import ctypes as C
import subprocess

import viprs_amd
from viprs_amd import _native


def test_library_loads_and_exports_every_declared_symbol():
    lib = _native.load()
    declared = _native.declared_symbols()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T vgpu_" in line}
    assert set(declared) == exported


def test_library_embeds_a_gfx950_code_object():
    blob = _native.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_status_strings():
    lib = _native.load()
    for s in range(0, -10, -1):
        assert lib.vgpu_status_string(s)
    assert lib.vgpu_status_string(-5) == b"buffer length does not match region"


def test_null_handles_are_rejected_without_a_device():
    lib = _native.load()
    h = C.c_void_p()
    assert lib.vgpu_pipeline_new(None, 4, 4, 1, 0, C.byref(h)) == viprs_amd.api.E_INVALID_ARGUMENT
    assert lib.vgpu_pipeline_reduce_h(None, 2.0, 4) == viprs_amd.api.E_INVALID_ARGUMENT
    assert lib.vgpu_op_required_input_region(None, _native.Region(0, 0, 1, 1), None) == viprs_amd.api.E_INVALID_ARGUMENT
    lib.vgpu_pipeline_destroy(None)
    lib.vgpu_op_destroy(None)
    lib.vgpu_context_destroy(None)
