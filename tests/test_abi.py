"""CPU checks of the C ABI: librtx.so loads and exports every function include/rtx.h declares, and the Python
binding covers the same surface.  No compute calls (there is no GPU here)."""
import ctypes
import os
import re

import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "rtx.h")
LIB = os.path.join(PKG, "librtx.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void\*|void)\s+(rtx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_reference_flow():
    names = declared_functions()
    for must in ["rtx_get_blas_build_sizes", "rtx_get_tlas_build_sizes", "rtx_cmd_build_blas", "rtx_cmd_build_tlas",
                 "rtx_cmd_trace_rays", "rtx_cmd_trace_ray_buffer", "rtx_sync"]:
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="librtx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="librtx.so not built")
def test_proc_addr_resolves_every_entry_point():
    lib = ctypes.CDLL(LIB)
    get = lib.rtx_get_proc_addr
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.c_char_p]
    addr = lambda n: ctypes.cast(getattr(lib, n), ctypes.c_void_p).value
    for n in declared_functions():
        assert get(n.encode()) == addr(n), n
    # reference calls with a single librtx replacement, and calls librtx splits (unresolved under Vulkan's name)
    assert get(b"vkCmdTraceRaysKHR") == addr("rtx_cmd_trace_rays")
    assert get(b"vkCmdCopyAccelerationStructureKHR") == addr("rtx_cmd_copy_as")
    assert get(b"vkQueueWaitIdle") == addr("rtx_sync")
    for n in [b"vkCmdBuildAccelerationStructuresKHR", b"vkGetAccelerationStructureBuildSizesKHR", b"rtx_nothing", b""]:
        assert get(n) is None
    assert get(None) is None


def test_python_binding_covers_header():
    import pyrtx
    assert sorted(pyrtx.SIGNATURES) == declared_functions()


def test_struct_sizes_match_header_layouts():
    import pyrtx
    from pyrtx import scenes
    assert ctypes.sizeof(pyrtx.BlasGeometry) == 64
    assert ctypes.sizeof(pyrtx.Camera) == 16          # std140 block of main.cpp:1003-1006
    assert pyrtx.INSTANCE_DTYPE.itemsize == 64        # VkAccelerationStructureInstanceKHR
    assert scenes.HIT_RECORD_DTYPE.itemsize == 48
    assert ctypes.sizeof(pyrtx.Sbt) == 40


def test_no_gpu_means_loud_failure():
    import torch
    import pyrtx
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(pyrtx.RtxError):
        pyrtx.Context(0)
