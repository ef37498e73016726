"""The backend's C ABI (include/kazan/kazan_backend.h) on the GPU: the
counterpart of the reference's Compiler::run -> CompiledCode::get(key)
(shader-compiler-backend/src/lib.rs:191-257), with kernels keyed by name.

- A compiled pipeline exports its three kernels by key; an unknown key
  gives NULL and leaves no error behind.
- The compiled code's object bytes load again as compiled code with the same
  kernels (kazan_code_load: a code object from a pipeline cache, say).
- kazan_code_launch runs a kernel of a code object loaded from elsewhere
  with its arguments packed as one struct, and the kernel's writes land.
- The NoOptimizations mode compiles and loads too.
"""
import ctypes

import pytest

from kazan_amd.scenes import generators as g
from kazan_amd.scenes.renderer import pipeline_desc

pytestmark = pytest.mark.gpu

KERNELS = ("kz_bin_count", "kz_raster", "kz_raster_persistent")


def _lib(driver):
    L = driver.lib
    L.kazan_backend_compile.restype = ctypes.c_int
    L.kazan_backend_compile.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.kazan_code_load.restype = ctypes.c_int
    L.kazan_code_load.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p),
                                  ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.kazan_code_get.restype = ctypes.c_void_p
    L.kazan_code_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L.kazan_code_module.restype = ctypes.c_void_p
    L.kazan_code_module.argtypes = [ctypes.c_void_p]
    L.kazan_code_object.restype = ctypes.c_void_p
    L.kazan_code_object.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    L.kazan_code_launch.restype = ctypes.c_int
    L.kazan_code_launch.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_size_t]
    L.kazan_code_destroy.argtypes = [ctypes.c_void_p]
    return L


def _compile(L, scene, opt=1):
    desc, keep = pipeline_desc(scene, 4)
    code = ctypes.c_void_p()
    log = ctypes.create_string_buffer(1 << 16)
    n = ctypes.c_size_t()
    rc = L.kazan_backend_compile(ctypes.byref(desc), opt, ctypes.byref(code), log, len(log), ctypes.byref(n))
    assert rc == 0, log.value.decode(errors="replace")
    return code


def test_compiled_code_get_by_key(driver):
    L = _lib(driver)
    code = _compile(L, g.jittered_grid(4, 4, 32, 32, 16))
    try:
        assert L.kazan_code_module(code)
        fns = {k: L.kazan_code_get(code, k.encode()) for k in KERNELS}
        assert all(fns.values()), fns
        assert len(set(fns.values())) == 3
        assert L.kazan_code_get(code, b"kz_bin_count") == fns["kz_bin_count"]  # the same function again
        assert not L.kazan_code_get(code, b"no_such_kernel")
        # the object's bytes load as compiled code with the same kernels
        size = ctypes.c_size_t()
        ptr = L.kazan_code_object(code, ctypes.byref(size))
        assert ptr and size.value > 10000
        again = ctypes.c_void_p()
        n = ctypes.c_size_t()
        assert L.kazan_code_load(ptr, size.value, ctypes.byref(again), None, 0, ctypes.byref(n)) == 0
        try:
            assert all(L.kazan_code_get(again, k.encode()) for k in KERNELS)
        finally:
            L.kazan_code_destroy(again)
        # bytes that are no code object are refused
        junk = ctypes.create_string_buffer(128)
        assert L.kazan_code_load(junk, 128, ctypes.byref(again), None, 0, ctypes.byref(n)) < 0
    finally:
        L.kazan_code_destroy(code)


FILL_SRC = r"""
#include <hip/hip_runtime.h>
extern "C" __global__ void fill(unsigned *out, unsigned n, unsigned v) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v + i;
}
"""


def test_launch_through_the_abi(driver, tmp_path):
    """A code object from elsewhere (a small kernel built here with hipcc)
    loads through kazan_code_load, is found by its key and runs through
    kazan_code_launch with its arguments packed as one struct."""
    import subprocess

    import numpy as np
    src = tmp_path / "fill.hip"
    src.write_text(FILL_SRC)
    obj = tmp_path / "fill.hsaco"
    subprocess.run(["/opt/rocm/bin/hipcc", "--genco", "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", str(src), "-o", str(obj)], check=True, capture_output=True)
    blob = obj.read_bytes()
    L = _lib(driver)
    hip = ctypes.CDLL("libamdhip64.so")
    code = ctypes.c_void_p()
    n = ctypes.c_size_t()
    assert L.kazan_code_load(blob, len(blob), ctypes.byref(code), None, 0, ctypes.byref(n)) == 0
    dev = ctypes.c_void_p()
    count = 1000
    assert hip.hipMalloc(ctypes.byref(dev), ctypes.c_size_t(count * 4)) == 0
    try:
        class Args(ctypes.Structure):
            _fields_ = [("out", ctypes.c_void_p), ("n", ctypes.c_uint32), ("v", ctypes.c_uint32)]
        args = Args(dev.value, count, 7)
        grid = (ctypes.c_uint32 * 3)((count + 255) // 256, 1, 1)
        block = (ctypes.c_uint32 * 3)(256, 1, 1)
        assert L.kazan_code_launch(code, b"fill", grid, block, 0, None, ctypes.byref(args), ctypes.sizeof(args)) == 0
        assert L.kazan_code_launch(code, b"no_such_kernel", grid, block, 0, None, ctypes.byref(args),
                                   ctypes.sizeof(args)) == -2
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(count, np.uint32)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, ctypes.c_size_t(count * 4), 2) == 0
        np.testing.assert_array_equal(out, 7 + np.arange(count, dtype=np.uint32))
    finally:
        hip.hipFree(dev)
        L.kazan_code_destroy(code)


def test_no_optimizations_mode(driver):
    L = _lib(driver)
    code = _compile(L, g.random_triangles(8, 32, 32), opt=0)
    try:
        assert all(L.kazan_code_get(code, k.encode()) for k in KERNELS)
    finally:
        L.kazan_code_destroy(code)
