"""The compute path's shader compile without a GPU (kazan_compile_compute_spirv,
the work vkCreateComputePipelines does; the reference's ComputePipeline::create,
vulkan-driver/src/pipeline.rs:122-154).

- Every compute shader of the scene library translates to IR whose text
  form parses back to the same text, lists its storage buffers by (set,
  binding), its workgroup variables and its barriers / atomics.
- Each compiles to a gfx950 code object; the generated kernel reads the
  built-ins from the launch (workgroup id offset by vkCmdDispatchBase's base).
- Modules a compute pipeline cannot take are refused with a message: no
  LocalSize, a workgroup over 1024 invocations, a graphics entry point.
"""
import ctypes

import numpy as np
import pytest

from kazan_amd import vk
from kazan_amd.scenes import compute_shaders as cs
from kazan_amd.scenes import spirv as sp

SHADERS = ["cs_saxpy", "cs_reduce", "cs_histogram", "cs_grid_vertices", "cs_array_length"]


def _words(code):
    return np.frombuffer(code, np.uint32).copy()


def _compile(driver, code, what):
    w = _words(code)
    return vk.tool_call(driver, "kazan_compile_compute_spirv", w.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_size_t(w.size), b"main", ctypes.c_uint32(what))


def _ir_text(driver, code):
    w = _words(code)
    return vk.tool_call(driver, "kazan_translate_to_ir_text", w.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_size_t(w.size), ctypes.c_uint32(5), b"main")


@pytest.mark.parametrize("name", SHADERS)
def test_compute_ir_text_round_trips(driver, name):
    rc, text = _ir_text(driver, getattr(cs, name)())
    assert rc == 0, text
    rc, back = vk.tool_call(driver, "kazan_ir_round_trip", text.encode(), ctypes.c_uint32(0))
    assert rc == 0, back
    assert back == text


def test_compute_ir_contents(driver):
    rc, text = _ir_text(driver, cs.cs_reduce())
    assert rc == 0, text
    assert "storage_buffer -> v: data_ptr: { set: 0, binding: 0 };" in text
    assert "storage_buffer -> sums: data_ptr: { set: 0, binding: 1 };" in text
    assert "size: 0x400 -> s: data_ptr;" in text  # shared float s[256]
    assert text.count("barrier [] -> [];") == 2
    assert "kind: local_invocation_id" in text and "kind: workgroup_id" in text
    rc, text = _ir_text(driver, cs.cs_histogram())
    assert rc == 0, text
    assert "atomic" in text and "push_constants" in text


@pytest.mark.parametrize("name", SHADERS)
def test_compute_compiles_for_gfx950(driver, name, tmp_path, monkeypatch):
    monkeypatch.setenv("KAZAN_PIPELINE_CACHE", str(tmp_path))
    code = getattr(cs, name)()
    rc, src = _compile(driver, code, 1)
    assert rc == 0, src
    assert 'extern "C" __global__' in src and "kz_compute" in src
    rc, msg = _compile(driver, code, 2)
    assert rc == 0, msg
    assert "cache_hit=0" in msg and int(msg.split("code_object_bytes=")[1].split()[0]) > 1000
    want = {"cs_reduce": "256x1x1", "cs_array_length": "1x1x1"}.get(name, "64x1x1")
    assert "local_size=" + want in msg
    rc, msg = _compile(driver, code, 2)  # the second compile is served by the cache
    assert rc == 0 and "cache_hit=1" in msg, msg


def _module(local_size, model=sp.ExecutionModelGLCompute, mode=True):
    m = sp.Module()
    m.capability(sp.CapabilityShader)
    m.set_memory_model()
    f = sp.Function(m)
    f.ret()
    m.entry_point(model, f.id, "main", [])
    if mode:
        m.execution_mode(f.id, sp.ExecutionModeLocalSize, *local_size)
    return m.to_bytes()


@pytest.mark.parametrize("code,msg", [
    (_module((64, 1, 1), mode=False), "without a LocalSize"),
    (_module((64, 32, 1)), "at most 1024"),
    (_module((64, 1, 1), model=sp.ExecutionModelVertex), "not found"),
])
def test_bad_compute_modules_are_refused(driver, code, msg):
    rc, out = _compile(driver, code, 0)
    assert rc != 0 and msg in out, out


def test_empty_kernel_compiles(driver):
    rc, out = _compile(driver, _module((8, 8, 1)), 0)
    assert rc == 0, out
