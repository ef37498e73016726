"""Error paths of the ICD, driven through its entry points as an application
would drive them (the reference returns these VkResults from the same
places; where it panics instead, Vulkan's listed code is returned).

Without a GPU: instance creation with a layer (api_impl.rs:2950), an
extension, or an API version the driver cannot serve; extension queries for
an unknown layer; an unknown entry point; and the loader manifest written
by the build (vulkan-driver/build.rs:153-174), followed the way a loader
follows it.

With a GPU (`gpu` marker): device creation with a layer or an extension
(api_impl.rs:1340-1351), shader modules whose size is no whole number of
words, pipelines whose SPIR-V is damaged or has unstructured control flow
(VK_ERROR_INVALID_SHADER_NV, a code vkCreateGraphicsPipelines may return),
memory of an unknown type or beyond any size (api_impl.rs:3460), mapping
device-local memory, and an image format the draw path has no support for.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from kazan_amd import vk

ERROR_OUT_OF_DEVICE_MEMORY = -2
ERROR_MEMORY_MAP_FAILED = -5
ERROR_LAYER_NOT_PRESENT = -6
ERROR_EXTENSION_NOT_PRESENT = -7
ERROR_INCOMPATIBLE_DRIVER = -9
ERROR_FORMAT_NOT_SUPPORTED = -11
ERROR_INVALID_SHADER_NV = -1000012000

_ref = ctypes.byref


def _instance_info(layers=0, exts=0, api=vk.VK_MAKE_VERSION(1, 1, 0), keep=None):
    app = vk.VkApplicationInfo(vk.ST_APPLICATION_INFO, None, b"errors", 1, b"kazan", 1, api)
    names = (ctypes.c_char_p * 1)(b"VK_not_a_real_name")
    keep.extend([app, names])
    return vk.VkInstanceCreateInfo(vk.ST_INSTANCE_CREATE_INFO, None, 0, ctypes.pointer(app), layers,
                                   names if layers else None, exts, names if exts else None)


@pytest.mark.parametrize("case,code", [("layer", ERROR_LAYER_NOT_PRESENT), ("extension", ERROR_EXTENSION_NOT_PRESENT),
                                       ("api_2_0", ERROR_INCOMPATIBLE_DRIVER)])
def test_create_instance_errors(driver, case, code):
    keep = []
    ci = _instance_info(layers=case == "layer", exts=case == "extension",
                        api=vk.VK_MAKE_VERSION(2, 0, 0) if case == "api_2_0" else vk.VK_MAKE_VERSION(1, 1, 0),
                        keep=keep)
    inst = vk.Handle()
    assert driver.vkCreateInstance(_ref(ci), None, _ref(inst)) == code
    assert not inst.value


def test_instance_succeeds_without_them(driver):
    keep = []
    inst = vk.Handle()
    assert driver.vkCreateInstance(_ref(_instance_info(keep=keep)), None, _ref(inst)) == 0
    driver.vkDestroyInstance(inst, None)


def _raw(driver, name, restype, *argtypes):
    addr = driver._gipa(None, name.encode())
    assert addr, name
    return ctypes.CFUNCTYPE(restype, *argtypes)(addr)


def test_extension_queries(driver):
    f = _raw(driver, "vkEnumerateInstanceExtensionProperties", ctypes.c_int32, ctypes.c_char_p,
             ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p)
    n = ctypes.c_uint32(99)
    assert f(b"VK_LAYER_not_here", _ref(n), None) == ERROR_LAYER_NOT_PRESENT
    assert f(None, _ref(n), None) == 0 and n.value == 0


def test_unknown_entry_point_resolves_to_null(driver):
    assert not driver._gipa(None, b"vkCreateFrobnicatorKHR")
    assert not driver._gipa(None, b"")


def test_loader_manifest(tmp_path):
    """The manifest build() writes, read as the Vulkan loader reads it: the
    library it names loads and negotiates the ICD interface."""
    from kazan_amd import build
    path = build.write_manifest(path=str(tmp_path / "kazan_driver.json"))
    with open(path) as f:
        m = json.load(f)
    assert m["file_format_version"] == "1.0.0"
    assert m["ICD"]["api_version"] == "1.1.0"
    lib_path = m["ICD"]["library_path"]
    assert os.path.isabs(lib_path) and os.path.samefile(lib_path, vk.DRIVER_PATH)
    lib = ctypes.CDLL(lib_path)
    neg = lib.vk_icdNegotiateLoaderICDInterfaceVersion
    neg.restype = ctypes.c_int32
    v = ctypes.c_uint32(5)
    assert neg(_ref(v)) == 0 and v.value == 5
    gipa = lib.vk_icdGetInstanceProcAddr
    gipa.restype = ctypes.c_void_p
    gipa.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    assert gipa(None, b"vkCreateInstance")


# --- with a device ---------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("case,code", [("layer", ERROR_LAYER_NOT_PRESENT), ("extension", ERROR_EXTENSION_NOT_PRESENT)])
def test_create_device_errors(ctx, case, code):
    prio = ctypes.c_float(1.0)
    qci = vk.VkDeviceQueueCreateInfo(vk.ST_DEVICE_QUEUE_CREATE_INFO, None, 0, 0, 1, ctypes.pointer(prio))
    names = (ctypes.c_char_p * 1)(b"VK_not_a_real_name")
    dci = vk.VkDeviceCreateInfo(vk.ST_DEVICE_CREATE_INFO, None, 0, 1, ctypes.pointer(qci),
                                1 if case == "layer" else 0, names if case == "layer" else None,
                                1 if case == "extension" else 0, names if case == "extension" else None, None)
    dev = vk.Handle()
    assert ctx.d.vkCreateDevice(ctx.physical, _ref(dci), None, _ref(dev)) == code
    assert not dev.value


@pytest.mark.gpu
def test_shader_module_size_not_whole_words(ctx):
    words = (ctypes.c_uint32 * 4)(0x07230203, 0x00010000, 0, 10)
    ci = vk.VkShaderModuleCreateInfo(vk.ST_SHADER_MODULE_CREATE_INFO, None, 0, 6, ctypes.addressof(words))
    m = vk.Handle()
    assert ctx.d.vkCreateShaderModule(ctx.device, _ref(ci), None, _ref(m)) == ERROR_INVALID_SHADER_NV


@pytest.mark.gpu
@pytest.mark.parametrize("damage", ["truncated", "bad_magic", "unstructured"])
def test_pipeline_from_bad_spirv(ctx, monkeypatch, damage):
    from kazan_amd.scenes import generators as g
    from kazan_amd.scenes import renderer as rd
    from kazan_amd.scenes import shaders
    good = np.frombuffer(shaders.fs_color(), np.uint32).copy()
    if damage == "truncated":
        bad = good.copy()
        starts, i = [], 5
        while i < len(bad):
            starts.append(i)
            i += int(bad[i]) >> 16
        last = starts[-1]  # the last instruction now claims words past the end
        bad[last] = (bad[last] & 0xFFFF) | (60 << 16)
    elif damage == "bad_magic":
        bad = good.copy()
        bad[0] = 0x12345678
    else:
        from test_control_flow import path_shader

        def build(b, _):
            t1, m1 = b.L(), b.L()
            b.f.selection_merge(m1)
            b.f.branch_cond(b.cond(), t1, m1)
            b.f.label(t1)
            b.f.branch(m1)
            b.f.label(m1)
            b.f.branch(t1)
        bad = np.frombuffer(path_shader(build), np.uint32).copy()
    monkeypatch.setitem(rd._FS, g.FS_COLOR, lambda: bad.tobytes())
    with pytest.raises(vk.VulkanError) as e:
        rd.SceneRenderer(ctx, g.random_triangles(8, 16, 16, seed=1))
    assert e.value.code == ERROR_INVALID_SHADER_NV


@pytest.mark.gpu
def test_memory_errors(ctx):
    d, dev = ctx.d, ctx.device
    mem = vk.Handle()
    ai = vk.VkMemoryAllocateInfo(vk.ST_MEMORY_ALLOCATE_INFO, None, 4096, 99)  # no memory type 99
    assert d.vkAllocateMemory(dev, _ref(ai), None, _ref(mem)) == ERROR_OUT_OF_DEVICE_MEMORY
    ai = vk.VkMemoryAllocateInfo(vk.ST_MEMORY_ALLOCATE_INFO, None, (1 << 63) + 4096, 0)  # beyond any size
    assert d.vkAllocateMemory(dev, _ref(ai), None, _ref(mem)) == ERROR_OUT_OF_DEVICE_MEMORY
    ai = vk.VkMemoryAllocateInfo(vk.ST_MEMORY_ALLOCATE_INFO, None, 4096, 0)  # device-local
    assert d.vkAllocateMemory(dev, _ref(ai), None, _ref(mem)) == 0
    p = ctypes.c_void_p()
    assert d.vkMapMemory(dev, mem, 0, 4096, 0, _ref(p)) == ERROR_MEMORY_MAP_FAILED
    d.vkFreeMemory(dev, mem, None)


@pytest.mark.gpu
def test_unsupported_image_format(ctx):
    with pytest.raises(vk.VulkanError) as e:
        ctx.image(16, 16, 109, 1)  # R32G32B32A32_SFLOAT: not an attachment or texture format here
    assert e.value.code == ERROR_FORMAT_NOT_SUPPORTED
