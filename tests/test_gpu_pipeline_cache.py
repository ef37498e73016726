"""Pipeline caches (vkCreatePipelineCache / vkGetPipelineCacheData /
vkMergePipelineCaches; stubs in the reference, vulkan-driver/src/api_impl.rs:3979):
the data of a cache holds the code objects of the pipelines created with it,
behind Vulkan's 32-byte header (vendor, device, pipelineCacheUUID), and a
cache created from that data in another cache directory puts the same code
objects there, so that pipeline creation finds them instead of compiling."""
import ctypes
import os
import struct

import numpy as np
import pytest

from kazan_amd import vk
from kazan_amd.scenes import generators as g

pytestmark = pytest.mark.gpu


def _cache(ctx, data=None):
    h = vk.Handle()
    buf = None if data is None else ctypes.create_string_buffer(data, len(data))
    ci = vk.VkPipelineCacheCreateInfo(vk.ST_PIPELINE_CACHE_CREATE_INFO, None, 0, 0 if data is None else len(data),
                                      None if buf is None else ctypes.addressof(buf))
    ctx.d.check("vkCreatePipelineCache", ctx.device, ctypes.byref(ci), None, ctypes.byref(h))
    return h


def _data(ctx, cache):
    n = ctypes.c_size_t(0)
    ctx.d.check("vkGetPipelineCacheData", ctx.device, cache, ctypes.byref(n), None)
    buf = ctypes.create_string_buffer(n.value)
    ctx.d.check("vkGetPipelineCacheData", ctx.device, cache, ctypes.byref(n), buf)
    return buf.raw[:n.value]


def _entries(data):
    (count,) = struct.unpack_from("<I", data, 32)
    at, out = 36, {}
    for _ in range(count):
        name = data[at:at + 32].rstrip(b"\0").decode()
        (n,) = struct.unpack_from("<Q", data, at + 32)
        out[name] = data[at + 40:at + 40 + n]
        at += 40 + n
    assert at == len(data)
    return out


def test_cache_data_round_trip(driver, monkeypatch, oracle, tmp_path):
    from kazan_amd.scenes.renderer import Context, SceneRenderer
    first, second = tmp_path / "first", tmp_path / "second"
    monkeypatch.setenv("KAZAN_PIPELINE_CACHE", str(first))
    monkeypatch.setenv("KAZAN_AUTOTUNE", "0")
    ctx = Context(driver)
    cache = _cache(ctx)
    ctx.pipeline_cache = cache
    s = g.random_triangles(200, 64, 48, seed=21)
    r = SceneRenderer(ctx, s)
    r.render(1)
    data = _data(ctx, cache)
    hsize, version, vendor, device = struct.unpack_from("<IIII", data, 0)
    props = ctx.props
    assert (hsize, version, vendor, device) == (32, 1, props.vendorID, props.deviceID)
    assert data[16:32] == bytes(props.pipelineCacheUUID)
    entries = _entries(data)
    assert entries, "the pipeline's code objects"
    for name, code in entries.items():
        assert (first / name).read_bytes() == code
    # a short buffer: the header and the entries that fit, VK_INCOMPLETE
    n = ctypes.c_size_t(36)
    buf = ctypes.create_string_buffer(36)
    assert ctx.d.vkGetPipelineCacheData(ctx.device, cache, ctypes.byref(n), buf) == vk.VK_INCOMPLETE
    assert n.value == 36 and struct.unpack_from("<I", buf.raw, 32)[0] == 0
    # another cache directory: creating a cache from the data fills it
    monkeypatch.setenv("KAZAN_PIPELINE_CACHE", str(second))
    loaded = _cache(ctx, data)
    for name, code in entries.items():
        assert (second / name).read_bytes() == code
    merged = _cache(ctx)
    ctx.d.check("vkMergePipelineCaches", ctx.device, merged, 1, ctypes.byref(loaded))
    assert _entries(_data(ctx, merged)) == entries
    # pixels from the cached code objects
    ctx.pipeline_cache = None
    r2 = SceneRenderer(ctx, s)
    r2.render(1)
    c, d, _ = r2.read_back()
    oc, od, _, _ = oracle.render(s)
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    # data of another device or format is ignored
    bad = bytearray(data)
    bad[16] ^= 0xFF
    third = tmp_path / "third"
    monkeypatch.setenv("KAZAN_PIPELINE_CACHE", str(third))
    _cache(ctx, bytes(bad))
    assert not third.exists() or not os.listdir(third)
    for h in (cache, loaded, merged):
        ctx.d.vkDestroyPipelineCache(ctx.device, h, None)
