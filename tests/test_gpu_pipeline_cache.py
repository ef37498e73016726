"""Pipeline caches (vkCreatePipelineCache / vkGetPipelineCacheData /
vkMergePipelineCaches; stubs in the reference, vulkan-driver/src/api_impl.rs:3979):
the data of a cache holds the code objects of the pipelines created with it,
behind Vulkan's 32-byte header (vendor, device, pipelineCacheUUID), each
entry with its size and a checksum. A cache created from that data holds
the same code objects, and a pipeline created with it loads them instead of
compiling; nothing is written to the shared on-disk cache, and a damaged
entry is dropped."""
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
        (n, _sum) = struct.unpack_from("<QQ", data, at + 32)
        out[name] = data[at + 48:at + 48 + n]
        at += 48 + n
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
    # another (empty) cache directory: a cache created from the data holds
    # the code objects and writes nothing to the directory
    monkeypatch.setenv("KAZAN_PIPELINE_CACHE", str(second))
    loaded = _cache(ctx, data)
    assert _entries(_data(ctx, loaded)) == entries
    assert not second.exists() or not os.listdir(second)
    merged = _cache(ctx)
    ctx.d.check("vkMergePipelineCaches", ctx.device, merged, 1, ctypes.byref(loaded))
    assert _entries(_data(ctx, merged)) == entries
    # pixels from the cached code objects: nothing is compiled (a compile
    # would write its code object into the empty directory)
    ctx.pipeline_cache = loaded
    r2 = SceneRenderer(ctx, s)
    r2.render(1)
    c, d, _ = r2.read_back()
    assert not second.exists() or not os.listdir(second)
    oc, od, _, ost = oracle.render(s)
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    assert np.abs(c.astype(np.int16) - oc.astype(np.int16)).max() <= 1
    # a damaged entry (one flipped code byte: the checksum no longer
    # matches) is dropped; the others stay
    name0 = sorted(entries)[0]
    at = data.index(name0.encode()) + 48 + 100
    damaged = bytearray(data)
    damaged[at] ^= 0x5A
    dcache = _cache(ctx, bytes(damaged))
    got = _entries(_data(ctx, dcache))
    assert name0 not in got and set(got) == set(entries) - {name0}
    # a pipeline created with the damaged cache compiles the missing variant
    ctx.pipeline_cache = dcache
    r3 = SceneRenderer(ctx, s)
    r3.render(1)
    c3, d3, _ = r3.read_back()
    np.testing.assert_array_equal(d3.view(np.uint32), od.view(np.uint32))
    ctx.pipeline_cache = None
    ctx.d.vkDestroyPipelineCache(ctx.device, dcache, None)
    # data of another device or format is ignored
    bad = bytearray(data)
    bad[16] ^= 0xFF
    other = _cache(ctx, bytes(bad))
    assert _entries(_data(ctx, other)) == {}
    for h in (cache, loaded, merged, other):
        ctx.d.vkDestroyPipelineCache(ctx.device, h, None)
