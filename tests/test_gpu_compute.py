"""Compute pipelines through the ICD on the GPU (vkCreateComputePipelines,
vkCmdBindDescriptorSets / vkCmdPushConstants on the compute bind point,
vkCmdDispatch, vkCmdDispatchBase; the reference's api_impl.rs entry points
and ComputePipeline, vulkan-driver/src/pipeline.rs:122-154).

Each kernel's result is restated in numpy in the same arithmetic order and
compared bit for bit:
- saxpy with push constants, a count that is no multiple of the workgroup
  (the invocations past it write nothing), and vkCmdDispatchBase (only the
  base's workgroups run, and they see the offset workgroup ids);
- a workgroup reduction through shared memory with barriers inside a loop
  (the same pairwise tree in numpy);
- a histogram of atomic adds (np.bincount);
- OpArrayLength of a run-time array after a 16-byte header and of a plain
  one, under bound ranges shorter than the buffer and VK_WHOLE_SIZE;
- a dispatch that writes a vertex buffer, drawn by the same command buffer:
  the frame equals the oracle's render of the vertices numpy computes.
"""
import numpy as np
import pytest

from kazan_amd.scenes import compute_shaders as cs
from kazan_amd.scenes import generators as g
from kazan_amd.scenes.compute import ComputeKernel, read_buffer

pytestmark = pytest.mark.gpu


def _push(*vals):
    out = np.zeros(len(vals), np.uint32)
    for i, v in enumerate(vals):
        out[i] = np.float32(v).view(np.uint32) if isinstance(v, float) else v
    return out


def test_saxpy(ctx):
    n, pad = 1000, 64
    r = np.random.default_rng(1)
    x = r.standard_normal(n + pad).astype(np.float32)
    y = r.standard_normal(n + pad).astype(np.float32)
    a = 2.5
    xb, yb = ctx.device_buffer_with(x), ctx.device_buffer_with(y)
    k = ComputeKernel(ctx, cs.cs_saxpy(), 2, push_bytes=8)
    try:
        k.bind([(xb[0], x.nbytes), (yb[0], y.nbytes)])
        k.run(((n + 63) // 64,), push=_push(a, n))
        got = read_buffer(ctx, yb[0], y.nbytes, np.float32)
        want = y.copy()
        want[:n] = np.float32(a) * x[:n] + y[:n]
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    finally:
        k.destroy()
        ctx.destroy_buffer(xb)
        ctx.destroy_buffer(yb)


def test_dispatch_base(ctx):
    n = 1024
    x = np.arange(n, dtype=np.float32)
    y = np.ones(n, np.float32)
    xb, yb = ctx.device_buffer_with(x), ctx.device_buffer_with(y)
    k = ComputeKernel(ctx, cs.cs_saxpy(), 2, push_bytes=8)
    try:
        k.bind([(xb[0], x.nbytes), (yb[0], y.nbytes)])
        k.run((4,), push=_push(1.0, n), base=(4, 0, 0))  # workgroups 4 .. 7: elements 256 .. 511
        got = read_buffer(ctx, yb[0], y.nbytes, np.float32)
        want = y.copy()
        want[256:512] = x[256:512] + 1.0
        np.testing.assert_array_equal(got, want)
    finally:
        k.destroy()
        ctx.destroy_buffer(xb)
        ctx.destroy_buffer(yb)


def test_workgroup_reduction_with_barriers(ctx):
    groups, G = 40, cs.REDUCE_GROUP
    v = np.random.default_rng(2).uniform(-1, 1, groups * G).astype(np.float32)
    vb = ctx.device_buffer_with(v)
    sb = ctx.device_buffer_with(np.zeros(groups, np.float32))
    k = ComputeKernel(ctx, cs.cs_reduce(), 2)
    try:
        k.bind([(vb[0], v.nbytes), (sb[0], groups * 4)])
        k.run((groups,))
        got = read_buffer(ctx, sb[0], groups * 4, np.float32)
        s = v.reshape(groups, G).copy()
        kk = G // 2
        while kk > 0:
            s[:, :kk] = s[:, :kk] + s[:, kk:2 * kk]
            kk >>= 1
        np.testing.assert_array_equal(got, s[:, 0])
        # and it is the sum, to float32 accuracy
        np.testing.assert_allclose(got, v.reshape(groups, G).astype(np.float64).sum(1), atol=1e-4)
    finally:
        k.destroy()
        ctx.destroy_buffer(vb)
        ctx.destroy_buffer(sb)


def test_atomic_histogram(ctx):
    n = 100_000
    v = np.random.default_rng(3).integers(0, 2**32, n, dtype=np.uint32)
    vb = ctx.device_buffer_with(v)
    hb = ctx.device_buffer_with(np.zeros(256, np.uint32))
    k = ComputeKernel(ctx, cs.cs_histogram(), 2, push_bytes=4)
    try:
        k.bind([(vb[0], v.nbytes), (hb[0], 1024)])
        k.run(((n + 63) // 64,), push=_push(n))
        got = read_buffer(ctx, hb[0], 1024, np.uint32)
        np.testing.assert_array_equal(got, np.bincount(v & 255, minlength=256))
        # a second dispatch adds to the counts (no clear in between)
        k.run(((n + 63) // 64,), push=_push(n))
        got2 = read_buffer(ctx, hb[0], 1024, np.uint32)
        np.testing.assert_array_equal(got2, 2 * np.bincount(v & 255, minlength=256))
    finally:
        k.destroy()
        ctx.destroy_buffer(vb)
        ctx.destroy_buffer(hb)


def _grid_vertices(n):
    """cs_grid_vertices restated: (3 n, 8) float32, position then colour."""
    t = np.arange(n, dtype=np.uint32)
    cx = (t % 16).astype(np.float32)
    cy = (t // 16).astype(np.float32)
    x0 = cx * np.float32(0.125) - np.float32(1.0)
    y0 = cy * np.float32(0.125) - np.float32(1.0)
    x1, y1 = x0 + np.float32(0.125), y0 + np.float32(0.125)
    col = np.stack([cx * np.float32(0.0625), cy * np.float32(0.0625), np.full(n, 0.5, np.float32),
                    np.ones(n, np.float32)], 1)
    out = np.zeros((n, 3, 8), np.float32)
    for k, (px, py) in enumerate(((x0, y0), (x1, y0), (x0, y1))):
        out[:, k, 0], out[:, k, 1], out[:, k, 2], out[:, k, 3] = px, py, 0.5, 1.0
        out[:, k, 4:8] = col
    return out.reshape(3 * n, 8)


def test_compute_writes_vertices_then_draw(ctx, oracle):
    from kazan_amd.scenes.renderer import SceneRenderer
    n = 256
    want_v = _grid_vertices(n)
    gpu_scene = g.Scene("compute_grid", 64, 64, g.VS_POS4_COLOR4, g.FS_COLOR, np.zeros_like(want_v))
    k = ComputeKernel(ctx, cs.cs_grid_vertices(), 1, push_bytes=4)
    try:
        def before_pass(rend, cb):
            k.bind([(rend.vbuf[0], want_v.nbytes)])
            k.record(cb, ((n + 63) // 64,), push=_push(n))

        r = SceneRenderer(ctx, gpu_scene, before_pass=before_pass)
        st = r.render(2)
        c, d, _ = r.read_back()
        got_v = read_buffer(ctx, r.vbuf[0], want_v.nbytes, np.float32).reshape(want_v.shape)
        np.testing.assert_array_equal(got_v, want_v)
        oc, od, _, ost = oracle.render(g.Scene("compute_grid", 64, 64, g.VS_POS4_COLOR4, g.FS_COLOR, want_v))
        assert st["fragments"] == 2 * ost["fragments"]
        np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
        assert np.abs(c.astype(np.int16) - oc.astype(np.int16)).max() <= 1
    finally:
        k.destroy()


@pytest.mark.parametrize("range_a,range_o,want", [
    (1000, 64, ((1000 - 16) // 16, 16)),   # ranges shorter than the buffers
    (8, 4096, (0, 1024)),                  # a range that ends before the array: 0 elements
    (None, None, ((4096 - 16) // 16, 1024)),  # VK_WHOLE_SIZE: to the end of each buffer
])
def test_array_length(ctx, range_a, range_o, want):
    whole = 0xFFFFFFFFFFFFFFFF
    ab = ctx.device_buffer_with(np.zeros(1024, np.uint32))
    ob = ctx.device_buffer_with(np.full(1024, 0xDEADBEEF, np.uint32))
    k = ComputeKernel(ctx, cs.cs_array_length(), 2)
    try:
        k.bind([(ab[0], whole if range_a is None else range_a), (ob[0], whole if range_o is None else range_o)])
        k.run((1,))
        got = read_buffer(ctx, ob[0], 4096, np.uint32)
        assert (int(got[0]), int(got[1])) == want
        assert (got[2:] == 0xDEADBEEF).all()
    finally:
        k.destroy()
        ctx.destroy_buffer(ab)
        ctx.destroy_buffer(ob)
