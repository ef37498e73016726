"""ctypes bridge to the CPU oracle (oracle/raster_oracle.cpp).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and bench.py's
cpu_baseline leg, as the checker. It takes the same Scene objects the Vulkan
renderer takes (kazan_amd/scenes/generators.py) and returns colour, depth,
the resolved image and the counters.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libkazan_oracle.so")

c_u32, c_i32, c_f32, c_u64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64


class KzoState(ctypes.Structure):
    _fields_ = ([(n, c_u32) for n in ("width", "height", "samples", "clear_color_enable", "clear_depth_enable")] +
                [("clear_color", c_f32 * 4), ("clear_depth", c_f32)] +
                [(n, c_f32) for n in ("vp_x", "vp_y", "vp_w", "vp_h", "vp_min_depth", "vp_max_depth")] +
                [("sc_x", c_i32), ("sc_y", c_i32), ("sc_w", c_u32), ("sc_h", c_u32)] +
                [(n, c_u32) for n in ("cull_mode", "front_face", "depth_test", "depth_write", "depth_compare",
                                      "blend_enable", "src_color", "dst_color", "color_op", "src_alpha", "dst_alpha",
                                      "alpha_op", "write_mask", "vs_kind", "fs_kind")] +
                [("tex", ctypes.c_void_p)] +
                [(n, c_u32) for n in ("tex_w", "tex_h", "tex_filter", "tex_addr_u", "tex_addr_v")] +
                [("ubo", ctypes.c_void_p), ("num_lights", c_u32), ("pad0", c_u32)])


class KzoDraw(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("indices", ctypes.c_void_p), ("count", c_u32), ("first", c_u32),
                ("vertex_offset", c_i32), ("pad", c_u32), ("mvp", c_f32 * 16)]


class KzoStats(ctypes.Structure):
    _fields_ = [(n, c_u64) for n in ("triangles_in", "triangles_setup", "triangles_clipped", "fragments",
                                     "depth_passed")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("oracle not built: make -C oracle")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.kzo_render.restype = ctypes.c_int
        _lib.kzo_render.argtypes = [ctypes.c_void_p, ctypes.c_void_p, c_u32, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _lib.kzo_snap.restype = c_i32
        _lib.kzo_snap.argtypes = [c_f32]
        _lib.kzo_i64_to_f32.restype = c_f32
        _lib.kzo_i64_to_f32.argtypes = [ctypes.c_int64]
        _lib.kzo_encode_unorm8.restype = c_u32
        _lib.kzo_encode_unorm8.argtypes = [c_f32]
        _lib.kzo_bilinear.argtypes = [ctypes.c_void_p, c_u32, c_u32, c_u32, c_u32, c_f32, c_f32, ctypes.c_void_p]
        _lib.kzo_vertex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, c_u32, ctypes.c_void_p, ctypes.c_void_p]
        _lib.kzo_fragment.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.kzo_fragment.restype = ctypes.c_int
        _lib.kzo_clip_polygon.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.kzo_clip_polygon.restype = ctypes.c_int
    return _lib


def make_state(scene):
    s = scene
    st = KzoState()
    st.width, st.height, st.samples = s.width, s.height, s.samples
    st.clear_color_enable = st.clear_depth_enable = 1
    for k in range(4):
        st.clear_color[k] = s.clear_color[k]
    st.clear_depth = s.clear_depth
    st.vp_x, st.vp_y, st.vp_w, st.vp_h, st.vp_min_depth, st.vp_max_depth = s.viewport
    st.sc_x, st.sc_y, st.sc_w, st.sc_h = s.scissor
    st.cull_mode, st.front_face = s.cull_mode, s.front_face
    st.depth_test, st.depth_write, st.depth_compare = s.depth_test, s.depth_write, s.depth_op
    (st.blend_enable, st.src_color, st.dst_color, st.color_op, st.src_alpha, st.dst_alpha, st.alpha_op) = s.blend
    st.write_mask = s.write_mask
    st.vs_kind, st.fs_kind = s.vs, s.fs
    keep = []
    if s.texture is not None:
        tex = np.ascontiguousarray(s.texture)
        keep.append(tex)
        st.tex = tex.ctypes.data
        st.tex_h, st.tex_w = tex.shape[:2]
        st.tex_filter = s.tex_filter
        st.tex_addr_u, st.tex_addr_v = [0 if a == 0 else 2 for a in s.tex_addr]
    if s.ubo is not None:
        ubo = np.ascontiguousarray(s.ubo, dtype=np.float32)
        keep.append(ubo)
        st.ubo = ubo.ctypes.data
        st.num_lights = s.num_lights
    return st, keep


def vertex(scene, draw_index, vi):
    """The oracle's vertex shader on vertex `vi` of draw `draw_index`:
    (clip position [4], varyings [8]) as float32 arrays."""
    st, keep = make_state(scene)
    draws = _draws(scene)
    pos = np.zeros(4, np.float32)
    var = np.zeros(8, np.float32)
    lib().kzo_vertex(ctypes.byref(st), ctypes.byref(draws[draw_index]), vi, pos.ctypes.data, var.ctypes.data)
    del keep
    return pos, var


def clip_polygon(scene, pos):
    """The oracle's R-CLIP on three clip positions (3, 4): the polygon's
    vertex clip positions, shape (n, 4) float32."""
    st, keep = make_state(scene)
    p = np.ascontiguousarray(pos, np.float32).reshape(12)
    poly = np.zeros((12, 4), np.float32)
    n = lib().kzo_clip_polygon(ctypes.byref(st), p.ctypes.data, poly.ctypes.data)
    del keep
    return poly[:n].copy()


def fragment(scene, var):
    """The oracle's fragment shader on one set of varyings: unclamped RGBA,
    or None when the shader discards the fragment."""
    st, keep = make_state(scene)
    v = np.zeros(8, np.float32)
    v[:len(var)] = var
    out = np.zeros(4, np.float32)
    kept = lib().kzo_fragment(ctypes.byref(st), v.ctypes.data, out.ctypes.data)
    del keep
    return out if kept else None


def _draws(s):
    draws = (KzoDraw * len(s.draws))()
    for i, dr in enumerate(s.draws):
        draws[i].vertices = s.vertices.ctypes.data
        draws[i].indices = s.indices.ctypes.data if s.indices is not None else None
        draws[i].count = dr["count"]
        draws[i].first = dr["first"]
        draws[i].vertex_offset = dr["vertex_offset"]
        if dr.get("mvp") is not None:
            m = np.ascontiguousarray(dr["mvp"], dtype=np.float32).ravel()
            for k in range(16):
                draws[i].mvp[k] = m[k]
    return draws


def render(scene, threads=0):
    """Returns (colour (H,W,S,4) u8, depth (H,W,S) f32, resolved (H,W,4) u8, stats dict)."""
    s = scene
    st, keep = make_state(s)
    draws = _draws(s)
    W, H, S = s.width, s.height, s.samples
    color = np.zeros((H, W, S, 4), np.uint8)
    depth = np.zeros((H, W, S), np.float32)
    resolved = np.zeros((H, W, 4), np.uint8)
    stats = KzoStats()
    rc = lib().kzo_render(ctypes.byref(st), draws, len(s.draws), color.ctypes.data, depth.ctypes.data,
                          resolved.ctypes.data, ctypes.byref(stats), threads)
    if rc != 0:
        raise RuntimeError("oracle render failed: %d" % rc)
    del keep
    return color, depth, resolved, {k: getattr(stats, k) for k, _ in KzoStats._fields_}
