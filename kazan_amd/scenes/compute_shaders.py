"""Compute shaders, assembled as SPIR-V 1.0 modules (GLCompute entry points,
storage buffers as Uniform blocks decorated BufferBlock).

Each function returns the module's bytes; the GLSL it encodes is in the
docstring. The tests restate each one's arithmetic in numpy, in the same
order, so that results compare bit for bit.
"""
from . import spirv as sp


def _compute_module(local_size, build_body):
    m = sp.Module()
    m.capability(sp.CapabilityShader)
    m.set_memory_model()
    iface = []

    def builtin(b, name):
        v = m.variable(sp.StorageInput, m.t_uvec(3), name)
        m.decorate(v, sp.DecorationBuiltIn, b)
        iface.append(v)
        return v

    f = build_body(m, builtin)
    m.entry_point(sp.ExecutionModelGLCompute, f.id, "main", iface)
    m.execution_mode(f.id, sp.ExecutionModeLocalSize, *local_size)
    return m


def _ssbo(m, elem, stride, binding, name, set_=0):
    """`layout(set, binding) buffer { elem data[]; }`: the variable id."""
    arr = m.t_runtime_array(elem)
    m.decorate(arr, sp.DecorationArrayStride, stride)
    st = m.t_struct(arr)
    m.decorate(st, sp.DecorationBufferBlock)
    m.member_decorate(st, 0, sp.DecorationOffset, 0)
    v = m.variable(sp.StorageUniform, st, name)
    m.decorate(v, sp.DecorationDescriptorSet, set_)
    m.decorate(v, sp.DecorationBinding, binding)
    return v


def _push_block(m, members, name="pc"):
    """`layout(push_constant) uniform { members }` with 4-byte members at
    offsets 0, 4, ...: the variable id."""
    st = m.t_struct(*members)
    m.decorate(st, sp.DecorationBlock)
    for i in range(len(members)):
        m.member_decorate(st, i, sp.DecorationOffset, 4 * i)
    return m.variable(sp.StoragePushConstant, st, name)


def _elem(f, m, var, elem, idx):
    return f.access_chain(m.t_ptr(sp.StorageUniform, elem), var, m.c_int(0), idx)


def _gid_x(f, m, gid_var):
    return f.extract(m.t_int(False), f.load(m.t_uvec(3), gid_var), 0)


def cs_saxpy():
    """layout(local_size_x = 64) in;
    layout(binding = 0) buffer X { float x[]; };
    layout(binding = 1) buffer Y { float y[]; };
    layout(push_constant) uniform P { float a; uint n; } pc;
    void main() {
        uint i = gl_GlobalInvocationID.x;
        if (i < pc.n) y[i] = pc.a * x[i] + y[i];
    }"""

    def body(m, builtin):
        fl, u = m.t_float(), m.t_int(False)
        gid = builtin(sp.BuiltInGlobalInvocationId, "gl_GlobalInvocationID")
        x = _ssbo(m, fl, 4, 0, "x")
        y = _ssbo(m, fl, 4, 1, "y")
        pc = _push_block(m, [fl, u])
        f = sp.Function(m)
        i = _gid_x(f, m, gid)
        n = f.load(u, f.access_chain(m.t_ptr(sp.StoragePushConstant, u), pc, m.c_int(1)))
        cond = f.binop(sp.OpULessThan, m.t_bool(), i, n)
        then, merge = f.new_label(), f.new_label()
        f.selection_merge(merge)
        f.branch_cond(cond, then, merge)
        f.label(then)
        a = f.load(fl, f.access_chain(m.t_ptr(sp.StoragePushConstant, fl), pc, m.c_int(0)))
        yp = _elem(f, m, y, fl, i)
        r = f.fadd(fl, f.fmul(fl, a, f.load(fl, _elem(f, m, x, fl, i))), f.load(fl, yp))
        f.store(yp, r)
        f.branch(merge)
        f.label(merge)
        f.ret()
        return f

    return _compute_module((64, 1, 1), body).to_bytes()


REDUCE_GROUP = 256


def cs_reduce():
    """layout(local_size_x = 256) in;
    layout(binding = 0) buffer In { float v[]; };
    layout(binding = 1) buffer Out { float sums[]; };
    shared float s[256];
    void main() {
        uint l = gl_LocalInvocationID.x;
        s[l] = v[gl_GlobalInvocationID.x];
        barrier();
        for (uint k = 128; k > 0; k >>= 1) {
            if (l < k) s[l] = s[l] + s[l + k];
            barrier();
        }
        if (l == 0) sums[gl_WorkGroupID.x] = s[0];
    }"""

    def body(m, builtin):
        fl, u, b = m.t_float(), m.t_int(False), m.t_bool()
        gid = builtin(sp.BuiltInGlobalInvocationId, "gl_GlobalInvocationID")
        lid = builtin(sp.BuiltInLocalInvocationId, "gl_LocalInvocationID")
        wid = builtin(sp.BuiltInWorkgroupId, "gl_WorkGroupID")
        v = _ssbo(m, fl, 4, 0, "v")
        sums = _ssbo(m, fl, 4, 1, "sums")
        shared = m.variable(sp.StorageWorkgroup, m.t_array(fl, m.c_int(REDUCE_GROUP, signed=False)), "s")
        ptr_s = m.t_ptr(sp.StorageWorkgroup, fl)
        wg = m.c_int(sp.ScopeWorkgroup, signed=False)
        sem = m.c_int(sp.SemanticsAcquireRelease | sp.SemanticsWorkgroupMemory, signed=False)
        f = sp.Function(m)
        k = f.var(u)
        l = f.extract(u, f.load(m.t_uvec(3), lid), 0)
        f.store(f.access_chain(ptr_s, shared, l), f.load(fl, _elem(f, m, v, fl, _gid_x(f, m, gid))))
        f.barrier(wg, wg, sem)
        f.store(k, m.c_int(REDUCE_GROUP // 2, signed=False))
        head, body_l, cont, merge = f.new_label(), f.new_label(), f.new_label(), f.new_label()
        f.branch(head)
        f.label(head)
        kv = f.load(u, k)
        f.loop_merge(merge, cont)
        f.branch_cond(f.binop(sp.OpUGreaterThan, b, kv, m.c_int(0, signed=False)), body_l, merge)
        f.label(body_l)
        add, add_merge = f.new_label(), f.new_label()
        f.selection_merge(add_merge)
        f.branch_cond(f.binop(sp.OpULessThan, b, l, kv), add, add_merge)
        f.label(add)
        sl = f.access_chain(ptr_s, shared, l)
        other = f.load(fl, f.access_chain(ptr_s, shared, f.binop(sp.OpIAdd, u, l, kv)))
        f.store(sl, f.fadd(fl, f.load(fl, sl), other))
        f.branch(add_merge)
        f.label(add_merge)
        f.barrier(wg, wg, sem)
        f.branch(cont)
        f.label(cont)
        f.store(k, f.binop(sp.OpShiftRightLogical, u, f.load(u, k), m.c_int(1, signed=False)))
        f.branch(head)
        f.label(merge)
        first, end = f.new_label(), f.new_label()
        f.selection_merge(end)
        f.branch_cond(f.binop(sp.OpIEqual, b, l, m.c_int(0, signed=False)), first, end)
        f.label(first)
        w = f.extract(u, f.load(m.t_uvec(3), wid), 0)
        f.store(_elem(f, m, sums, fl, w), f.load(fl, f.access_chain(ptr_s, shared, m.c_int(0, signed=False))))
        f.branch(end)
        f.label(end)
        f.ret()
        return f

    return _compute_module((REDUCE_GROUP, 1, 1), body).to_bytes()


def cs_histogram():
    """layout(local_size_x = 64) in;
    layout(binding = 0) buffer In { uint v[]; };
    layout(binding = 1) buffer Hist { uint h[]; };
    layout(push_constant) uniform P { uint n; } pc;
    void main() {
        uint i = gl_GlobalInvocationID.x;
        if (i < pc.n) atomicAdd(h[v[i] & 255u], 1u);
    }"""

    def body(m, builtin):
        u, b = m.t_int(False), m.t_bool()
        gid = builtin(sp.BuiltInGlobalInvocationId, "gl_GlobalInvocationID")
        v = _ssbo(m, u, 4, 0, "v")
        h = _ssbo(m, u, 4, 1, "h")
        pc = _push_block(m, [u])
        f = sp.Function(m)
        i = _gid_x(f, m, gid)
        n = f.load(u, f.access_chain(m.t_ptr(sp.StoragePushConstant, u), pc, m.c_int(0)))
        then, merge = f.new_label(), f.new_label()
        f.selection_merge(merge)
        f.branch_cond(f.binop(sp.OpULessThan, b, i, n), then, merge)
        f.label(then)
        bucket = f.binop(sp.OpBitwiseAnd, u, f.load(u, _elem(f, m, v, u, i)), m.c_int(255, signed=False))
        f.atomic_iadd(u, _elem(f, m, h, u, bucket), m.c_int(sp.ScopeDevice, signed=False),
                      m.c_int(0, signed=False), m.c_int(1, signed=False))
        f.branch(merge)
        f.label(merge)
        f.ret()
        return f

    return _compute_module((64, 1, 1), body).to_bytes()


GRID_CELLS = 16


def cs_grid_vertices():
    """Writes a triangle list for VS_POS4_COLOR4 (per vertex: position,
    colour): triangle t covers the lower-left half of cell (t % 16, t / 16)
    of a 16 x 16 grid over NDC [-1, 1]^2.
    layout(local_size_x = 64) in;
    layout(binding = 0) buffer V { vec4 v[]; };
    layout(push_constant) uniform P { uint n; } pc;
    void main() {
        uint t = gl_GlobalInvocationID.x;
        if (t < pc.n) {
            float cx = float(t % 16u), cy = float(t / 16u);
            float x0 = cx * 0.125 - 1.0, y0 = cy * 0.125 - 1.0;
            vec4 col = vec4(cx * 0.0625, cy * 0.0625, 0.5, 1.0);
            v[6u * t + 0u] = vec4(x0, y0, 0.5, 1.0);          v[6u * t + 1u] = col;
            v[6u * t + 2u] = vec4(x0 + 0.125, y0, 0.5, 1.0);  v[6u * t + 3u] = col;
            v[6u * t + 4u] = vec4(x0, y0 + 0.125, 0.5, 1.0);  v[6u * t + 5u] = col;
        }
    }"""

    def body(m, builtin):
        fl, u, b, v4 = m.t_float(), m.t_int(False), m.t_bool(), m.t_vec(4)
        gid = builtin(sp.BuiltInGlobalInvocationId, "gl_GlobalInvocationID")
        vb = _ssbo(m, v4, 16, 0, "v")
        pc = _push_block(m, [u])
        f = sp.Function(m)
        t = _gid_x(f, m, gid)
        n = f.load(u, f.access_chain(m.t_ptr(sp.StoragePushConstant, u), pc, m.c_int(0)))
        then, merge = f.new_label(), f.new_label()
        f.selection_merge(merge)
        f.branch_cond(f.binop(sp.OpULessThan, b, t, n), then, merge)
        f.label(then)
        cells = m.c_int(GRID_CELLS, signed=False)
        cx = f._result(sp.OpConvertUToF, fl, f.binop(sp.OpUMod, u, t, cells))
        cy = f._result(sp.OpConvertUToF, fl, f.binop(sp.OpUDiv, u, t, cells))
        step, one, half = m.c_float(0.125), m.c_float(1.0), m.c_float(0.5)
        x0 = f.fsub(fl, f.fmul(fl, cx, step), one)
        y0 = f.fsub(fl, f.fmul(fl, cy, step), one)
        x1 = f.fadd(fl, x0, step)
        y1 = f.fadd(fl, y0, step)
        col = f.construct(v4, f.fmul(fl, cx, m.c_float(0.0625)), f.fmul(fl, cy, m.c_float(0.0625)), half, one)
        base = f.binop(sp.OpIMul, u, m.c_int(6, signed=False), t)
        for k, (px, py) in enumerate(((x0, y0), (x1, y0), (x0, y1))):
            pos = f.construct(v4, px, py, half, one)
            f.store(_elem(f, m, vb, v4, f.binop(sp.OpIAdd, u, base, m.c_int(2 * k, signed=False))), pos)
            f.store(_elem(f, m, vb, v4, f.binop(sp.OpIAdd, u, base, m.c_int(2 * k + 1, signed=False))), col)
        f.branch(merge)
        f.label(merge)
        f.ret()
        return f

    return _compute_module((64, 1, 1), body).to_bytes()


def cs_array_length():
    """layout(local_size_x = 1) in;
    layout(binding = 0) buffer A { uint count; vec4 v[]; };   // v at offset 16, stride 16
    layout(binding = 1) buffer B { uint o[]; };
    void main() { o[0] = v.length(); o[1] = o.length(); }"""

    def body(m, builtin):
        u, v4 = m.t_int(False), m.t_vec(4)
        arr = m.t_runtime_array(v4)
        m.decorate(arr, sp.DecorationArrayStride, 16)
        st = m.t_struct(u, arr)
        m.decorate(st, sp.DecorationBufferBlock)
        m.member_decorate(st, 0, sp.DecorationOffset, 0)
        m.member_decorate(st, 1, sp.DecorationOffset, 16)
        a = m.variable(sp.StorageUniform, st, "a")
        m.decorate(a, sp.DecorationDescriptorSet, 0)
        m.decorate(a, sp.DecorationBinding, 0)
        ob = _ssbo(m, u, 4, 1, "o")
        f = sp.Function(m)
        na = f.op(sp.OpArrayLength, u, a, 1)
        no = f.op(sp.OpArrayLength, u, ob, 0)
        f.store(_elem(f, m, ob, u, m.c_int(0)), na)
        f.store(_elem(f, m, ob, u, m.c_int(1)), no)
        f.ret()
        return f

    return _compute_module((1, 1, 1), body).to_bytes()
