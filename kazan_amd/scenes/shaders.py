"""The scene shaders, assembled as SPIR-V modules.

Each function returns the module's bytes; the GLSL it encodes is in the
docstring. The CPU oracle restates the same arithmetic in the same order
(oracle/raster_oracle.cpp: run_vs / run_fs).
"""
from . import spirv as sp


def _vertex_module(build_body):
    m = sp.Module()
    m.capability(sp.CapabilityShader)
    m.set_memory_model()
    iface = []
    fn_holder = {}

    def var(storage, pointee, name, location=None, builtin=None):
        v = m.variable(storage, pointee, name)
        if location is not None:
            m.decorate(v, sp.DecorationLocation, location)
        if builtin is not None:
            m.decorate(v, sp.DecorationBuiltIn, builtin)
        if storage in (sp.StorageInput, sp.StorageOutput):
            iface.append(v)
        return v

    f = build_body(m, var)
    m.entry_point(sp.ExecutionModelVertex, f.id, "main", iface)
    fn_holder["f"] = f
    return m


def _fragment_module(build_body):
    m = sp.Module()
    m.capability(sp.CapabilityShader)
    m.set_memory_model()
    iface = []

    def var(storage, pointee, name, location=None, builtin=None):
        v = m.variable(storage, pointee, name)
        if location is not None:
            m.decorate(v, sp.DecorationLocation, location)
        if builtin is not None:
            m.decorate(v, sp.DecorationBuiltIn, builtin)
        if storage in (sp.StorageInput, sp.StorageOutput):
            iface.append(v)
        return v

    f = build_body(m, var)
    m.entry_point(sp.ExecutionModelFragment, f.id, "main", iface)
    m.execution_mode(f.id, sp.ExecutionModeOriginUpperLeft)
    return m


def vs_pos4_color4():
    """layout(location=0) in vec4 pos; layout(location=1) in vec4 color;
    layout(location=0) out vec4 v_color;
    void main() { gl_Position = pos; v_color = color; }"""

    def body(m, var):
        v4 = m.t_vec(4)
        pos_in = var(sp.StorageInput, v4, "pos", location=0)
        col_in = var(sp.StorageInput, v4, "color", location=1)
        pos_out = var(sp.StorageOutput, v4, "gl_Position", builtin=sp.BuiltInPosition)
        col_out = var(sp.StorageOutput, v4, "v_color", location=0)
        f = sp.Function(m)
        f.store(pos_out, f.load(v4, pos_in))
        f.store(col_out, f.load(v4, col_in))
        f.ret()
        return f

    return _vertex_module(body).to_bytes()


def vs_pos3_uv2():
    """layout(location=0) in vec3 pos; layout(location=1) in vec2 uv;
    layout(location=0) out vec2 v_uv;
    void main() { gl_Position = vec4(pos, 1.0); v_uv = uv; }"""

    def body(m, var):
        v2, v3, v4, f32 = m.t_vec(2), m.t_vec(3), m.t_vec(4), m.t_float()
        pos_in = var(sp.StorageInput, v3, "pos", location=0)
        uv_in = var(sp.StorageInput, v2, "uv", location=1)
        pos_out = var(sp.StorageOutput, v4, "gl_Position", builtin=sp.BuiltInPosition)
        uv_out = var(sp.StorageOutput, v2, "v_uv", location=0)
        one = m.c_float(1.0)
        f = sp.Function(m)
        p = f.load(v3, pos_in)
        x, y, z = (f.extract(f32, p, i) for i in range(3))
        f.store(pos_out, f.construct(v4, x, y, z, one))
        f.store(uv_out, f.load(v2, uv_in))
        f.ret()
        return f

    return _vertex_module(body).to_bytes()


def _push_mvp(m):
    """push_constant block { mat4 mvp; } (column-major, stride 16)."""
    mat4 = m.t_mat4()
    blk = m.t_struct(mat4)
    m.decorate(blk, sp.DecorationBlock)
    m.member_decorate(blk, 0, sp.DecorationOffset, 0)
    m.member_decorate(blk, 0, sp.DecorationColMajor)
    m.member_decorate(blk, 0, sp.DecorationMatrixStride, 16)
    pc = m.variable(sp.StoragePushConstant, blk, "pc")
    return pc, mat4


def vs_mvp_pos3_color4():
    """layout(push_constant) uniform PC { mat4 mvp; } pc;
    layout(location=0) in vec3 pos; layout(location=1) in vec4 color;
    layout(location=0) out vec4 v_color;
    void main() { gl_Position = pc.mvp * vec4(pos, 1.0); v_color = color; }"""

    def body(m, var):
        v3, v4, f32 = m.t_vec(3), m.t_vec(4), m.t_float()
        pc, mat4 = _push_mvp(m)
        pos_in = var(sp.StorageInput, v3, "pos", location=0)
        col_in = var(sp.StorageInput, v4, "color", location=1)
        pos_out = var(sp.StorageOutput, v4, "gl_Position", builtin=sp.BuiltInPosition)
        col_out = var(sp.StorageOutput, v4, "v_color", location=0)
        one = m.c_float(1.0)
        zero_i = m.c_int(0)
        f = sp.Function(m)
        mvp = f.load(mat4, f.access_chain(m.t_ptr(sp.StoragePushConstant, mat4), pc, zero_i))
        p = f.load(v3, pos_in)
        x, y, z = (f.extract(f32, p, i) for i in range(3))
        f.store(pos_out, f.mat_times_vec(v4, mvp, f.construct(v4, x, y, z, one)))
        f.store(col_out, f.load(v4, col_in))
        f.ret()
        return f

    return _vertex_module(body).to_bytes()


def vs_mvp_pos3_normal3():
    """layout(push_constant) uniform PC { mat4 mvp; } pc;
    layout(location=0) in vec3 pos; layout(location=1) in vec3 normal;
    layout(location=0) out vec3 v_pos; layout(location=1) out vec3 v_normal;
    void main() { gl_Position = pc.mvp * vec4(pos, 1.0); v_pos = pos; v_normal = normal; }"""

    def body(m, var):
        v3, v4, f32 = m.t_vec(3), m.t_vec(4), m.t_float()
        pc, mat4 = _push_mvp(m)
        pos_in = var(sp.StorageInput, v3, "pos", location=0)
        nrm_in = var(sp.StorageInput, v3, "normal", location=1)
        pos_out = var(sp.StorageOutput, v4, "gl_Position", builtin=sp.BuiltInPosition)
        vpos_out = var(sp.StorageOutput, v3, "v_pos", location=0)
        vnrm_out = var(sp.StorageOutput, v3, "v_normal", location=1)
        one = m.c_float(1.0)
        zero_i = m.c_int(0)
        f = sp.Function(m)
        mvp = f.load(mat4, f.access_chain(m.t_ptr(sp.StoragePushConstant, mat4), pc, zero_i))
        p = f.load(v3, pos_in)
        x, y, z = (f.extract(f32, p, i) for i in range(3))
        f.store(pos_out, f.mat_times_vec(v4, mvp, f.construct(v4, x, y, z, one)))
        f.store(vpos_out, p)
        f.store(vnrm_out, f.load(v3, nrm_in))
        f.ret()
        return f

    return _vertex_module(body).to_bytes()


def fs_color():
    """layout(location=0) in vec4 v_color; layout(location=0) out vec4 o;
    void main() { o = v_color; }"""

    def body(m, var):
        v4 = m.t_vec(4)
        col_in = var(sp.StorageInput, v4, "v_color", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        f.store(out, f.load(v4, col_in))
        f.ret()
        return f

    return _fragment_module(body).to_bytes()


def fs_texture():
    """layout(set=0, binding=0) uniform sampler2D tex;
    layout(location=0) in vec2 v_uv; layout(location=0) out vec4 o;
    void main() { o = texture(tex, v_uv); }"""

    def body(m, var):
        v2, v4 = m.t_vec(2), m.t_vec(4)
        si = m.t_sampled_image_2d()
        tex = m.variable(sp.StorageUniformConstant, si, "tex")
        m.decorate(tex, sp.DecorationDescriptorSet, 0)
        m.decorate(tex, sp.DecorationBinding, 0)
        uv_in = var(sp.StorageInput, v2, "v_uv", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        s = f.load(si, tex)
        uv = f.load(v2, uv_in)
        f.store(out, f.sample(v4, s, uv))
        f.ret()
        return f

    return _fragment_module(body).to_bytes()


def fs_phong(num_lights=8):
    """layout(set=0, binding=0) uniform Lights { vec4 pos[N]; vec4 color[N]; vec4 ambient; } L;
    layout(location=0) in vec3 v_pos; layout(location=1) in vec3 v_normal;
    layout(location=0) out vec4 o;
    void main() {
      vec3 n = v_normal * (1.0 / sqrt(dot(v_normal, v_normal)));
      vec3 c = L.ambient.xyz;
      for (int i = 0; i < N; ++i) {
        vec3 Lv = L.pos[i].xyz - v_pos;
        float d2 = dot(Lv, Lv);
        vec3 l = Lv * (1.0 / sqrt(d2));
        float ndl = max(dot(n, l), 0.0);
        vec3 hv = vec3(l.x, l.y, l.z + 1.0);
        vec3 h = hv * (1.0 / sqrt(dot(hv, hv)));
        float s = max(dot(n, h), 0.0); s = s*s; s = s*s; s = s*s; s = s*s;
        float att = 1.0 / (1.0 + d2);
        float k = (ndl + s) * att;
        c = c + L.color[i].xyz * k;
      }
      o = vec4(c, 1.0);
    }
    The loop is a structured SPIR-V loop as an optimising compiler writes it:
    OpLoopMerge in the header, the counter and the colour carried by OpPhi,
    the exit test in its own block, the increment in the continue target, and
    L.pos[i] / L.color[i] as run-time access-chain indices."""

    def body(m, var):
        f32, v3, v4, i32, b = m.t_float(), m.t_vec(3), m.t_vec(4), m.t_int(), m.t_bool()
        arr = m.t_array(v4, m.c_int(num_lights, signed=False))
        m.decorate(arr, sp.DecorationArrayStride, 16)
        blk = m.t_struct(arr, arr, v4)
        m.decorate(blk, sp.DecorationBlock)
        m.member_decorate(blk, 0, sp.DecorationOffset, 0)
        m.member_decorate(blk, 1, sp.DecorationOffset, 16 * num_lights)
        m.member_decorate(blk, 2, sp.DecorationOffset, 32 * num_lights)
        ubo = m.variable(sp.StorageUniform, blk, "L")
        m.decorate(ubo, sp.DecorationDescriptorSet, 0)
        m.decorate(ubo, sp.DecorationBinding, 0)
        vpos_in = var(sp.StorageInput, v3, "v_pos", location=0)
        vnrm_in = var(sp.StorageInput, v3, "v_normal", location=1)
        out = var(sp.StorageOutput, v4, "o", location=0)
        one, zero = m.c_float(1.0), m.c_float(0.0)
        pv4 = m.t_ptr(sp.StorageUniform, v4)
        f = sp.Function(m)
        entry = None
        vp = f.load(v3, vpos_in)
        vn = f.load(v3, vnrm_in)
        rn = f.fdiv(f32, one, f.ext(f32, sp.GLSL_Sqrt, f.dot(vn, vn)))
        n = f.vec_times_scalar(v3, vn, rn)
        amb = f.load(v4, f.access_chain(pv4, ubo, m.c_int(2)))
        c0 = f.shuffle(v3, amb, amb, 0, 1, 2)
        header, cond, loop_body, cont, merge = (f.new_label() for _ in range(5))
        entry = f.entry_label
        f.branch(header)
        # header: the carried counter and colour
        f.label(header)
        i_next, c_next = m.id(), m.id()  # defined in the continue target and the body, used by the phis
        i = f.phi(i32, (m.c_int(0), entry), (i_next, cont))
        c = f.phi(v3, (c0, entry), (c_next, cont))
        f.loop_merge(merge, cont)
        f.branch(cond)
        f.label(cond)
        f.branch_cond(f.binop(sp.OpSLessThan, b, i, m.c_int(num_lights)), loop_body, merge)
        f.label(loop_body)
        lp = f.load(v4, f.access_chain(pv4, ubo, m.c_int(0), i))
        lc = f.load(v4, f.access_chain(pv4, ubo, m.c_int(1), i))
        Lv = f.fsub(v3, f.shuffle(v3, lp, lp, 0, 1, 2), vp)
        d2 = f.dot(Lv, Lv)
        l = f.vec_times_scalar(v3, Lv, f.fdiv(f32, one, f.ext(f32, sp.GLSL_Sqrt, d2)))
        ndl = f.ext(f32, sp.GLSL_FMax, f.dot(n, l), zero)
        lz = f.extract(f32, l, 2)
        hv = f.construct(v3, f.extract(f32, l, 0), f.extract(f32, l, 1), f.fadd(f32, lz, one))
        h = f.vec_times_scalar(v3, hv, f.fdiv(f32, one, f.ext(f32, sp.GLSL_Sqrt, f.dot(hv, hv))))
        sv = f.ext(f32, sp.GLSL_FMax, f.dot(n, h), zero)
        for _ in range(4):
            sv = f.fmul(f32, sv, sv)
        att = f.fdiv(f32, one, f.fadd(f32, one, d2))
        k = f.fmul(f32, f.fadd(f32, ndl, sv), att)
        f.result_into(c_next, sp.OpFAdd, v3, c, f.vec_times_scalar(v3, f.shuffle(v3, lc, lc, 0, 1, 2), k))
        f.branch(cont)
        f.label(cont)
        f.result_into(i_next, sp.OpIAdd, i32, i, m.c_int(1))
        f.branch(header)
        f.label(merge)
        cx, cy, cz = (f.extract(f32, c, j) for j in range(3))
        f.store(out, f.construct(v4, cx, cy, cz, one))
        f.ret()
        return f

    return _fragment_module(body).to_bytes()


def fs_alpha_test():
    """layout(location=0) in vec4 v_color; layout(location=0) out vec4 o;
    void main() { if (v_color.a < 0.5) discard; o = v_color; }"""

    def body(m, var):
        f32, v4, b = m.t_float(), m.t_vec(4), m.t_bool()
        col_in = var(sp.StorageInput, v4, "v_color", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        c = f.load(v4, col_in)
        lt = f.binop(sp.OpFOrdLessThan, b, f.extract(f32, c, 3), m.c_float(0.5))
        kill, merge = f.new_label(), f.new_label()
        f.selection_merge(merge)
        f.branch_cond(lt, kill, merge)
        f.label(kill)
        f.kill()
        f.label(merge)
        f.store(out, c)
        f.ret()
        return f

    return _fragment_module(body).to_bytes()
