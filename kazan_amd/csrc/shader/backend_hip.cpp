// backend_hip.cpp — shader IR -> HIP device functions.
//
// Every IR value becomes one C++ local per component; every pointer is
// resolved at compile time to (storage root, constant byte offset), since the
// translator only produces constant access chains. Float operations are
// emitted one rounding at a time in IR order and the kernel is compiled with
// -ffp-contract=off and IEEE division / square root, so the generated code
// computes exactly what the IR (and the CPU oracle's restatement) says.
#include "backend_hip.h"

#include <cstdio>
#include <cstring>
#include <set>
#include <sstream>

namespace kz {

namespace {

using ir::Op;
using ir::Scalar;
using ir::Type;

enum class Root { None, In, Out, Position, VertexIndex, InstanceIndex, FragCoord, Push, Ubo, Local, Texture };

struct PtrRef {
    Root root = Root::None;
    uint32_t index = 0;   // ubo/texture slot or local number
    uint32_t offset = 0;  // bytes
};

std::string flit(uint32_t bits) {
    float f;
    std::memcpy(&f, &bits, 4);
    char b[64];
    if (f != f) {
        std::snprintf(b, sizeof b, "__uint_as_float(0x%08Xu)", bits);
    } else if (f == __builtin_inff() || f == -__builtin_inff()) {
        std::snprintf(b, sizeof b, "__uint_as_float(0x%08Xu)", bits);
    } else {
        std::snprintf(b, sizeof b, "%af", (double)f);
    }
    return b;
}

struct StageGen {
    const ir::Module &m;
    const PipelineCodegenInput &pin;
    std::string &err;
    std::ostringstream os;
    std::map<uint32_t, PtrRef> ptrs;
    std::map<uint32_t, uint32_t> local_index;  // pointer value -> local number
    bool is_fs;

    StageGen(const ir::Module &mm, const PipelineCodegenInput &p, std::string &e, bool fs)
        : m(mm), pin(p), err(e), is_fs(fs) {}

    bool fail(const std::string &s) {
        if (err.empty()) err = s;
        return false;
    }

    std::string comp(uint32_t id, int c) {
        const ir::Value &v = m.value(id);
        const int cc = v.type.lanes > 1 ? c : 0;
        if (v.is_const) {
            const uint32_t bits = v.const_bits[cc];
            switch (v.type.scalar) {
            case Scalar::F32: return flit(bits);
            case Scalar::Bool: return bits ? "true" : "false";
            default: {
                char b[32];
                std::snprintf(b, sizeof b, "((int32_t)0x%08Xu)", bits);
                return b;
            }
            }
        }
        if (v.type.lanes > 1) return "v" + std::to_string(id) + "_" + std::to_string(cc);
        return "v" + std::to_string(id);
    }

    std::string ctype(Scalar s) {
        switch (s) {
        case Scalar::F32: return "float";
        case Scalar::I32: return "int32_t";
        case Scalar::Bool: return "bool";
        default: return "float";
        }
    }

    std::string lhs(uint32_t id, int c) {
        const ir::Value &v = m.value(id);
        return "const " + ctype(v.type.scalar) + " " +
               (v.type.lanes > 1 ? "v" + std::to_string(id) + "_" + std::to_string(c) : "v" + std::to_string(id));
    }

    // the storage expression of 32-bit word `w` of a root, as float bits
    bool word_expr(const PtrRef &p, uint32_t w, bool for_store, std::string &out) {
        const std::string ws = std::to_string(w);
        switch (p.root) {
        case Root::In: out = "kz_in[" + ws + "]"; return true;
        case Root::Out: out = "kz_out[" + ws + "]"; return true;
        case Root::Position: out = "kz_pos[" + ws + "]"; return true;
        case Root::Local: out = "kz_local" + std::to_string(p.index) + "[" + ws + "]"; return true;
        case Root::Push:
            if (for_store) return fail("store to push constants");
            if (is_fs) return fail("push constants in a fragment shader are not supported");
            out = "__uint_as_float(d.push[" + ws + "])";
            return true;
        case Root::Ubo:
            if (for_store) return fail("store to a uniform buffer");
            out = "((const float *)f.ubo[" + std::to_string(p.index) + "])[" + ws + "]";
            return true;
        case Root::VertexIndex:
            out = "__int_as_float((int32_t)kz_vertex_index)";
            return !for_store || fail("store to a built-in input");
        case Root::InstanceIndex:
            out = "__int_as_float((int32_t)d.instance)";
            return !for_store || fail("store to a built-in input");
        case Root::FragCoord:
            if (w > 3) return fail("frag coord index");
            out = "kz_frag_coord[" + ws + "]";
            return !for_store || fail("store to a built-in input");
        default: return fail("unresolved pointer");
        }
    }

    bool emit_ins(const ir::Instruction &ins) {
        auto res = [&]() { return ins.results.at(0); };
        auto lanes = [&](uint32_t id) { return (int)m.value(id).type.lanes; };
        switch (ins.op) {
        case Op::PtrOffset: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end()) return fail("pointer offset from an unknown pointer");
            PtrRef p = it->second;
            p.offset += ins.fields[0];
            ptrs[res()] = p;
            return true;
        }
        case Op::Load: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end()) return fail("load from an unknown pointer");
            const PtrRef &p = it->second;
            const ir::Value &r = m.value(res());
            if (r.type.scalar == Scalar::Texture) return fail("texture load");
            if (p.offset % 4) return fail("unaligned load");
            for (int c = 0; c < r.type.lanes; ++c) {
                std::string e;
                if (!word_expr(p, p.offset / 4 + c, false, e)) return false;
                if (r.type.scalar == Scalar::F32) os << "    " << lhs(res(), c) << " = " << e << ";\n";
                else if (r.type.scalar == Scalar::I32) os << "    " << lhs(res(), c) << " = __float_as_int(" << e << ");\n";
                else os << "    " << lhs(res(), c) << " = __float_as_int(" << e << ") != 0;\n";
            }
            return true;
        }
        case Op::Store: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end()) return fail("store to an unknown pointer");
            const PtrRef &p = it->second;
            const uint32_t v = ins.args[1];
            const ir::Value &vv = m.value(v);
            if (p.offset % 4) return fail("unaligned store");
            for (int c = 0; c < vv.type.lanes; ++c) {
                std::string e;
                if (!word_expr(p, p.offset / 4 + c, true, e)) return false;
                if (vv.type.scalar == Scalar::F32) os << "    " << e << " = " << comp(v, c) << ";\n";
                else if (vv.type.scalar == Scalar::I32) os << "    " << e << " = __int_as_float(" << comp(v, c) << ");\n";
                else os << "    " << e << " = __int_as_float(" << comp(v, c) << " ? 1 : 0);\n";
            }
            return true;
        }
        case Op::FDiv:
            if (is_fs) {  // fragment shaders: Vulkan's 2.5 ULP division (kz_fs_div)
                for (int c = 0; c < lanes(res()); ++c)
                    os << "    " << lhs(res(), c) << " = kz_fs_div(" << comp(ins.args[0], c) << ", " << comp(ins.args[1], c)
                       << ");\n";
                return true;
            }
            [[fallthrough]];
        case Op::FAdd:
        case Op::FSub:
        case Op::FMul:
        case Op::Add: {
            const char *o = ins.op == Op::FSub ? " - " : ins.op == Op::FMul ? " * " : ins.op == Op::FDiv ? " / " : " + ";
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = " << comp(ins.args[0], c) << o << comp(ins.args[1], c) << ";\n";
            return true;
        }
        case Op::FNeg:
            for (int c = 0; c < lanes(res()); ++c) os << "    " << lhs(res(), c) << " = -" << comp(ins.args[0], c) << ";\n";
            return true;
        case Op::FMax:
        case Op::FMin:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = " << (ins.op == Op::FMax ? "fmaxf(" : "fminf(") << comp(ins.args[0], c)
                   << ", " << comp(ins.args[1], c) << ");\n";
            return true;
        case Op::FClamp:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = fminf(fmaxf(" << comp(ins.args[0], c) << ", " << comp(ins.args[1], c)
                   << "), " << comp(ins.args[2], c) << ");\n";
            return true;
        case Op::Sqrt:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << (is_fs ? " = kz_fs_sqrt(" : " = __builtin_sqrtf(") << comp(ins.args[0], c)
                   << ");\n";
            return true;
        case Op::Floor:
        case Op::Ceil:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << (ins.op == Op::Floor ? " = __builtin_floorf(" : " = __builtin_ceilf(")
                   << comp(ins.args[0], c) << ");\n";
            return true;
        case Op::Dot: {
            const int n = lanes(ins.args[0]);
            std::string e = comp(ins.args[0], 0) + " * " + comp(ins.args[1], 0);
            for (int c = 1; c < n; ++c) e = "(" + e + ") + " + comp(ins.args[0], c) + " * " + comp(ins.args[1], c);
            os << "    " << lhs(res(), 0) << " = " << e << ";\n";
            return true;
        }
        case Op::Splat:
            for (int c = 0; c < lanes(res()); ++c) os << "    " << lhs(res(), c) << " = " << comp(ins.args[0], 0) << ";\n";
            return true;
        case Op::MakeVector:
            for (int c = 0; c < lanes(res()); ++c) os << "    " << lhs(res(), c) << " = " << comp(ins.args.at(c), 0) << ";\n";
            return true;
        case Op::Extract:
            os << "    " << lhs(res(), 0) << " = " << comp(ins.args[0], (int)ins.fields[0]) << ";\n";
            return true;
        case Op::Shuffle: {
            const int na = lanes(ins.args[0]);
            for (int c = 0; c < lanes(res()); ++c) {
                const int l = (int)ins.fields[c];
                os << "    " << lhs(res(), c) << " = "
                   << (l < na ? comp(ins.args[0], l) : comp(ins.args[1], l - na)) << ";\n";
            }
            return true;
        }
        case Op::Sample: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end() || it->second.root != Root::Texture) return fail("sample from an unknown texture");
            const std::string r = "smp" + std::to_string(res());
            os << "    float " << r << "[4];\n";
            os << "    kz_sample(f.tex[" << it->second.index << "], " << comp(ins.args[1], 0) << ", "
               << comp(ins.args[1], 1) << ", " << r << ");\n";
            for (int c = 0; c < 4; ++c) os << "    " << lhs(res(), c) << " = " << r << "[" << c << "];\n";
            return true;
        }
        case Op::ConvSToF:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = (float)" << comp(ins.args[0], c) << ";\n";
            return true;
        case Op::ConvFToS:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = (int32_t)" << comp(ins.args[0], c) << ";\n";
            return true;
        case Op::Select:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = " << comp(ins.args[0], c) << " ? " << comp(ins.args[1], c) << " : "
                   << comp(ins.args[2], c) << ";\n";
            return true;
        case Op::FOrdLt:
        case Op::FOrdGt:
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = " << comp(ins.args[0], c) << (ins.op == Op::FOrdLt ? " < " : " > ")
                   << comp(ins.args[1], c) << ";\n";
            return true;
        case Op::Kill: return fail("fragment discard (OpKill) is not supported yet");
        case Op::BreakBlock: return true;
        }
        return fail("unknown IR instruction");
    }

    bool generate(const std::string &fn_name) {
        // interface roots
        ptrs[m.user_inputs.pointer] = {Root::In, 0, 0};
        ptrs[m.user_outputs.pointer] = {Root::Out, 0, 0};
        for (const auto &b : m.built_in_outputs) {
            if (b.kind != ir::BuiltIn::VertexPosition) return fail("unsupported built-in output");
            ptrs[b.pointer] = {Root::Position, 0, 0};
        }
        for (const auto &b : m.built_in_inputs) {
            Root r = b.kind == ir::BuiltIn::VertexIndex     ? Root::VertexIndex
                     : b.kind == ir::BuiltIn::InstanceIndex ? Root::InstanceIndex
                                                            : Root::FragCoord;
            ptrs[b.pointer] = {r, 0, 0};
        }
        for (const auto &r : m.resources) {
            if (r.kind == ir::ResourceKind::PushConstants) {
                ptrs[r.pointer] = {Root::Push, 0, 0};
            } else if (r.kind == ir::ResourceKind::UniformBuffer) {
                auto it = pin.ubo_slot.find({r.set, r.binding});
                if (it == pin.ubo_slot.end()) return fail("uniform buffer without a descriptor binding in the layout");
                ptrs[r.pointer] = {Root::Ubo, it->second, 0};
            } else {
                auto it = pin.tex_slot.find({r.set, r.binding});
                if (it == pin.tex_slot.end()) return fail("sampler without a descriptor binding in the layout");
                ptrs[r.pointer] = {Root::Texture, it->second, 0};
            }
        }
        uint32_t nlocal = 0;
        std::ostringstream decl;
        auto add_locals = [&](const std::vector<ir::Variable> &vars) {
            for (const auto &v : vars) {
                const uint32_t words = (std::max(v.size, 4u) + 3) / 4;
                decl << "    float kz_local" << nlocal << "[" << words << "] = {};\n";
                ptrs[v.pointer] = {Root::Local, nlocal, 0};
                ++nlocal;
            }
        };
        add_locals(m.invocation_globals);
        add_locals(m.main.locals);
        if (is_fs)
            os << "__device__ __forceinline__ void " << fn_name
               << "(const KzFrame &f, const float *kz_in, float *kz_out, const float *kz_frag_coord) {\n";
        else
            os << "__device__ __forceinline__ void " << fn_name
               << "(const KzFrame &f, const KzDraw &d, uint32_t kz_vertex_index, const float *kz_in, float *kz_out, "
                  "float *kz_pos) {\n";
        os << "    (void)f; (void)kz_in; (void)kz_out;\n";
        os << decl.str();
        for (const auto &ins : m.main.body.body)
            if (!emit_ins(ins)) return false;
        os << "}\n\n";
        return true;
    }
};

uint32_t block_words(const ir::InterfaceBlock &b) {
    uint32_t w = (b.size + 3) / 4;
    return w < 4 ? 4 : w;
}

bool format_components(uint32_t fmt, int &n, int &kind) {
    // kind: 0 = f32, 1 = u32 bits, 2 = unorm8
    switch (fmt) {
    case 100: n = 1; kind = 0; return true;  // R32_SFLOAT
    case 103: n = 2; kind = 0; return true;  // R32G32_SFLOAT
    case 106: n = 3; kind = 0; return true;  // R32G32B32_SFLOAT
    case 109: n = 4; kind = 0; return true;  // R32G32B32A32_SFLOAT
    case 98: n = 1; kind = 1; return true;   // R32_UINT
    case 37: n = 4; kind = 2; return true;   // R8G8B8A8_UNORM
    }
    return false;
}

}  // namespace

bool generate_pipeline_source(const PipelineCodegenInput &in, PipelineCodegenOutput &out, std::string &err) {
    if (!in.vs) {
        err = "pipeline without a vertex shader";
        return false;
    }
    std::ostringstream os;
    StageGen vg(*in.vs, in, err, false);
    if (!vg.generate("kz_vs_main")) return false;
    os << vg.os.str();
    if (in.fs) {
        StageGen fg(*in.fs, in, err, true);
        if (!fg.generate("kz_fs_main")) return false;
        os << fg.os.str();
        for (const auto &b : in.fs->built_in_inputs)
            if (b.kind == ir::BuiltIn::FragCoord) out.fs_reads_frag_coord = true;
    }

    // varying linkage: every float slot the fragment shader reads, in
    // offset order, packed densely
    std::vector<uint32_t> slots;
    if (in.fs) {
        std::set<uint32_t> s;
        for (const auto &mem : in.fs->user_inputs.members) {
            if (mem.type.scalar != Scalar::F32) {
                err = "non-float fragment inputs are not supported";
                return false;
            }
            for (int c = 0; c < mem.type.lanes; ++c) s.insert(mem.offset / 4 + c);
        }
        slots.assign(s.begin(), s.end());
    }
    std::set<uint32_t> vs_written;
    for (const auto &mem : in.vs->user_outputs.members)
        for (int c = 0; c < mem.type.lanes; ++c) vs_written.insert(mem.offset / 4 + c);
    out.num_varyings = (uint32_t)slots.size();
    const uint32_t vin_words = block_words(in.vs->user_inputs), vout_words = block_words(in.vs->user_outputs);

    os << "__device__ __forceinline__ void kz_vertex(const KzFrame &f, const KzDraw &d, uint32_t vi, float pos[4], "
          "float *vary) {\n";
    os << "    float kz_in[" << vin_words << "];\n";
    // vertex fetch: defaults (0, 0, 0, 1) per location
    for (uint32_t w = 0; w < vin_words; ++w) os << "    kz_in[" << w << "] = " << ((w % 4) == 3 ? "1.0f" : "0.0f") << ";\n";
    for (const auto &mem : in.vs->user_inputs.members) {
        const uint32_t loc = mem.offset / 16;
        const VertexAttributeDesc *a = nullptr;
        for (const auto &att : in.attributes)
            if (att.location == loc) a = &att;
        if (!a) continue;
        const VertexBindingDesc *b = nullptr;
        for (const auto &bb : in.bindings)
            if (bb.binding == a->binding) b = &bb;
        if (!b) {
            err = "vertex attribute refers to a missing binding";
            return false;
        }
        int n, kind;
        if (!format_components(a->format, n, kind)) {
            err = "unsupported vertex attribute format " + std::to_string(a->format);
            return false;
        }
        os << "    {\n        const uint8_t *src = d.vb[" << a->binding << "] + (size_t)("
           << (b->input_rate == 1 ? "d.instance" : "vi") << ") * " << b->stride << "u + " << a->offset << "u;\n";
        for (int c = 0; c < n; ++c) {
            const uint32_t w = loc * 4 + c;
            if (kind == 0) os << "        kz_in[" << w << "] = ((const float *)src)[" << c << "];\n";
            else if (kind == 1) os << "        kz_in[" << w << "] = __uint_as_float(((const uint32_t *)src)[" << c << "]);\n";
            else os << "        kz_in[" << w << "] = (float)src[" << c << "] * (1.0f / 255.0f);\n";
        }
        os << "    }\n";
    }
    os << "    float kz_out[" << vout_words << "];\n";
    os << "    for (int i = 0; i < " << vout_words << "; ++i) kz_out[i] = 0.0f;\n";
    os << "    pos[0] = pos[1] = pos[2] = 0.0f; pos[3] = 1.0f;\n";
    os << "    kz_vs_main(f, d, vi, kz_in, kz_out, pos);\n";
    for (size_t j = 0; j < slots.size(); ++j) {
        if (vs_written.count(slots[j])) os << "    vary[" << j << "] = kz_out[" << slots[j] << "];\n";
        else os << "    vary[" << j << "] = 0.0f;\n";
    }
    os << "    (void)vary;\n}\n\n";

    os << "__device__ __forceinline__ void kz_fragment(const KzFrame &f, const float *vary, const float *frag_coord, "
          "float o[4]) {\n";
    if (in.fs) {
        const uint32_t fin_words = block_words(in.fs->user_inputs), fout_words = block_words(in.fs->user_outputs);
        os << "    float kz_in[" << fin_words << "];\n";
        os << "    for (int i = 0; i < " << fin_words << "; ++i) kz_in[i] = 0.0f;\n";
        for (size_t j = 0; j < slots.size(); ++j) os << "    kz_in[" << slots[j] << "] = vary[" << j << "];\n";
        os << "    float kz_out[" << fout_words << "];\n";
        os << "    for (int i = 0; i < " << fout_words << "; ++i) kz_out[i] = 0.0f;\n";
        os << "    kz_fs_main(f, kz_in, kz_out, frag_coord);\n";
        os << "    o[0] = kz_out[0]; o[1] = kz_out[1]; o[2] = kz_out[2]; o[3] = kz_out[3];\n";
    } else {
        os << "    (void)f; (void)vary; (void)frag_coord; o[0] = o[1] = o[2] = o[3] = 0.0f;\n";
    }
    os << "}\n";
    out.source = os.str();
    return true;
}

}  // namespace kz
