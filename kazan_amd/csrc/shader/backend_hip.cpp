// backend_hip.cpp — shader IR -> HIP device functions.
//
// Every IR value becomes one C++ variable per component, declared at the top
// of the stage function; every pointer is resolved at compile time to a
// storage root plus a constant byte offset, plus a run-time word index for
// dynamically indexed access chains. Structured control flow lowers to
// labels and gotos: a Block's breaks jump to a label right after it, a
// Loop's continues to a label right before it, a Switch to the chosen
// block's label. Returning from the stage function is a break from block 0;
// the fragment stage returns false when it discards (OpKill), the planned
// ABI's bool result (docs/Example Translation from SPIR-V to LLVM IR.md
// :190-198). Float operations are emitted one rounding at a time in IR
// order and the kernel is compiled with -ffp-contract=off and IEEE division /
// square root, so the generated code computes exactly what the IR (and the
// CPU oracle's restatement) says.
#include "backend_hip.h"

#include "../engine/device_common.h"

#include <cstdio>
#include <cstring>
#include <functional>
#include <set>
#include <sstream>

namespace kz {

namespace {

using ir::Op;
using ir::Scalar;
using ir::Type;

enum class Root {
    None, In, Out, Position, VertexIndex, InstanceIndex, FragCoord, Push, Ubo, Local, Texture,
    Storage, Workgroup, ComputeId,  // compute: storage buffers, workgroup memory, the invocation's ids
};

struct PtrRef {
    Root root = Root::None;
    uint32_t index = 0;   // ubo/texture slot or local number
    uint32_t offset = 0;  // bytes
    std::string dyn;      // run-time word index added to offset / 4 ("" = none)
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

uint32_t block_words_of(const ir::InterfaceBlock &b) {
    uint32_t w = (b.size + 3) / 4;
    return w < 4 ? 4 : w;
}

struct StageGen {
    const ir::Module &m;
    const PipelineCodegenInput &pin;
    std::string &err;
    std::ostringstream os;
    std::map<uint32_t, PtrRef> ptrs;
    std::map<uint32_t, uint32_t> local_index;  // pointer value -> local number
    std::vector<uint32_t> local_words;         // words of each local
    std::vector<uint32_t> wg_words;            // words of each workgroup variable
    uint32_t in_words = 4, out_words = 4;
    bool is_fs;
    bool is_cs = false;

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

    // values are declared at the top of the function (structured code
    // jumps between scopes) and assigned where the IR computes them
    std::string lhs(uint32_t id, int c) {
        const ir::Value &v = m.value(id);
        return v.type.lanes > 1 ? "v" + std::to_string(id) + "_" + std::to_string(c) : "v" + std::to_string(id);
    }

    // A word index: constant, or constant + run-time part clamped into an
    // array of n words (an out-of-range index of a shader's own arrays
    // stays inside them instead of touching other memory).
    static std::string index_expr(uint32_t w, const std::string &dyn, uint32_t n) {
        if (dyn.empty()) return std::to_string(w);
        return "kz_word_index(" + std::to_string(w) + dyn + ", " + std::to_string(n) + ")";
    }

    // the storage expression of 32-bit word `w` of a root, as float bits
    bool word_expr(const PtrRef &p, uint32_t w, bool for_store, std::string &out) {
        const std::string ws = std::to_string(w);
        switch (p.root) {
        case Root::In: out = "kz_in[" + index_expr(w, p.dyn, in_words) + "]"; return true;
        case Root::Out: out = "kz_out[" + index_expr(w, p.dyn, out_words) + "]"; return true;
        case Root::Position: out = "kz_pos[" + index_expr(w, p.dyn, 4) + "]"; return true;
        case Root::Local:
            out = "kz_local" + std::to_string(p.index) + "[" + index_expr(w, p.dyn, local_words.at(p.index)) + "]";
            return true;
        case Root::Push:
            if (for_store) return fail("store to push constants");
            if (is_fs) return fail("push constants in a fragment shader are not supported");
            out = std::string("__uint_as_float(") + (is_cs ? "f" : "d") + ".push[" +
                  index_expr(w, p.dyn, KZ_PUSH_WORDS) + "])";
            return true;
        case Root::Storage:  // (a storage buffer's size is the application's: indices are not clamped)
            out = "((float *)f.ssbo[" + std::to_string(p.index) + "])[" + ws + p.dyn + "]";
            return true;
        case Root::Workgroup:
            out = "kz_wg" + std::to_string(p.index) + "[" + index_expr(w, p.dyn, wg_words.at(p.index)) + "]";
            return true;
        case Root::ComputeId:  // index: which id; the word selects x, y or z
            if (w > 2) return fail("compute built-in index");
            out = "__int_as_float((int32_t)kz_ids[" + std::to_string(p.index * 3 + w) + "])";
            return !for_store || fail("store to a built-in input");
        case Root::Ubo:
            if (for_store) return fail("store to a uniform buffer");
            out = "((const float *)f.ubo[" + std::to_string(p.index) + "])[" + ws + p.dyn + "]";
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

    void lanes_expr(uint32_t res, const std::function<std::string(int)> &e) {
        for (int c = 0; c < m.value(res).type.lanes; ++c) os << "    " << lhs(res, c) << " = " << e(c) << ";\n";
    }

    bool emit_block(uint32_t b) {
        for (const auto &ins : m.main.blocks.at(b).body)
            if (!emit_ins(ins)) return false;
        return true;
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
        case Op::PtrIndex: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end()) return fail("pointer index from an unknown pointer");
            if (ins.fields[0] % 4) return fail("pointer index stride is not a whole number of words");
            PtrRef p = it->second;
            if (p.root == Root::Texture) return fail("indexed texture");
            p.dyn += " + " + comp(ins.args[1], 0) + " * " + std::to_string(ins.fields[0] / 4);
            ptrs[res()] = p;
            return true;
        }
        case Op::Block:
            os << "    {\n";
            if (!emit_block(ins.fields[0])) return false;
            os << "    }\n    kzb" << ins.fields[0] << ":;\n";
            return true;
        case Op::Loop:
            os << "    kzl" << ins.fields[0] << ":;\n    {\n";
            if (!emit_block(ins.fields[0])) return false;
            os << "    }\n    goto kzl" << ins.fields[0] << ";\n";
            return true;
        case Op::BreakBlock:
            if (ins.fields.at(0) == 0) os << "    return true;\n";
            else os << "    goto kzb" << ins.fields[0] << ";\n";
            return true;
        case Op::ContinueLoop: os << "    goto kzl" << ins.fields.at(0) << ";\n"; return true;
        case Op::Switch: {
            const Scalar st = m.value(ins.args[0]).type.scalar;
            os << "    switch ((int32_t)" << comp(ins.args[0], 0) << ") {\n";
            for (size_t i = 1; i + 1 < ins.fields.size(); i += 2) {
                const uint32_t lit = st == Scalar::Bool ? (ins.fields[i] != 0) : ins.fields[i];
                os << "    case (int32_t)0x" << std::hex << lit << std::dec << "u: goto kzb" << ins.fields[i + 1] << ";\n";
            }
            os << "    default: goto kzb" << ins.fields.at(0) << ";\n    }\n";
            return true;
        }
        case Op::Kill:
            if (!is_fs) return fail("discard outside a fragment shader");
            os << "    return false;\n";
            return true;
        case Op::Unreachable: os << "    return true;\n"; return true;
        case Op::BufferSize: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end() || it->second.root != Root::Storage) return fail("buffer size of a non-storage-buffer pointer");
            os << "    " << lhs(res(), 0) << " = (int32_t)f.ssbo_bytes[" << it->second.index << "];\n";
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
        case Op::FMul: {
            const char *o = ins.op == Op::FSub ? " - " : ins.op == Op::FMul ? " * " : ins.op == Op::FDiv ? " / " : " + ";
            for (int c = 0; c < lanes(res()); ++c)
                os << "    " << lhs(res(), c) << " = " << comp(ins.args[0], c) << o << comp(ins.args[1], c) << ";\n";
            return true;
        }
        // integer arithmetic wraps (two's complement), as SPIR-V defines it
        case Op::Add:
        case Op::ISub:
        case Op::IMul: {
            const char *o = ins.op == Op::Add ? " + " : ins.op == Op::ISub ? " - " : " * ";
            const auto a = ins.args[0], b = ins.args[1];
            lanes_expr(res(), [&](int c) { return "(int32_t)((uint32_t)" + comp(a, c) + o + "(uint32_t)" + comp(b, c) + ")"; });
            return true;
        }
        case Op::SDiv:
        case Op::UDiv:
        case Op::SRem:
        case Op::SMod:
        case Op::UMod:
        case Op::Shl:
        case Op::ShrL:
        case Op::ShrA:
        case Op::SMin:
        case Op::SMax:
        case Op::UMin:
        case Op::UMax:
        case Op::FMod:
        case Op::FRem:
        case Op::Pow: {
            static const std::map<Op, const char *> fn = {
                {Op::SDiv, "kz_sdiv"}, {Op::UDiv, "kz_udiv"}, {Op::SRem, "kz_srem"}, {Op::SMod, "kz_smod"},
                {Op::UMod, "kz_umod"}, {Op::Shl, "kz_shl"}, {Op::ShrL, "kz_lshr"}, {Op::ShrA, "kz_ashr"},
                {Op::SMin, "kz_smin"}, {Op::SMax, "kz_smax"}, {Op::UMin, "kz_umin"}, {Op::UMax, "kz_umax"},
                {Op::FMod, "kz_fmod_glsl"}, {Op::FRem, "__builtin_fmodf"}, {Op::Pow, "__builtin_powf"}};
            const auto a = ins.args[0], b = ins.args[1];
            const std::string f = fn.at(ins.op);
            lanes_expr(res(), [&](int c) { return f + "(" + comp(a, c) + ", " + comp(b, c) + ")"; });
            return true;
        }
        case Op::SClamp:
        case Op::UClamp: {
            const std::string f = ins.op == Op::SClamp ? "kz_sclamp" : "kz_uclamp";
            lanes_expr(res(), [&](int c) {
                return f + "(" + comp(ins.args[0], c) + ", " + comp(ins.args[1], c) + ", " + comp(ins.args[2], c) + ")";
            });
            return true;
        }
        case Op::And:
        case Op::Or:
        case Op::Xor: {
            const bool bl = m.value(ins.args[0]).type.scalar == Scalar::Bool;
            const char *o = ins.op == Op::And ? (bl ? " && " : " & ") : ins.op == Op::Or ? (bl ? " || " : " | ")
                                                                                          : (bl ? " != " : " ^ ");
            lanes_expr(res(), [&](int c) { return "(" + comp(ins.args[0], c) + o + comp(ins.args[1], c) + ")"; });
            return true;
        }
        case Op::Not: {
            const bool bl = m.value(ins.args[0]).type.scalar == Scalar::Bool;
            lanes_expr(res(), [&](int c) { return std::string(bl ? "!" : "~") + comp(ins.args[0], c); });
            return true;
        }
        case Op::INeg:
            lanes_expr(res(), [&](int c) { return "(int32_t)(0u - (uint32_t)" + comp(ins.args[0], c) + ")"; });
            return true;
        case Op::SAbs:
            lanes_expr(res(), [&](int c) { return "kz_sabs(" + comp(ins.args[0], c) + ")"; });
            return true;
        case Op::FSign:
        case Op::Trunc:
        case Op::RoundEven:
        case Op::Exp2:
        case Op::Log2:
        case Op::Exp:
        case Op::Log:
        case Op::Sin:
        case Op::Cos: {
            static const std::map<Op, const char *> fn = {
                {Op::FSign, "kz_fsign"}, {Op::Trunc, "__builtin_truncf"}, {Op::RoundEven, "__builtin_rintf"},
                {Op::Exp2, "__builtin_exp2f"}, {Op::Log2, "__builtin_log2f"}, {Op::Exp, "__builtin_expf"},
                {Op::Log, "__builtin_logf"}, {Op::Sin, "__builtin_sinf"}, {Op::Cos, "__builtin_cosf"}};
            const std::string f = fn.at(ins.op);
            lanes_expr(res(), [&](int c) { return f + "(" + comp(ins.args[0], c) + ")"; });
            return true;
        }
        case Op::Atan2:
        case Op::BitCount:
        case Op::BitReverse:
        case Op::FindILsb:
        case Op::FindUMsb:
        case Op::FindSMsb:
        case Op::BitFieldInsert:
        case Op::BitFieldSExtract:
        case Op::BitFieldUExtract:
        case Op::F32ToF16:
        case Op::F16ToF32:
        case Op::Ldexp:
        case Op::FrexpMant:
        case Op::FrexpExp: {
            static const std::map<Op, const char *> fn = {
                {Op::Ldexp, "kz_ldexp"}, {Op::FrexpMant, "kz_frexp_mant"}, {Op::FrexpExp, "kz_frexp_exp"},
                {Op::F32ToF16, "kz_f32_to_f16_bits"}, {Op::F16ToF32, "kz_f16_bits_to_f32"},
                {Op::Atan2, "atan2f"}, {Op::BitCount, "kz_bit_count"}, {Op::BitReverse, "kz_bit_reverse"},
                {Op::FindILsb, "kz_find_lsb"}, {Op::FindUMsb, "kz_find_umsb"}, {Op::FindSMsb, "kz_find_smsb"},
                {Op::BitFieldInsert, "kz_bitfield_insert"}, {Op::BitFieldSExtract, "kz_bitfield_sextract"},
                {Op::BitFieldUExtract, "kz_bitfield_uextract"}};
            const std::string f = fn.at(ins.op);
            lanes_expr(res(), [&](int c) {
                std::string e = f + "(";
                for (size_t k = 0; k < ins.args.size(); ++k) e += (k ? ", " : "") + comp(ins.args[k], c);
                return e + ")";
            });
            return true;
        }
        case Op::FCmp: {
            static const char *const form[] = {
                "(%a == %b)", "(%a < %b || %a > %b)", "(%a < %b)", "(%a <= %b)", "(%a > %b)", "(%a >= %b)",
                "!(%a < %b || %a > %b)", "(%a != %b)", "!(%a >= %b)", "!(%a > %b)", "!(%a <= %b)", "!(%a < %b)",
                "(%a == %a && %b == %b)", "(%a != %a || %b != %b)"};
            if (ins.fields.at(0) >= sizeof form / sizeof form[0]) return fail("float comparison predicate");
            lanes_expr(res(), [&](int c) {
                std::string e = form[ins.fields[0]];
                for (size_t k; (k = e.find("%a")) != std::string::npos;) e.replace(k, 2, comp(ins.args[0], c));
                for (size_t k; (k = e.find("%b")) != std::string::npos;) e.replace(k, 2, comp(ins.args[1], c));
                return e;
            });
            return true;
        }
        case Op::ICmp: {
            static const char *const op[] = {" == ", " != ", " < ", " <= ", " > ", " >= ", " < ", " <= ", " > ", " >= "};
            const uint32_t pr = ins.fields.at(0);
            if (pr >= sizeof op / sizeof op[0]) return fail("integer comparison predicate");
            const bool u = pr >= ir::IULt;
            lanes_expr(res(), [&](int c) {
                const std::string a = comp(ins.args[0], c), b = comp(ins.args[1], c);
                return u ? "((uint32_t)" + a + op[pr] + "(uint32_t)" + b + ")" : "(" + a + op[pr] + b + ")";
            });
            return true;
        }
        case Op::ConvUToF:
            lanes_expr(res(), [&](int c) { return "(float)(uint32_t)" + comp(ins.args[0], c); });
            return true;
        case Op::ConvFToU:
            lanes_expr(res(), [&](int c) { return "(int32_t)kz_f2u(" + comp(ins.args[0], c) + ")"; });
            return true;
        case Op::Bitcast: {
            const Scalar from = m.value(ins.args[0]).type.scalar, to = m.value(res()).type.scalar;
            lanes_expr(res(), [&](int c) {
                const std::string a = comp(ins.args[0], c);
                if (from == to) return a;
                return to == Scalar::F32 ? "__int_as_float(" + a + ")" : "__float_as_int(" + a + ")";
            });
            return true;
        }
        case Op::Any:
        case Op::All: {
            std::string e;
            for (int c = 0; c < lanes(ins.args[0]); ++c)
                e += (c ? (ins.op == Op::Any ? " || " : " && ") : "") + comp(ins.args[0], c);
            os << "    " << lhs(res(), 0) << " = (" << e << ");\n";
            return true;
        }
        case Op::IsNan:
            lanes_expr(res(), [&](int c) { return "(" + comp(ins.args[0], c) + " != " + comp(ins.args[0], c) + ")"; });
            return true;
        case Op::IsInf:
            lanes_expr(res(), [&](int c) { return "(__builtin_fabsf(" + comp(ins.args[0], c) + ") == __builtin_inff())"; });
            return true;
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
        case Op::Barrier:
            if (!is_cs) return fail("barrier outside a compute shader");
            os << "    __syncthreads();\n";
            return true;
        case Op::Atomic: {
            auto it = ptrs.find(ins.args[0]);
            if (it == ptrs.end()) return fail("atomic on an unknown pointer");
            const PtrRef &p = it->second;
            if (p.root != Root::Storage && p.root != Root::Workgroup && p.root != Root::Local)
                return fail("atomic on memory other than a storage buffer, workgroup or function variable");
            if (p.offset % 4) return fail("unaligned atomic");
            std::string e;
            if (!word_expr(p, p.offset / 4, true, e)) return false;
            const std::string v = comp(ins.args[1], 0);
            static const char *const fn[] = {"atomicAdd", "atomicSub", "atomicMin", "atomicMax", "atomicMin",
                                             "atomicMax", "atomicAnd", "atomicOr",  "atomicXor", "atomicExch"};
            const uint32_t k = ins.fields.at(0);
            if (k > ir::AExchange) return fail("unknown atomic");
            const bool u = k == ir::AUMin || k == ir::AUMax;
            os << "    " << lhs(res(), 0) << " = (int32_t)" << fn[k] << "((" << (u ? "unsigned" : "int") << " *)&("
               << e << "), (" << (u ? "unsigned" : "int") << ")" << v << ");\n";
            return true;
        }
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
            switch (b.kind) {
            case ir::BuiltIn::VertexIndex: ptrs[b.pointer] = {Root::VertexIndex, 0, 0}; break;
            case ir::BuiltIn::InstanceIndex: ptrs[b.pointer] = {Root::InstanceIndex, 0, 0}; break;
            case ir::BuiltIn::FragCoord: ptrs[b.pointer] = {Root::FragCoord, 0, 0}; break;
            // kz_ids: global id, local id, workgroup id, workgroup count (x, y, z each), local index
            case ir::BuiltIn::GlobalInvocationId: ptrs[b.pointer] = {Root::ComputeId, 0, 0}; break;
            case ir::BuiltIn::LocalInvocationId: ptrs[b.pointer] = {Root::ComputeId, 1, 0}; break;
            case ir::BuiltIn::WorkgroupId: ptrs[b.pointer] = {Root::ComputeId, 2, 0}; break;
            case ir::BuiltIn::NumWorkgroups: ptrs[b.pointer] = {Root::ComputeId, 3, 0}; break;
            case ir::BuiltIn::LocalInvocationIndex: ptrs[b.pointer] = {Root::ComputeId, 4, 0}; break;
            default: return fail("unsupported built-in input");
            }
        }
        for (const auto &r : m.resources) {
            if (r.kind == ir::ResourceKind::PushConstants) {
                ptrs[r.pointer] = {Root::Push, 0, 0};
            } else if (r.kind == ir::ResourceKind::UniformBuffer) {
                auto it = pin.ubo_slot.find({r.set, r.binding});
                if (it == pin.ubo_slot.end()) return fail("uniform buffer without a descriptor binding in the layout");
                ptrs[r.pointer] = {Root::Ubo, it->second, 0};
            } else if (r.kind == ir::ResourceKind::StorageBuffer) {
                if (!is_cs) return fail("storage buffers are supported in compute shaders only");
                auto it = pin.ssbo_slot.find({r.set, r.binding});
                if (it == pin.ssbo_slot.end()) return fail("storage buffer without a descriptor binding in the layout");
                ptrs[r.pointer] = {Root::Storage, it->second, 0};
            } else {
                if (is_cs) return fail("combined image samplers in compute shaders are not supported");
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
                ptrs[v.pointer] = {Root::Local, nlocal, 0, ""};
                local_words.push_back(words);
                ++nlocal;
            }
        };
        add_locals(m.invocation_globals);
        add_locals(m.main.locals);
        for (const auto &v : m.workgroup_variables) {  // (declared at file scope by the caller)
            const uint32_t words = (std::max(v.size, 4u) + 3) / 4;
            ptrs[v.pointer] = {Root::Workgroup, (uint32_t)wg_words.size(), 0, ""};
            wg_words.push_back(words);
        }
        // every value the function computes, declared up front
        for (const auto &blk : m.main.blocks)
            for (const auto &ins : blk.body)
                for (uint32_t r : ins.results) {
                    const ir::Value &v = m.value(r);
                    if (v.type.scalar == Scalar::DataPtr || v.type.scalar == Scalar::Texture) continue;
                    decl << "    " << ctype(v.type.scalar);
                    for (int c = 0; c < v.type.lanes; ++c) decl << (c ? ", " : " ") << lhs(r, c);
                    decl << ";\n";
                }
        in_words = block_words_of(m.user_inputs);
        out_words = block_words_of(m.user_outputs);
        if (is_cs)
            os << "__device__ __forceinline__ bool " << fn_name
               << "(const KzComputeFrame &f, const uint32_t *kz_ids, const float *kz_in, float *kz_out) {\n";
        else if (is_fs)
            os << "__device__ __forceinline__ bool " << fn_name
               << "(const KzFrame &f, const float *kz_in, float *kz_out, const float *kz_frag_coord) {\n";
        else
            os << "__device__ __forceinline__ bool " << fn_name
               << "(const KzFrame &f, const KzDraw &d, uint32_t kz_vertex_index, const float *kz_in, float *kz_out, "
                  "float *kz_pos) {\n";
        os << "    (void)f; (void)kz_in; (void)kz_out;\n";
        os << decl.str();
        if (!emit_block(0)) return false;
        os << "    return true;\n}\n\n";
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

    // false: the fragment shader discarded the fragment
    os << "__device__ __forceinline__ bool kz_fragment(const KzFrame &f, const float *vary, const float *frag_coord, "
          "float o[4]) {\n";
    if (in.fs) {
        const uint32_t fin_words = block_words(in.fs->user_inputs), fout_words = block_words(in.fs->user_outputs);
        os << "    float kz_in[" << fin_words << "];\n";
        os << "    for (int i = 0; i < " << fin_words << "; ++i) kz_in[i] = 0.0f;\n";
        for (size_t j = 0; j < slots.size(); ++j) os << "    kz_in[" << slots[j] << "] = vary[" << j << "];\n";
        os << "    float kz_out[" << fout_words << "];\n";
        os << "    for (int i = 0; i < " << fout_words << "; ++i) kz_out[i] = 0.0f;\n";
        os << "    const bool keep = kz_fs_main(f, kz_in, kz_out, frag_coord);\n";
        os << "    o[0] = kz_out[0]; o[1] = kz_out[1]; o[2] = kz_out[2]; o[3] = kz_out[3];\n";
        os << "    return keep;\n";
    } else {
        os << "    (void)f; (void)vary; (void)frag_coord; o[0] = o[1] = o[2] = o[3] = 0.0f;\n";
        os << "    return true;\n";
    }
    os << "}\n";
    out.source = os.str();
    return true;
}

bool generate_compute_source(const ir::Module &cs, const PipelineCodegenInput &in, std::string &out,
                             std::string &err) {
    const uint32_t lx = cs.local_size[0], ly = cs.local_size[1], lz = cs.local_size[2];
    const uint64_t n = (uint64_t)lx * ly * lz;
    if (cs.stage != ir::Stage::Compute || n == 0 || n > 1024) {
        err = "compute shader without a workgroup size of 1 .. 1024 invocations";
        return false;
    }
    std::ostringstream os;
    // workgroup memory: one array per variable, shared by the workgroup
    for (size_t i = 0; i < cs.workgroup_variables.size(); ++i) {
        const uint32_t words = (std::max(cs.workgroup_variables[i].size, 4u) + 3) / 4;
        os << "__shared__ float kz_wg" << i << "[" << words << "];\n";
    }
    StageGen g(cs, in, err, false);
    g.is_cs = true;
    if (!g.generate("kz_cs_main")) return false;
    os << g.os.str();
    // the kernel: one workgroup per dispatched group, one thread per invocation
    os << "extern \"C\" __global__ __launch_bounds__(" << n << ") void kz_compute(KzComputeFrame f) {\n";
    os << "    const uint32_t L[3] = {" << lx << "u, " << ly << "u, " << lz << "u};\n";
    os << "    const uint32_t lid[3] = {threadIdx.x, threadIdx.y, threadIdx.z};\n";
    os << "    const uint32_t wid[3] = {blockIdx.x + f.base[0], blockIdx.y + f.base[1], blockIdx.z + f.base[2]};\n";
    os << "    uint32_t kz_ids[13];\n";
    os << "    for (int k = 0; k < 3; ++k) {\n";
    os << "        kz_ids[k] = wid[k] * L[k] + lid[k];\n";
    os << "        kz_ids[3 + k] = lid[k];\n";
    os << "        kz_ids[6 + k] = wid[k];\n";
    os << "    }\n";
    os << "    kz_ids[9] = gridDim.x; kz_ids[10] = gridDim.y; kz_ids[11] = gridDim.z;\n";
    os << "    kz_ids[12] = lid[0] + L[0] * (lid[1] + L[1] * lid[2]);\n";
    os << "    float kz_in[4] = {}, kz_out[4] = {};\n";
    os << "    kz_cs_main(f, kz_ids, kz_in, kz_out);\n";
    os << "}\n";
    out = os.str();
    return true;
}

}  // namespace kz
