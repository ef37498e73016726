// ir.cpp — construction helpers and the text form of the shader IR.
//
// The layout of the text (indentation, section names, `size: fixed 0x..;`,
// `align: 0x..;`, `-> name: data_ptr;`, hints block, `entry_point: main;`)
// is the one fixed by the reference's golden files
// shader-compiler-ir/test_data/test_assert_ir_matches_file.kazan-ir and
// shader-compiler-translate-spirv-to-ir/test_data/trivial_test.kazan-ir.
#include "ir.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace kz {
namespace ir {

namespace {

const char *scalar_text(Scalar s) {
    switch (s) {
    case Scalar::Void: return "void";
    case Scalar::Bool: return "bool";
    case Scalar::I32: return "i32";
    case Scalar::F32: return "f32";
    case Scalar::DataPtr: return "data_ptr";
    case Scalar::Texture: return "texture";
    }
    return "?";
}

std::string hex(uint32_t v) {
    char b[16];
    std::snprintf(b, sizeof b, "0x%X", v);
    return b;
}

const char *op_text(Op op) {  // (op_name below)
    switch (op) {
    case Op::Add: return "add";
    case Op::FAdd: return "fadd";
    case Op::FSub: return "fsub";
    case Op::FMul: return "fmul";
    case Op::FDiv: return "fdiv";
    case Op::FNeg: return "fneg";
    case Op::FMax: return "fmax";
    case Op::FMin: return "fmin";
    case Op::FClamp: return "fclamp";
    case Op::Sqrt: return "sqrt";
    case Op::Floor: return "floor";
    case Op::Ceil: return "ceil";
    case Op::Dot: return "dot";
    case Op::Splat: return "splat";
    case Op::MakeVector: return "make_vector";
    case Op::Extract: return "extract";
    case Op::Shuffle: return "shuffle";
    case Op::Load: return "load";
    case Op::Store: return "store";
    case Op::PtrOffset: return "ptr_offset";
    case Op::Sample: return "sample";
    case Op::ConvSToF: return "convert_s_to_f";
    case Op::ConvFToS: return "convert_f_to_s";
    case Op::Select: return "select";
    case Op::FOrdLt: return "ford_lt";
    case Op::FOrdGt: return "ford_gt";
    case Op::Kill: return "kill";
    case Op::BreakBlock: return "break";
    case Op::ISub: return "sub";
    case Op::IMul: return "mul";
    case Op::SDiv: return "sdiv";
    case Op::UDiv: return "udiv";
    case Op::SRem: return "srem";
    case Op::SMod: return "smod";
    case Op::UMod: return "umod";
    case Op::Shl: return "shl";
    case Op::ShrL: return "lshr";
    case Op::ShrA: return "ashr";
    case Op::And: return "and";
    case Op::Or: return "or";
    case Op::Xor: return "xor";
    case Op::Not: return "not";
    case Op::INeg: return "neg";
    case Op::SMin: return "smin";
    case Op::SMax: return "smax";
    case Op::UMin: return "umin";
    case Op::UMax: return "umax";
    case Op::SAbs: return "sabs";
    case Op::SClamp: return "sclamp";
    case Op::UClamp: return "uclamp";
    case Op::FMod: return "fmod";
    case Op::FRem: return "frem";
    case Op::FSign: return "fsign";
    case Op::Trunc: return "trunc";
    case Op::RoundEven: return "round_even";
    case Op::Exp2: return "exp2";
    case Op::Log2: return "log2";
    case Op::Exp: return "exp";
    case Op::Log: return "log";
    case Op::Pow: return "pow";
    case Op::Sin: return "sin";
    case Op::Cos: return "cos";
    case Op::FCmp: return "fcmp";
    case Op::ICmp: return "icmp";
    case Op::ConvUToF: return "convert_u_to_f";
    case Op::ConvFToU: return "convert_f_to_u";
    case Op::Bitcast: return "bitcast";
    case Op::Any: return "any";
    case Op::All: return "all";
    case Op::IsNan: return "is_nan";
    case Op::IsInf: return "is_inf";
    case Op::Atan2: return "atan2";
    case Op::BitCount: return "bit_count";
    case Op::BitReverse: return "bit_reverse";
    case Op::FindILsb: return "find_lsb";
    case Op::FindUMsb: return "find_umsb";
    case Op::FindSMsb: return "find_smsb";
    case Op::BitFieldInsert: return "bitfield_insert";
    case Op::BitFieldSExtract: return "bitfield_sextract";
    case Op::BitFieldUExtract: return "bitfield_uextract";
    case Op::BufferSize: return "buffer_size";
    case Op::F32ToF16: return "f32_to_f16_bits";
    case Op::F16ToF32: return "f16_bits_to_f32";
    case Op::Ldexp: return "ldexp";
    case Op::FrexpMant: return "frexp_mantissa";
    case Op::FrexpExp: return "frexp_exponent";
    case Op::PtrIndex: return "ptr_index";
    case Op::Block: return "block";
    case Op::Loop: return "loop";
    case Op::ContinueLoop: return "continue";
    case Op::Switch: return "branch";
    case Op::Unreachable: return "unreachable";
    case Op::Barrier: return "barrier";
    case Op::Atomic: return "atomic";
    }
    return "?";
}

const char *fcmp_text(uint32_t p) {
    static const char *const k[] = {"oeq", "one", "olt", "ole", "ogt", "oge", "ueq",
                                     "une", "ult", "ule", "ugt", "uge", "ord", "uno"};
    return p < sizeof k / sizeof k[0] ? k[p] : "?";
}

const char *atomic_text(uint32_t p) {
    static const char *const k[] = {"add", "sub", "smin", "smax", "umin", "umax", "and", "or", "xor", "exchange"};
    return p < sizeof k / sizeof k[0] ? k[p] : "?";
}

const char *icmp_text(uint32_t p) {
    static const char *const k[] = {"eq", "ne", "slt", "sle", "sgt", "sge", "ult", "ule", "ugt", "uge"};
    return p < sizeof k / sizeof k[0] ? k[p] : "?";
}

struct Printer {
    std::string out;
    int indent = 0;
    void line(const std::string &s) {
        out.append(4 * indent, ' ');
        out += s;
        out += '\n';
    }
};

std::string value_use(const Module &m, uint32_t id) {
    const Value &v = m.value(id);
    if (!v.is_const) return v.name;
    std::string s;
    if (v.type.lanes > 1) s += "<";
    for (int i = 0; i < v.type.lanes; ++i) {
        if (i) s += ", ";
        if (v.type.scalar == Scalar::F32) {
            float f;
            std::memcpy(&f, &v.const_bits[i], 4);
            char b[48];
            std::snprintf(b, sizeof b, "%s_f32", hex(v.const_bits[i]).c_str());
            s += b;
        } else {
            s += hex(v.const_bits[i]) + "_" + scalar_text(v.type.scalar);
        }
    }
    if (v.type.lanes > 1) s += ">";
    return s;
}

// (`size` is printed when the storage is larger than one value of the type:
// arrays and structs that the IR keeps in memory)
std::string variable_text(const Module &m, const Type &t, uint32_t align, uint32_t ptr, uint32_t size = 0) {
    std::string s = type_text(t) + ", align: " + hex(align);
    if (size && size != t.size_bytes()) s += ", size: " + hex(size);
    return s + " -> " + m.value(ptr).name + ": data_ptr";
}

const char *builtin_text(BuiltIn b) {  // (builtin_of below is its inverse)
    switch (b) {
    case BuiltIn::VertexPosition: return "vertex_position";
    case BuiltIn::VertexIndex: return "vertex_index";
    case BuiltIn::InstanceIndex: return "instance_index";
    case BuiltIn::FragCoord: return "fragment_coord";
    case BuiltIn::GlobalInvocationId: return "global_invocation_id";
    case BuiltIn::LocalInvocationId: return "local_invocation_id";
    case BuiltIn::WorkgroupId: return "workgroup_id";
    case BuiltIn::NumWorkgroups: return "num_workgroups";
    case BuiltIn::LocalInvocationIndex: return "local_invocation_index";
    }
    return "?";
}

void print_block(Printer &p, const Module &m, const InterfaceBlock &b, const char *name) {
    p.line(std::string(name) + " {");
    ++p.indent;
    p.line("-> " + m.value(b.pointer).name + ": data_ptr;");
    p.line("size: fixed " + hex(b.size) + ";");
    p.line("align: " + hex(b.alignment) + ";");
    for (const auto &mem : b.members) p.line(hex(mem.offset) + ": " + type_text(mem.type) + ": {};");
    --p.indent;
    p.line("}");
}

}  // namespace

std::string type_text(const Type &t) {
    if (t.lanes > 1) return "<" + std::to_string(t.lanes) + " x " + scalar_text(t.scalar) + ">";
    return scalar_text(t.scalar);
}

namespace {

// `want`, or want_2, want_3, ... : the first not in `taken`
template <class Taken>
std::string unique_name(const std::string &want, const Taken &taken) {
    if (!taken(want)) return want;
    for (uint32_t k = 2;; ++k) {
        const std::string n = want + "_" + std::to_string(k);
        if (!taken(n)) return n;
    }
}

}  // namespace

uint32_t Module::new_value(const Type &t, const std::string &name) {
    Value v;
    v.id = (uint32_t)values.size() + 1;
    v.type = t;
    const std::string want = name.empty() ? "v" + std::to_string(v.id) : name;
    v.name = unique_name(want, [&](const std::string &n) { return value_names.count(n) != 0; });
    value_names.insert(v.name);
    values.push_back(v);
    return v.id;
}

uint32_t Module::new_const(const Type &t) {
    Value v;
    v.id = (uint32_t)values.size() + 1;
    v.type = t;
    v.is_const = true;
    values.push_back(v);
    return v.id;
}

uint32_t Function::new_block(const std::string &name, bool loop_body) {
    const std::string n = unique_name(name, [&](const std::string &c) {
        for (const Block &b : blocks)
            if (b.name == c) return true;
        return false;
    });
    blocks.push_back(Block{n, {}, loop_body});
    return (uint32_t)blocks.size() - 1;
}

const char *op_name(Op op) { return op_text(op); }

bool builtin_of(const std::string &s, BuiltIn &out) {
    for (BuiltIn b : {BuiltIn::VertexPosition, BuiltIn::VertexIndex, BuiltIn::InstanceIndex, BuiltIn::FragCoord,
                      BuiltIn::GlobalInvocationId, BuiltIn::LocalInvocationId, BuiltIn::WorkgroupId,
                      BuiltIn::NumWorkgroups, BuiltIn::LocalInvocationIndex})
        if (s == builtin_text(b)) {
            out = b;
            return true;
        }
    return false;
}

bool fcmp_of(const std::string &s, uint32_t &out) {
    for (uint32_t p = 0; p <= FUno; ++p)
        if (s == fcmp_text(p)) {
            out = p;
            return true;
        }
    return false;
}

bool atomic_of(const std::string &s, uint32_t &out) {
    for (uint32_t p = 0; p <= AExchange; ++p)
        if (s == atomic_text(p)) {
            out = p;
            return true;
        }
    return false;
}

bool icmp_of(const std::string &s, uint32_t &out) {
    for (uint32_t p = 0; p <= IUGe; ++p)
        if (s == icmp_text(p)) {
            out = p;
            return true;
        }
    return false;
}

uint32_t Module::new_const_f32(float f) {
    const uint32_t id = new_const(Type::f32());
    Value &v = value(id);
    std::memcpy(&v.const_bits[0], &f, 4);
    return id;
}

uint32_t Module::new_const_i32(int32_t i) {
    const uint32_t id = new_const(Type::i32());
    Value &v = value(id);
    v.const_bits[0] = (uint32_t)i;
    return id;
}

std::string instruction_text(const Module &m, const Instruction &ins) {
    std::string s = op_text(ins.op);
    auto block_name = [&](uint32_t b) { return b < m.main.blocks.size() ? m.main.blocks[b].name : "?"; };
    switch (ins.op) {
    case Op::BreakBlock:
    case Op::ContinueLoop: s += " " + block_name(ins.fields.at(0)) + "[]"; break;
    case Op::Kill:
    case Op::Unreachable: break;
    case Op::Switch: {
        s += " " + value_use(m, ins.args.at(0)) + " {";
        for (size_t i = 1; i + 1 < ins.fields.size(); i += 2)
            s += " " + hex(ins.fields[i]) + " -> " + block_name(ins.fields[i + 1]) + ",";
        s += " default -> " + block_name(ins.fields.at(0)) + " }";
        break;
    }
    default: {
        s += " [";
        for (size_t i = 0; i < ins.args.size(); ++i) {
            if (i) s += ", ";
            s += value_use(m, ins.args[i]);
        }
        s += "]";
        switch (ins.op) {
        case Op::Load:
        case Op::Store: s += ", align: " + hex(ins.fields.empty() ? 4 : ins.fields[0]); break;
        case Op::PtrOffset: s += ", offset: " + hex(ins.fields[0]); break;
        case Op::PtrIndex: s += ", stride: " + hex(ins.fields[0]); break;
        case Op::Extract: s += ", index: " + std::to_string(ins.fields[0]); break;
        case Op::FCmp: s += ", " + std::string(fcmp_text(ins.fields[0])); break;
        case Op::ICmp: s += ", " + std::string(icmp_text(ins.fields[0])); break;
        case Op::Atomic: s += ", " + std::string(atomic_text(ins.fields[0])); break;
        case Op::Shuffle: {
            s += ", lanes: [";
            for (size_t i = 0; i < ins.fields.size(); ++i) s += (i ? ", " : "") + std::to_string(ins.fields[i]);
            s += "]";
            break;
        }
        default: break;
        }
        s += " -> [";
        for (size_t i = 0; i < ins.results.size(); ++i) {
            if (i) s += ", ";
            const Value &v = m.value(ins.results[i]);
            s += v.name + ": " + type_text(v.type);
        }
        s += "]";
    }
    }
    if (ins.has_location)
        s += " @ \"" + ins.loc.file + "\":" + std::to_string(ins.loc.line) + ":" + std::to_string(ins.loc.column);
    return s;
}

namespace {

// Block and Loop print their bodies nested, as the reference's text form
// does (block.rs: `block NAME -> [] { ... }`, `loop NAME[] -> [] { -> [];
// BODY { ... } }`).
void print_body(Printer &p, const Module &m, uint32_t b);

void print_instruction(Printer &p, const Module &m, const Instruction &ins) {
    if (ins.op == Op::Block) {
        const uint32_t b = ins.fields.at(0);
        p.line("block " + m.main.blocks[b].name + " -> [] {");
        ++p.indent;
        print_body(p, m, b);
        --p.indent;
        p.line("};");
    } else if (ins.op == Op::Loop) {
        const uint32_t b = ins.fields.at(0);
        p.line("loop " + m.main.blocks[b].name + "[] -> [] {");
        ++p.indent;
        p.line("-> [];");
        p.line(m.main.blocks[b].name + " {");
        ++p.indent;
        print_body(p, m, b);
        --p.indent;
        p.line("}");
        --p.indent;
        p.line("};");
    } else {
        p.line(instruction_text(m, ins) + ";");
    }
}

void print_body(Printer &p, const Module &m, uint32_t b) {
    for (const auto &ins : m.main.blocks[b].body) print_instruction(p, m, ins);
}

}  // namespace

std::string module_text(const Module &m) {
    Printer p;
    p.line("module {");
    ++p.indent;
    p.line("target_properties {");
    ++p.indent;
    p.line("data_pointer_underlying_type: i64,");
    p.line("function_pointer_underlying_type: i64,");
    --p.indent;
    p.line("}");
    p.line("built_in_inputs {");
    ++p.indent;
    for (const auto &v : m.built_in_inputs)
        p.line(variable_text(m, v.type, v.alignment, v.pointer) + ": { kind: " + builtin_text(v.kind) + " };");
    --p.indent;
    p.line("}");
    print_block(p, m, m.user_inputs, "user_inputs_block");
    p.line("built_in_outputs {");
    ++p.indent;
    for (const auto &v : m.built_in_outputs)
        p.line(variable_text(m, v.type, v.alignment, v.pointer) + ": { kind: " + builtin_text(v.kind) + " };");
    --p.indent;
    p.line("}");
    print_block(p, m, m.user_outputs, "user_outputs_block");
    p.line("invocation_global_variables {");
    ++p.indent;
    for (const auto &v : m.invocation_globals) p.line(variable_text(m, v.type, v.alignment, v.pointer, v.size) + ";");
    --p.indent;
    p.line("}");
    if (!m.workgroup_variables.empty()) {
        p.line("workgroup_variables {");
        ++p.indent;
        for (const auto &v : m.workgroup_variables)
            p.line(variable_text(m, v.type, v.alignment, v.pointer, v.size) + ";");
        --p.indent;
        p.line("}");
    }
    if (!m.resources.empty()) {
        p.line("resources {");
        ++p.indent;
        for (const auto &r : m.resources) {
            const char *k = r.kind == ResourceKind::PushConstants   ? "push_constants"
                            : r.kind == ResourceKind::UniformBuffer ? "uniform_buffer"
                            : r.kind == ResourceKind::StorageBuffer ? "storage_buffer"
                                                                     : "combined_image_sampler";
            std::string s = std::string(k) + " -> " + m.value(r.pointer).name + ": " + type_text(m.value(r.pointer).type);
            if (r.kind != ResourceKind::PushConstants)
                s += ": { set: " + std::to_string(r.set) + ", binding: " + std::to_string(r.binding) + " }";
            p.line(s + ";");
        }
        --p.indent;
        p.line("}");
    }
    p.line("fn " + m.main.name + "[] -> [] {");
    ++p.indent;
    p.line("hints {");
    ++p.indent;
    p.line("inlining_hint: none,");
    p.line("side_effects: normal,");
    --p.indent;
    p.line("}");
    p.line("{");
    ++p.indent;
    for (const auto &v : m.main.locals) p.line(variable_text(m, v.type, v.alignment, v.pointer, v.size) + ";");
    --p.indent;
    p.line("}");
    p.line(m.main.body().name + " {");
    ++p.indent;
    print_body(p, m, 0);
    --p.indent;
    p.line("}");
    --p.indent;
    p.line("}");
    p.line("entry_point: " + m.main.name + ";");
    --p.indent;
    p.line("}");
    return p.out;
}

}  // namespace ir
}  // namespace kz
