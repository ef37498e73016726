// ir_text.cpp — parser of the shader IR's text form (the inverse of
// module_text in ir.cpp).
//
// Counterpart of the reference's IR text parser
// (shader-compiler-ir/src/text.rs, FromText for Module and for each
// instruction, used by assert_ir_matches_file!, macros.rs:39-80): a module
// printed by module_text parses back to a module that prints identically,
// and so do the reference's golden files. Grammar (one production per
// printed construct):
//
//   module      = "module" "{" target built_in_inputs user_block
//                 built_in_outputs user_block globals [workgroup_vars]
//                 [resources] function "entry_point" ":" NAME ";" "}"
//
// (The stage and a compute shader's local size are not part of the text,
// as in the reference's form; whoever parses a module supplies them.)
//   variable    = type "," "align" ":" HEX ["," "size" ":" HEX] "->" NAME ":" "data_ptr"
//   user_block  = SECTION "{" "->" NAME ":" "data_ptr" ";" "size" ":" "fixed"
//                 HEX ";" "align" ":" HEX ";" { HEX ":" type ":" "{" "}" ";" } "}"
//   function    = "fn" NAME "[" "]" "->" "[" "]" "{" hints "{" { variable ";" } "}"
//                 NAME "{" body "}" "}"
//   instruction = "block" NAME "->" "[" "]" "{" body "}"
//               | "loop" NAME "[" "]" "->" "[" "]" "{" "->" "[" "]" ";" NAME "{" body "}" "}"
//               | ("break" | "continue") NAME "[" "]" | "kill" | "unreachable"
//               | "branch" use "{" { HEX "->" NAME "," } "default" "->" NAME "}"
//               | OP "[" [use { "," use }] "]" { "," field } "->" "[" [result { "," result }] "]"
//               followed by an optional location  "@" STRING ":" INT ":" INT
//   use         = NAME | const | "<" const { "," const } ">"
//   const       = HEX "_" scalar_type
//
// Value names are unique in a module (Module::new_value makes them so), so a
// use names the one definition; break and continue name an enclosing block
// or loop, innermost first.
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <map>

#include "ir.h"

namespace kz {
namespace ir {

namespace {

struct Tok {
    enum Kind { End, Ident, Hex, Dec, Const, Str, Punct } kind = End;
    std::string text;  // identifier, string contents, punctuation
    uint64_t num = 0;  // Hex, Dec, Const (bits)
    std::string suffix;  // Const: its scalar type
    int line = 1;
};

class Lexer {
  public:
    explicit Lexer(const std::string &s) : s_(s) {}

    bool next(Tok &t, std::string &err) {
        skip();
        t = Tok{};
        t.line = line_;
        if (p_ >= s_.size()) return true;
        const char c = s_[p_];
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t q = p_;
            while (q < s_.size() && (std::isalnum((unsigned char)s_[q]) || s_[q] == '_')) ++q;
            t.kind = Tok::Ident;
            t.text = s_.substr(p_, q - p_);
            p_ = q;
            return true;
        }
        if (std::isdigit((unsigned char)c)) {
            size_t q = p_;
            if (c == '0' && q + 1 < s_.size() && (s_[q + 1] == 'x' || s_[q + 1] == 'X')) {
                q += 2;
                const size_t d0 = q;
                while (q < s_.size() && std::isxdigit((unsigned char)s_[q])) ++q;
                if (q == d0) return fail(err, "hex number without digits");
                t.num = std::strtoull(s_.substr(d0, q - d0).c_str(), nullptr, 16);
                t.kind = Tok::Hex;
                if (q < s_.size() && s_[q] == '_') {  // a constant: 0x3F800000_f32
                    size_t r = q + 1;
                    while (r < s_.size() && (std::isalnum((unsigned char)s_[r]) || s_[r] == '_')) ++r;
                    t.kind = Tok::Const;
                    t.suffix = s_.substr(q + 1, r - q - 1);
                    q = r;
                }
            } else {
                while (q < s_.size() && std::isdigit((unsigned char)s_[q])) ++q;
                t.num = std::strtoull(s_.substr(p_, q - p_).c_str(), nullptr, 10);
                t.kind = Tok::Dec;
            }
            p_ = q;
            return true;
        }
        if (c == '"') {
            size_t q = p_ + 1;
            while (q < s_.size() && s_[q] != '"') ++q;
            if (q >= s_.size()) return fail(err, "unterminated string");
            t.kind = Tok::Str;
            t.text = s_.substr(p_ + 1, q - p_ - 1);
            p_ = q + 1;
            return true;
        }
        if (c == '-' && p_ + 1 < s_.size() && s_[p_ + 1] == '>') {
            t.kind = Tok::Punct;
            t.text = "->";
            p_ += 2;
            return true;
        }
        if (std::strchr("{}[]<>:;,@", c)) {
            t.kind = Tok::Punct;
            t.text = std::string(1, c);
            ++p_;
            return true;
        }
        return fail(err, std::string("unexpected character '") + c + "'");
    }

  private:
    void skip() {
        while (p_ < s_.size()) {
            if (s_[p_] == '\n') {
                ++line_;
                ++p_;
            } else if (std::isspace((unsigned char)s_[p_])) {
                ++p_;
            } else if (s_[p_] == '#') {  // comment to the end of the line
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else {
                break;
            }
        }
    }
    bool fail(std::string &err, const std::string &what) {
        err = "line " + std::to_string(line_) + ": " + what;
        return false;
    }
    const std::string &s_;
    size_t p_ = 0;
    int line_ = 1;
};

bool scalar_of(const std::string &s, Scalar &out) {
    static const std::map<std::string, Scalar> k = {{"void", Scalar::Void}, {"bool", Scalar::Bool},
                                                   {"i32", Scalar::I32},   {"f32", Scalar::F32},
                                                   {"data_ptr", Scalar::DataPtr}, {"texture", Scalar::Texture}};
    auto it = k.find(s);
    if (it == k.end()) return false;
    out = it->second;
    return true;
}

// op name -> opcode, for the instructions that print as OP [args] ... -> [results]
const std::map<std::string, Op> &op_names() {
    static const std::map<std::string, Op> k = [] {
        std::map<std::string, Op> m;
        const Op plain[] = {Op::Add,   Op::FAdd,     Op::FSub,      Op::FMul,     Op::FDiv,     Op::FNeg,
                            Op::FMax,  Op::FMin,     Op::FClamp,    Op::Sqrt,     Op::Floor,    Op::Ceil,
                            Op::Dot,   Op::Splat,    Op::MakeVector, Op::Extract, Op::Shuffle,  Op::Load,
                            Op::Store, Op::PtrOffset, Op::Sample,   Op::ConvSToF, Op::ConvFToS, Op::Select,
                            Op::FOrdLt, Op::FOrdGt,  Op::ISub,      Op::IMul,     Op::SDiv,     Op::UDiv,
                            Op::SRem,  Op::SMod,     Op::UMod,      Op::Shl,      Op::ShrL,     Op::ShrA,
                            Op::And,   Op::Or,       Op::Xor,       Op::Not,      Op::INeg,     Op::SMin,
                            Op::SMax,  Op::UMin,     Op::UMax,      Op::SAbs,     Op::SClamp,   Op::UClamp,
                            Op::FMod,  Op::FRem,     Op::FSign,     Op::Trunc,    Op::RoundEven, Op::Exp2,
                            Op::Log2,  Op::Exp,      Op::Log,       Op::Pow,      Op::Sin,      Op::Cos,
                            Op::FCmp,  Op::ICmp,     Op::ConvUToF,  Op::ConvFToU, Op::Bitcast,  Op::Any,
                            Op::All,   Op::IsNan,    Op::IsInf,     Op::PtrIndex, Op::Barrier,  Op::Atomic,
                            Op::Atan2, Op::BitCount, Op::BitReverse, Op::FindILsb, Op::FindUMsb, Op::FindSMsb,
                            Op::BitFieldInsert, Op::BitFieldSExtract, Op::BitFieldUExtract, Op::BufferSize,
                            Op::F32ToF16, Op::F16ToF32, Op::Ldexp, Op::FrexpMant, Op::FrexpExp};
        for (Op op : plain) m[op_name(op)] = op;
        return m;
    }();
    return k;
}

class Parser {
  public:
    Parser(const std::string &text, Module &m, std::string &err) : lex_(text), m_(m), err_(err) {}

    bool module() {
        if (!advance() || !word("module") || !punct("{")) return false;
        if (!word("target_properties") || !punct("{") || !word("data_pointer_underlying_type") || !punct(":") ||
            !word("i64") || !punct(",") || !word("function_pointer_underlying_type") || !punct(":") ||
            !word("i64") || !punct(",") || !punct("}"))
            return false;
        if (!interface_vars("built_in_inputs", m_.built_in_inputs)) return false;
        if (!user_block("user_inputs_block", m_.user_inputs)) return false;
        if (!interface_vars("built_in_outputs", m_.built_in_outputs)) return false;
        if (!user_block("user_outputs_block", m_.user_outputs)) return false;
        if (!word("invocation_global_variables") || !punct("{")) return false;
        while (!is_punct("}")) {
            Variable v;
            if (!variable(v) || !punct(";")) return false;
            m_.invocation_globals.push_back(v);
        }
        if (!punct("}")) return false;
        if (is_word("workgroup_variables")) {
            if (!advance() || !punct("{")) return false;
            while (!is_punct("}")) {
                Variable v;
                if (!variable(v) || !punct(";")) return false;
                m_.workgroup_variables.push_back(v);
            }
            if (!punct("}")) return false;
        }
        if (is_word("resources") && !resources()) return false;
        if (!function()) return false;
        if (!word("entry_point") || !punct(":")) return false;
        std::string entry;
        if (!name(entry) || !punct(";") || !punct("}")) return false;
        if (entry != m_.main.name) return fail("entry point " + entry + " is not the module's function");
        if (t_.kind != Tok::End) return fail("text after the module");
        return true;
    }

    // "[defs]\n instruction": value definitions, then one instruction (the
    // reference's per-instruction text tests, instructions_impl.rs:560-622)
    bool instruction_with_defs(Instruction &ins) {
        if (!advance() || !punct("[")) return false;
        while (!is_punct("]")) {
            std::string n;
            Type t;
            if (!name(n) || !punct(":") || !type(t)) return false;
            if (!define(n, t)) return false;
            if (!is_punct("]") && !punct(",")) return false;
        }
        if (!punct("]")) return false;
        open_.push_back(0);
        if (!instruction(ins)) return false;
        if (t_.kind != Tok::End) return fail("text after the instruction");
        return true;
    }

  private:
    bool advance() { return lex_.next(t_, err_); }
    bool fail(const std::string &what) {
        err_ = "line " + std::to_string(t_.line) + ": " + what + (t_.text.empty() ? "" : " (at '" + t_.text + "')");
        return false;
    }
    bool is_punct(const char *p) const { return t_.kind == Tok::Punct && t_.text == p; }
    bool is_word(const char *w) const { return t_.kind == Tok::Ident && t_.text == w; }
    bool punct(const char *p) {
        if (!is_punct(p)) return fail(std::string("expected '") + p + "'");
        return advance();
    }
    bool word(const char *w) {
        if (!is_word(w)) return fail(std::string("expected '") + w + "'");
        return advance();
    }
    bool name(std::string &out) {
        if (t_.kind != Tok::Ident) return fail("expected a name");
        out = t_.text;
        return advance();
    }
    bool hex(uint32_t &out) {
        if (t_.kind != Tok::Hex) return fail("expected a hex number");
        out = (uint32_t)t_.num;
        return advance();
    }
    bool dec(uint32_t &out) {
        if (t_.kind != Tok::Dec) return fail("expected a number");
        out = (uint32_t)t_.num;
        return advance();
    }

    bool type(Type &out) {
        if (is_punct("<")) {
            uint32_t n = 0;
            Scalar s;
            if (!advance() || !dec(n) || !word("x")) return false;
            if (t_.kind != Tok::Ident || !scalar_of(t_.text, s)) return fail("expected a scalar type");
            if (n < 2 || n > 4) return fail("vector of " + std::to_string(n) + " lanes");
            if (!advance() || !punct(">")) return false;
            out = Type{s, (uint8_t)n};
            return true;
        }
        Scalar s;
        if (t_.kind != Tok::Ident || !scalar_of(t_.text, s)) return fail("expected a type");
        out = Type{s, 1};
        return advance();
    }

    bool define(const std::string &n, const Type &t, uint32_t *id = nullptr) {
        if (names_.count(n)) return fail("value " + n + " is defined twice");
        const uint32_t v = m_.new_value(t, n);
        if (m_.value(v).name != n) return fail("value name " + n + " is taken");
        names_[n] = v;
        if (id) *id = v;
        return true;
    }

    bool constant(uint32_t &bits, Scalar &s) {
        if (t_.kind != Tok::Const || !scalar_of(t_.suffix, s) || s == Scalar::Void || s == Scalar::DataPtr ||
            s == Scalar::Texture)
            return fail("expected a constant");
        bits = (uint32_t)t_.num;
        return advance();
    }

    // a value use: a defined name or a constant (scalar or vector)
    bool use(uint32_t &id) {
        if (t_.kind == Tok::Ident) {
            auto it = names_.find(t_.text);
            if (it == names_.end()) return fail("use of undefined value " + t_.text);
            id = it->second;
            return advance();
        }
        std::vector<uint32_t> bits;
        Scalar s = Scalar::Void;
        if (is_punct("<")) {
            if (!advance()) return false;
            for (;;) {
                uint32_t b;
                Scalar si;
                if (!constant(b, si)) return false;
                if (!bits.empty() && si != s) return fail("vector constant of mixed types");
                s = si;
                bits.push_back(b);
                if (is_punct(">")) break;
                if (!punct(",")) return false;
            }
            if (!advance()) return false;
            if (bits.size() < 2 || bits.size() > 4) return fail("vector constant of " + std::to_string(bits.size()) + " lanes");
        } else {
            uint32_t b;
            if (!constant(b, s)) return false;
            bits.push_back(b);
        }
        id = m_.new_const(Type{s, (uint8_t)bits.size()});
        Value &v = m_.value(id);
        for (size_t i = 0; i < bits.size(); ++i) v.const_bits[i] = bits[i];
        return true;
    }

    bool variable(Variable &v) {
        std::string n;
        if (!type(v.type) || !punct(",") || !word("align") || !punct(":") || !hex(v.alignment)) return false;
        v.size = v.type.size_bytes();
        if (is_punct(",") && (!advance() || !word("size") || !punct(":") || !hex(v.size))) return false;
        if (!punct("->") || !name(n) || !punct(":") || !word("data_ptr")) return false;
        return define(n, Type::ptr(), &v.pointer);
    }

    bool interface_vars(const char *section, std::vector<InterfaceVariable> &out) {
        if (!word(section) || !punct("{")) return false;
        while (!is_punct("}")) {
            Variable v;
            std::string kind;
            if (!variable(v) || !punct(":") || !punct("{") || !word("kind") || !punct(":") || !name(kind) ||
                !punct("}") || !punct(";"))
                return false;
            InterfaceVariable iv;
            iv.type = v.type;
            iv.alignment = v.alignment;
            iv.pointer = v.pointer;
            if (!builtin_of(kind, iv.kind)) return fail("unknown built-in " + kind);
            out.push_back(iv);
        }
        return punct("}");
    }

    bool user_block(const char *section, InterfaceBlock &b) {
        std::string n;
        if (!word(section) || !punct("{") || !punct("->") || !name(n) || !punct(":") || !word("data_ptr") ||
            !punct(";") || !define(n, Type::ptr(), &b.pointer))
            return false;
        if (!word("size") || !punct(":") || !word("fixed") || !hex(b.size) || !punct(";") || !word("align") ||
            !punct(":") || !hex(b.alignment) || !punct(";"))
            return false;
        while (!is_punct("}")) {
            BlockMember mem;
            if (!hex(mem.offset) || !punct(":") || !type(mem.type) || !punct(":") || !punct("{") || !punct("}") ||
                !punct(";"))
                return false;
            b.members.push_back(mem);
        }
        return punct("}");
    }

    bool resources() {
        if (!word("resources") || !punct("{")) return false;
        while (!is_punct("}")) {
            Resource r;
            std::string kind, n;
            Type t;
            if (!name(kind) || !punct("->") || !name(n) || !punct(":") || !type(t)) return false;
            if (kind == "push_constants") r.kind = ResourceKind::PushConstants;
            else if (kind == "uniform_buffer") r.kind = ResourceKind::UniformBuffer;
            else if (kind == "combined_image_sampler") r.kind = ResourceKind::CombinedImageSampler;
            else if (kind == "storage_buffer") r.kind = ResourceKind::StorageBuffer;
            else return fail("unknown resource kind " + kind);
            if (r.kind != ResourceKind::PushConstants) {
                if (!punct(":") || !punct("{") || !word("set") || !punct(":") || !dec(r.set) || !punct(",") ||
                    !word("binding") || !punct(":") || !dec(r.binding) || !punct("}"))
                    return false;
            }
            if (!punct(";") || !define(n, t, &r.pointer)) return false;
            m_.resources.push_back(r);
        }
        return punct("}");
    }

    bool function() {
        if (!word("fn") || !name(m_.main.name) || !punct("[") || !punct("]") || !punct("->") || !punct("[") ||
            !punct("]") || !punct("{"))
            return false;
        if (!word("hints") || !punct("{") || !word("inlining_hint") || !punct(":") || !word("none") || !punct(",") ||
            !word("side_effects") || !punct(":") || !word("normal") || !punct(",") || !punct("}"))
            return false;
        if (!punct("{")) return false;
        while (!is_punct("}")) {
            Variable v;
            if (!variable(v) || !punct(";")) return false;
            m_.main.locals.push_back(v);
        }
        if (!punct("}") || !name(m_.main.blocks[0].name) || !punct("{")) return false;
        open_.push_back(0);
        if (!body(0)) return false;
        open_.pop_back();
        return punct("}") && punct("}");
    }

    // an enclosing block (or loop body, for `continue`) by name, innermost first
    bool enclosing(const std::string &n, bool loop, uint32_t &out) {
        for (size_t i = open_.size(); i-- > 0;) {
            const Block &b = m_.main.blocks[open_[i]];
            if (b.name == n && (!loop || b.loop_body)) {
                out = open_[i];
                return true;
            }
        }
        return fail(std::string(loop ? "continue to " : "break from ") + n + ", which does not enclose it");
    }

    bool new_block(const std::string &n, bool loop_body, uint32_t &out) {
        for (const Block &b : m_.main.blocks)
            if (b.name == n) return fail("block " + n + " is defined twice");
        out = m_.main.new_block(n, loop_body);
        return true;
    }

    bool body(uint32_t b) {
        while (!is_punct("}")) {
            Instruction ins;
            if (!instruction(ins) || !punct(";")) return false;
            m_.main.blocks[b].body.push_back(std::move(ins));
        }
        return true;
    }

    bool instruction(Instruction &ins) {
        if (t_.kind != Tok::Ident) return fail("expected an instruction");
        const std::string op = t_.text;
        if (!advance()) return false;
        if (op == "block") {
            std::string n;
            uint32_t b;
            if (!name(n) || !punct("->") || !punct("[") || !punct("]") || !punct("{") || !new_block(n, false, b))
                return false;
            ins.op = Op::Block;
            ins.fields = {b};
            open_.push_back(b);
            if (!body(b)) return false;
            open_.pop_back();
            return punct("}");
        }
        if (op == "loop") {
            std::string n, n2;
            uint32_t b;
            if (!name(n) || !punct("[") || !punct("]") || !punct("->") || !punct("[") || !punct("]") || !punct("{") ||
                !punct("->") || !punct("[") || !punct("]") || !punct(";") || !name(n2) || !punct("{"))
                return false;
            if (n2 != n) return fail("loop " + n + " has a body named " + n2);
            if (!new_block(n, true, b)) return false;
            ins.op = Op::Loop;
            ins.fields = {b};
            open_.push_back(b);
            if (!body(b)) return false;
            open_.pop_back();
            return punct("}") && punct("}");
        }
        if (op == "break" || op == "continue") {
            std::string n;
            uint32_t b;
            if (!name(n) || !punct("[") || !punct("]") || !enclosing(n, op == "continue", b)) return false;
            ins.op = op == "break" ? Op::BreakBlock : Op::ContinueLoop;
            ins.fields = {b};
            return location(ins);
        }
        if (op == "kill" || op == "unreachable") {
            ins.op = op == "kill" ? Op::Kill : Op::Unreachable;
            return location(ins);
        }
        if (op == "branch") {
            uint32_t sel;
            if (!use(sel) || !punct("{")) return false;
            ins.op = Op::Switch;
            ins.args = {sel};
            std::vector<uint32_t> cases;
            while (t_.kind == Tok::Hex) {
                uint32_t lit, b;
                std::string n;
                if (!hex(lit) || !punct("->") || !name(n) || !punct(",") || !enclosing(n, false, b)) return false;
                cases.push_back(lit);
                cases.push_back(b);
            }
            std::string n;
            uint32_t d;
            if (!word("default") || !punct("->") || !name(n) || !punct("}") || !enclosing(n, false, d)) return false;
            ins.fields.push_back(d);
            ins.fields.insert(ins.fields.end(), cases.begin(), cases.end());
            return location(ins);
        }
        auto it = op_names().find(op);
        if (it == op_names().end()) return fail("unknown instruction " + op);
        ins.op = it->second;
        if (!punct("[")) return false;
        while (!is_punct("]")) {
            uint32_t a;
            if (!use(a)) return false;
            ins.args.push_back(a);
            if (!is_punct("]") && !punct(",")) return false;
        }
        if (!punct("]")) return false;
        if (!fields(ins)) return false;
        if (!punct("->") || !punct("[")) return false;
        std::vector<std::pair<std::string, Type>> res;
        while (!is_punct("]")) {
            std::string n;
            Type t;
            if (!name(n) || !punct(":") || !type(t)) return false;
            res.push_back({n, t});
            if (!is_punct("]") && !punct(",")) return false;
        }
        if (!punct("]")) return false;
        for (const auto &r : res) {
            uint32_t id;
            if (!define(r.first, r.second, &id)) return false;
            ins.results.push_back(id);
        }
        return location(ins);
    }

    // the immediates printed after an instruction's arguments
    bool fields(Instruction &ins) {
        switch (ins.op) {
        case Op::Load:
        case Op::Store:
        case Op::PtrOffset:
        case Op::PtrIndex: {
            const char *key = ins.op == Op::PtrOffset ? "offset" : ins.op == Op::PtrIndex ? "stride" : "align";
            uint32_t v;
            if (!punct(",") || !word(key) || !punct(":") || !hex(v)) return false;
            ins.fields = {v};
            return true;
        }
        case Op::Extract: {
            uint32_t v;
            if (!punct(",") || !word("index") || !punct(":") || !dec(v)) return false;
            ins.fields = {v};
            return true;
        }
        case Op::FCmp:
        case Op::ICmp:
        case Op::Atomic: {
            std::string p;
            if (!punct(",") || !name(p)) return false;
            uint32_t pred;
            const bool known = ins.op == Op::FCmp ? fcmp_of(p, pred) : ins.op == Op::ICmp ? icmp_of(p, pred)
                                                                                         : atomic_of(p, pred);
            if (!known) return fail("unknown predicate " + p);
            ins.fields = {pred};
            return true;
        }
        case Op::Shuffle: {
            if (!punct(",") || !word("lanes") || !punct(":") || !punct("[")) return false;
            while (!is_punct("]")) {
                uint32_t v;
                if (!dec(v)) return false;
                ins.fields.push_back(v);
                if (!is_punct("]") && !punct(",")) return false;
            }
            return punct("]");
        }
        default: return true;
        }
    }

    bool location(Instruction &ins) {
        if (!is_punct("@")) return true;
        if (!advance()) return false;
        if (t_.kind != Tok::Str) return fail("expected a file name");
        ins.has_location = true;
        ins.loc.file = t_.text;
        return advance() && punct(":") && dec(ins.loc.line) && punct(":") && dec(ins.loc.column);
    }

    Lexer lex_;
    Tok t_;
    Module &m_;
    std::string &err_;
    std::map<std::string, uint32_t> names_;
    std::vector<uint32_t> open_;  // enclosing blocks, outermost first
};

}  // namespace

bool parse_module_text(const std::string &text, Module &m, std::string &err) {
    m = Module{};
    m.main.blocks.clear();
    m.main.blocks.push_back(Block{});
    Parser p(text, m, err);
    return p.module();
}

bool parse_instruction_text(const std::string &text, Module &m, Instruction &ins, std::string &err) {
    m = Module{};
    Parser p(text, m, err);
    return p.instruction_with_defs(ins);
}

}  // namespace ir
}  // namespace kz
