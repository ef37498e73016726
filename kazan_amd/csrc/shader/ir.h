// ir.h — the shader IR between SPIR-V and the HIP backend.
//
// It follows the shape of the reference's shader-compiler-ir
// (shader-compiler-ir/src/module.rs:128-150): a module holds target
// properties, the built-in and user input/output interface blocks
// (interface.rs:134-141; user blocks are structs in memory whose members sit
// at location*16 + component*4, io_layout.rs:6-13), the invocation-global
// variables, functions made of structured blocks, and the entry point.
// Values are SSA, typed, and printed in the reference's text form
// (`add [arg0, arg1] -> [result: i32]`, instructions_impl.rs:608-622).
//
// The reference defines two data instructions (add, load;
// instructions_impl.rs:395-432). This IR adds the ones the draw path's
// shaders need — float arithmetic, store, pointer offsets, vector
// construction/extraction, dot, the GLSL.std.450 subset, texture sampling —
// and a `resources` section for push constants, uniform buffers and
// combined image samplers, which the reference's IR does not model yet.
#pragma once

#include <cstdint>
#include <memory>
#include <set>
#include <string>
#include <vector>

namespace kz {
namespace ir {

enum class Scalar : uint8_t { Void, Bool, I32, F32, DataPtr, Texture };

struct Type {
    Scalar scalar = Scalar::Void;
    uint8_t lanes = 1;  // > 1: vector
    bool operator==(const Type &o) const { return scalar == o.scalar && lanes == o.lanes; }
    bool operator!=(const Type &o) const { return !(*this == o); }
    static Type f32(int n = 1) { return Type{Scalar::F32, (uint8_t)n}; }
    static Type i32(int n = 1) { return Type{Scalar::I32, (uint8_t)n}; }
    static Type boolean(int n = 1) { return Type{Scalar::Bool, (uint8_t)n}; }
    static Type ptr() { return Type{Scalar::DataPtr, 1}; }
    static Type texture() { return Type{Scalar::Texture, 1}; }
    uint32_t size_bytes() const { return scalar == Scalar::DataPtr ? 8u : 4u * lanes; }
};

std::string type_text(const Type &t);

struct Value {
    uint32_t id = 0;
    Type type;
    std::string name;
    bool is_const = false;
    uint32_t const_bits[4] = {0, 0, 0, 0};
};

enum class Op : uint8_t {
    Add,      // integer add
    FAdd,
    FSub,
    FMul,
    FDiv,
    FNeg,
    FMax,
    FMin,
    FClamp,   // [x, lo, hi]
    Sqrt,
    Floor,
    Ceil,
    Dot,      // ((a0*b0 + a1*b1) + a2*b2) + a3*b3
    Splat,    // scalar -> vector
    MakeVector,
    Extract,  // field: index
    Shuffle,  // fields: lane indices into concat(a, b)
    Load,     // field: alignment
    Store,    // field: alignment
    PtrOffset,  // field: byte offset
    Sample,   // [texture, coord] -> <4 x f32>
    ConvSToF,
    ConvFToS,
    Select,   // [cond, a, b]
    FOrdLt,
    FOrdGt,
    Kill,     // discard the fragment (ends the invocation)
    BreakBlock,  // field: target block; jumps to the first instruction after that block
    // integer arithmetic (And, Or, Xor, Not also on bool: the logical forms)
    ISub,
    IMul,
    SDiv,
    UDiv,
    SRem,     // sign of the dividend
    SMod,     // sign of the divisor
    UMod,
    Shl,
    ShrL,
    ShrA,
    And,
    Or,
    Xor,
    Not,
    INeg,
    SMin,
    SMax,
    UMin,
    UMax,
    SAbs,
    SClamp,
    UClamp,
    // float
    FMod,     // x - y * floor(x / y)
    FRem,     // sign of x
    FSign,
    Trunc,
    RoundEven,
    Exp2,
    Log2,
    Exp,
    Log,
    Pow,
    Sin,
    Cos,
    FCmp,     // field: predicate (FCmpPred); result bool lanes
    ICmp,     // field: predicate (ICmpPred); result bool lanes
    ConvUToF,
    ConvFToU,
    Bitcast,
    Any,
    All,
    IsNan,
    IsInf,
    Atan2,             // [y, x]
    BitCount,
    BitReverse,
    FindILsb,          // -1 for 0
    FindUMsb,          // -1 for 0
    FindSMsb,          // -1 for 0 and -1
    BitFieldInsert,    // [base, insert, offset, count]
    BitFieldSExtract,  // [base, offset, count]
    BitFieldUExtract,  // [base, offset, count]
    BufferSize,        // [storage buffer pointer] -> bytes bound to it (i32)
    F32ToF16,          // f32 -> i32 holding the binary16 bits (round to nearest even)
    F16ToF32,          // i32 whose low 16 bits are binary16 -> f32
    Ldexp,             // [x, exponent: i32] -> x * 2^exponent
    FrexpMant,         // x = mantissa * 2^exponent, |mantissa| in [0.5, 1) (0 for 0)
    FrexpExp,          // -> exponent (i32)
    PtrIndex,  // [ptr, index] + field: stride in bytes -> ptr + index * stride
    Block,     // field: block; runs the block's body
    Loop,      // field: the loop's body block; runs it again on each continue
    ContinueLoop,  // field: the loop's body block
    Switch,    // [selector] + fields: default block, then (literal, block) pairs: breaks from the chosen block
    Unreachable,
    Barrier,   // workgroup control + memory barrier (compute)
    Atomic,    // [ptr, value] + field: AtomicOp -> [old value: i32]
};

enum AtomicOp : uint32_t { AAdd, ASub, ASMin, ASMax, AUMin, AUMax, AAnd, AOr, AXor, AExchange };

enum FCmpPred : uint32_t { FOEq, FONe, FOLt, FOLe, FOGt, FOGe, FUEq, FUNe, FULt, FULe, FUGt, FUGe, FOrd, FUno };
enum ICmpPred : uint32_t { IEq, INe, ISLt, ISLe, ISGt, ISGe, IULt, IULe, IUGt, IUGe };

struct Location {
    std::string file;
    uint32_t line = 0, column = 0;
};

struct Instruction {
    Op op;
    std::vector<uint32_t> args;     // value ids
    std::vector<uint32_t> results;  // value ids
    std::vector<uint32_t> fields;   // immediates
    bool has_location = false;
    Location loc;
};

enum class BuiltIn : uint8_t {
    VertexPosition,
    VertexIndex,
    InstanceIndex,
    FragCoord,
    // compute (all <3 x i32> but the index)
    GlobalInvocationId,
    LocalInvocationId,
    WorkgroupId,
    NumWorkgroups,
    LocalInvocationIndex,
};

struct InterfaceVariable {
    Type type;
    uint32_t alignment = 4;
    uint32_t pointer = 0;  // value id of its data_ptr
    BuiltIn kind = BuiltIn::VertexPosition;
};

struct BlockMember {
    uint32_t offset = 0;
    Type type;
};

struct InterfaceBlock {
    uint32_t pointer = 0;  // value id
    uint32_t size = 0;
    uint32_t alignment = 16;
    std::vector<BlockMember> members;
};

enum class ResourceKind : uint8_t { PushConstants, UniformBuffer, CombinedImageSampler, StorageBuffer };

struct Resource {
    ResourceKind kind;
    uint32_t set = 0, binding = 0;
    uint32_t pointer = 0;  // value id (data_ptr or texture)
};

struct Variable {
    Type type;
    uint32_t alignment = 4;
    uint32_t pointer = 0;
    uint32_t size = 0;  // bytes of storage
};

// A block of structured code. The reference IR nests blocks inside the
// instructions that own them (shader-compiler-ir/src/block.rs:76-184); here
// a function keeps them in one table and Block / Loop / BreakBlock /
// ContinueLoop / Switch refer to them by index. Block 0 is the function's
// body: breaking from it returns.
struct Block {
    std::string name;
    std::vector<Instruction> body;
    bool loop_body = false;  // the body of a Loop (its name is the loop's)
};

struct Function {
    std::string name;
    std::vector<Variable> locals;
    std::vector<Block> blocks{Block{}};
    Block &body() { return blocks[0]; }
    const Block &body() const { return blocks[0]; }
    // (a name already taken gets a suffix: names are unique in a function,
    // so the text form names each break and continue target unambiguously)
    uint32_t new_block(const std::string &name, bool loop_body = false);
};

enum class Stage : uint8_t { Vertex, Fragment, Compute };

struct Module {
    Stage stage = Stage::Vertex;
    std::vector<InterfaceVariable> built_in_inputs, built_in_outputs;
    InterfaceBlock user_inputs, user_outputs;
    std::vector<Variable> invocation_globals;
    std::vector<Variable> workgroup_variables;  // compute: shared by a workgroup's invocations
    std::vector<Resource> resources;
    uint32_t local_size[3] = {0, 0, 0};         // compute: the workgroup's size (not in the text form)
    std::vector<Value> values;  // index = id - 1
    Function main;

    std::set<std::string> value_names;

    Value &value(uint32_t id) { return values[id - 1]; }
    const Value &value(uint32_t id) const { return values[id - 1]; }
    // (a name already taken gets a suffix: value names are unique in a
    // module, so the text form names each value unambiguously; an empty
    // name becomes v<id>)
    uint32_t new_value(const Type &t, const std::string &name);
    // a constant: printed by its bits, it takes no name
    uint32_t new_const(const Type &t);
    uint32_t new_const_f32(float v);
    uint32_t new_const_i32(int32_t v);
};

// The module in the reference's textual form.
std::string module_text(const Module &m);
std::string instruction_text(const Module &m, const Instruction &ins);

// The text form parsed back (ir_text.cpp): a module, or value definitions
// "[name: type, ...]" followed by one instruction. False with a message
// (line number first) on malformed text.
bool parse_module_text(const std::string &text, Module &m, std::string &err);
bool parse_instruction_text(const std::string &text, Module &m, Instruction &ins, std::string &err);

// Names of the text form.
const char *op_name(Op op);
bool builtin_of(const std::string &s, BuiltIn &out);
bool fcmp_of(const std::string &s, uint32_t &out);
bool icmp_of(const std::string &s, uint32_t &out);
bool atomic_of(const std::string &s, uint32_t &out);

}  // namespace ir
}  // namespace kz
