// translate.cpp — SPIR-V -> shader IR for the draw path's shader subset.
#include "translate.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <set>

#include "spirv_id_map.h"
#include "spirv_reader.h"

namespace kz {

namespace {

using ir::Instruction;
using ir::Op;
using ir::Type;

enum SOp : uint16_t {
    OpNop = 0, OpSource = 3, OpSourceExtension = 4, OpName = 5, OpMemberName = 6, OpString = 7, OpLine = 8,
    OpExtension = 10, OpExtInstImport = 11, OpExtInst = 12, OpMemoryModel = 14, OpEntryPoint = 15,
    OpExecutionMode = 16, OpCapability = 17, OpTypeVoid = 19, OpTypeBool = 20, OpTypeInt = 21, OpTypeFloat = 22,
    OpTypeVector = 23, OpTypeMatrix = 24, OpTypeImage = 25, OpTypeSampler = 26, OpTypeSampledImage = 27,
    OpTypeArray = 28, OpTypeStruct = 30, OpTypePointer = 32, OpTypeFunction = 33, OpConstantTrue = 41,
    OpConstantFalse = 42, OpConstant = 43, OpConstantComposite = 44, OpSpecConstantTrue = 48,
    OpSpecConstantFalse = 49, OpSpecConstant = 50, OpFunction = 54, OpFunctionEnd = 56, OpVariable = 59,
    OpLoad = 61, OpStore = 62, OpAccessChain = 65, OpInBoundsAccessChain = 66, OpDecorate = 71,
    OpMemberDecorate = 72, OpVectorShuffle = 79, OpCompositeConstruct = 80, OpCompositeExtract = 81,
    OpImageSampleImplicitLod = 87, OpImageSampleExplicitLod = 88, OpConvertFToS = 110, OpConvertSToF = 111,
    OpFNegate = 127, OpIAdd = 128, OpFAdd = 129, OpFSub = 131, OpFMul = 133, OpFDiv = 136,
    OpVectorTimesScalar = 142, OpMatrixTimesVector = 145, OpDot = 148, OpSelect = 169, OpFOrdLessThan = 184,
    OpFOrdGreaterThan = 186, OpNoLine = 317, OpModuleProcessed = 330, OpSelectionMerge = 247, OpLoopMerge = 246,
    OpLabel = 248, OpBranch = 249, OpBranchConditional = 250, OpKill = 252, OpReturn = 253,
    OpUndef = 1, OpFunctionParameter = 55, OpFunctionCall = 57, OpCompositeInsert = 82, OpCopyObject = 83,
    OpConvertFToU = 109, OpConvertUToF = 112, OpBitcast = 124, OpSNegate = 126, OpISub = 130, OpIMul = 132,
    OpUDiv = 134, OpSDiv = 135, OpUMod = 137, OpSRem = 138, OpSMod = 139, OpFRem = 140, OpFMod = 141,
    OpAny = 154, OpAll = 155, OpIsNan = 156, OpIsInf = 157, OpLogicalEqual = 164, OpLogicalNotEqual = 165,
    OpLogicalOr = 166, OpLogicalAnd = 167, OpLogicalNot = 168, OpIEqual = 170, OpINotEqual = 171,
    OpUGreaterThan = 172, OpSGreaterThan = 173, OpUGreaterThanEqual = 174, OpSGreaterThanEqual = 175,
    OpULessThan = 176, OpSLessThan = 177, OpULessThanEqual = 178, OpSLessThanEqual = 179, OpFOrdEqual = 180,
    OpFUnordEqual = 181, OpFOrdNotEqual = 182, OpFUnordNotEqual = 183, OpFUnordLessThan = 185,
    OpFUnordGreaterThan = 187, OpFOrdLessThanEqual = 188, OpFUnordLessThanEqual = 189,
    OpFOrdGreaterThanEqual = 190, OpFUnordGreaterThanEqual = 191, OpShiftRightLogical = 194,
    OpShiftRightArithmetic = 195, OpShiftLeftLogical = 196, OpBitwiseOr = 197, OpBitwiseXor = 198,
    OpBitwiseAnd = 199, OpNot = 200, OpPhi = 245, OpSwitch = 251, OpReturnValue = 254, OpUnreachable = 255,
    OpTypeRuntimeArray = 29, OpControlBarrier = 224, OpMemoryBarrier = 225, OpAtomicLoad = 227, OpAtomicStore = 228,
    OpAtomicExchange = 229, OpAtomicIIncrement = 232, OpAtomicIDecrement = 233, OpAtomicIAdd = 234,
    OpAtomicISub = 235, OpAtomicSMin = 236, OpAtomicUMin = 237, OpAtomicSMax = 238, OpAtomicUMax = 239,
    OpAtomicAnd = 240, OpAtomicOr = 241, OpAtomicXor = 242, OpTranspose = 84, OpMatrixTimesScalar = 143,
    OpVectorTimesMatrix = 144, OpMatrixTimesMatrix = 146, OpOuterProduct = 147, OpBitFieldInsert = 201,
    OpBitFieldSExtract = 202, OpBitFieldUExtract = 203, OpBitReverse = 204, OpBitCount = 205, OpArrayLength = 68,
};

enum Storage : uint32_t {
    SC_UniformConstant = 0, SC_Input = 1, SC_Uniform = 2, SC_Output = 3, SC_Workgroup = 4, SC_Private = 6,
    SC_Function = 7, SC_PushConstant = 9, SC_StorageBuffer = 12,
};

enum Deco : uint32_t {
    D_SpecId = 1, D_Block = 2, D_BufferBlock = 3, D_ColMajor = 5, D_ArrayStride = 6, D_MatrixStride = 7, D_BuiltIn = 11, D_Flat = 14,
    D_Location = 30, D_Component = 31, D_Binding = 33, D_DescriptorSet = 34, D_Offset = 35,
};

enum BuiltInS : uint32_t { BI_Position = 0, BI_PointSize = 1, BI_FragCoord = 15, BI_NumWorkgroups = 24,
                           BI_WorkgroupSize = 25, BI_WorkgroupId = 26, BI_LocalInvocationId = 27,
                           BI_GlobalInvocationId = 28, BI_LocalInvocationIndex = 29, BI_VertexIndex = 42,
                           BI_InstanceIndex = 43 };

struct SType {
    enum K { Void, Bool, Int, Float, Vector, Matrix, Array, Struct, Pointer, Function, Image, SampledImage, Sampler };
    K k = Void;
    uint32_t width = 32, is_signed = 0;
    uint32_t elem = 0, count = 0;  // vector/matrix/array
    std::vector<uint32_t> members;
    uint32_t storage = 0, pointee = 0;
};

// A translated SPIR-V id.
struct Val {
    std::vector<uint32_t> ir;  // one IR value; four for a matrix (columns)
    uint32_t type = 0;         // SPIR-V type id of the value (pointee type for pointers)
    bool is_ptr = false;
    uint32_t storage = 0;
    uint32_t mat_stride = 16;  // pointers to matrices: column stride
    int layout = 0;            // pointers: 0 natural, 1 explicit (Offset/strides), 2 I/O locations
    std::string builtin_struct;  // pointer to a gl_PerVertex-style block
    uint32_t struct_var = 0;
};

// One basic block of a SPIR-V function: its instructions [begin, end) (the
// OpLabel excluded, the merge instruction included as a no-op) and its
// terminator at `end`.
struct CfgBlock {
    size_t begin = 0, end = 0;
    uint32_t merge = 0, cont = 0;  // OpSelectionMerge / OpLoopMerge operands
    bool loop = false, selection = false;
    bool emitted = false;
    // values stored into phi locals when control leaves this block:
    // (phi result id, incoming value id)
    std::vector<std::pair<uint32_t, uint32_t>> phi_out;
};

// The control-flow graph of one function being translated (the entry
// point, or a callee being inlined).
struct FnCfg {
    uint32_t fn = 0, ret_type = 0;
    std::vector<uint32_t> params;
    std::map<uint32_t, CfgBlock> blocks;
    std::vector<uint32_t> order;  // labels in declaration order
    uint32_t ret_block = 0;       // IR block whose break is a return
    uint32_t ret_local = 0;       // IR pointer receiving OpReturnValue (inlined calls)
};

// An enclosing structured construct: which branch targets are jumps out of
// or within it, and the IR blocks those jumps break from.
struct Construct {
    enum Kind { Selection, Loop, Switch } kind = Selection;
    uint32_t merge_label = 0, merge_block = 0;
    uint32_t cont_label = 0, cont_block = 0;    // loop: continue target and the block whose end reaches it
    uint32_t header_label = 0, loop_block = 0;  // loop: the back edge continues loop_block
    std::vector<std::pair<uint32_t, uint32_t>> cases;  // switch: case label -> block whose end reaches it
};

struct Translator {
    const SpirvModule &sm;
    ir::Module &m;
    std::string &err;
    // per-id tables, bounded by the module header's `bound`
    IdMap<SType> types;
    IdMap<std::string> names;
    IdMap<std::vector<std::vector<uint32_t>>> decos;  // id -> [deco, params...]
    std::map<std::pair<uint32_t, uint32_t>, std::vector<std::vector<uint32_t>>> mdecos;
    IdMap<Val> vals;
    IdMap<std::pair<uint32_t, std::vector<uint32_t>>> consts;  // id -> (type, words)
    std::map<uint32_t, std::vector<uint8_t>> spec;  // by SpecId
    uint32_t glsl_set = 0;
    std::string file_name;
    IdMap<std::string> strings;
    // builtin member map for block variables: (var id, member) -> IR pointer
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> block_builtin_ptr;
    ir::Location cur_loc;
    bool have_loc = false;
    // structured emission state
    uint32_t cur_block = 0;
    FnCfg *fcfg = nullptr;
    std::vector<Construct> constructs;
    size_t construct_base = 0;  // constructs of the function being emitted start here
    IdMap<uint32_t> phi_local;  // phi result id -> IR local pointer
    int inline_depth = 0;

    Translator(const SpirvModule &s, ir::Module &mm, std::string &e)
        : sm(s), m(mm), err(e), types(s.bound), names(s.bound), decos(s.bound), vals(s.bound), consts(s.bound),
          strings(s.bound), phi_local(s.bound) {}

    // an id outside [1, bound) that any table was asked to hold, 0 if none
    uint32_t out_of_bounds() const {
        for (uint32_t id : {types.out_of_bounds(), names.out_of_bounds(), decos.out_of_bounds(), vals.out_of_bounds(),
                            consts.out_of_bounds(), strings.out_of_bounds(), phi_local.out_of_bounds()})
            if (id) return id;
        return 0;
    }

    // (an id outside the header's bound, once seen, is the error to report:
    // what fails after it is its consequence)
    bool fail(const std::string &s) {
        if (err.empty()) {
            const uint32_t id = out_of_bounds();
            err = id ? "SPIR-V id %" + std::to_string(id) + " is out of bounds (header bound " +
                           std::to_string(sm.bound) + ")"
                     : s;
        }
        return false;
    }

    const std::vector<uint32_t> *deco(uint32_t id, uint32_t d) const {
        auto it = decos.find(id);
        if (it == decos.end()) return nullptr;
        for (const auto &v : it->second)
            if (v[0] == d) return &v;
        return nullptr;
    }
    const std::vector<uint32_t> *mdeco(uint32_t id, uint32_t mem, uint32_t d) const {
        auto it = mdecos.find({id, mem});
        if (it == mdecos.end()) return nullptr;
        for (const auto &v : it->second)
            if (v[0] == d) return &v;
        return nullptr;
    }

    // IR type of a SPIR-V scalar/vector type; Void for anything else.
    Type ir_type(uint32_t tid) const {
        auto it = types.find(tid);
        if (it == types.end()) return Type{};
        const SType &t = it->second;
        switch (t.k) {
        case SType::Bool: return Type::boolean();
        case SType::Int: return Type::i32();
        case SType::Float: return Type::f32();
        case SType::Vector: {
            Type e = ir_type(t.elem);
            e.lanes = (uint8_t)t.count;
            return e;
        }
        case SType::SampledImage: return Type::texture();
        default: return Type{};
        }
    }

    // Byte size of a type under a layout (0 natural, 1 explicit, 2 I/O).
    uint32_t size_of(uint32_t tid, int layout) const {
        const SType &t = types.at(tid);
        switch (t.k) {
        case SType::Bool:
        case SType::Int:
        case SType::Float: return 4;
        case SType::Vector: return layout == 2 ? 16 : 4 * t.count;
        case SType::Matrix: return t.count * (layout == 2 ? 16 : 16);
        case SType::Array: {
            uint32_t n = const_u32(t.count);
            uint32_t stride = array_stride(tid, layout);
            return n * stride;
        }
        case SType::Struct: {
            uint32_t end = 0;
            for (uint32_t i = 0; i < t.members.size(); ++i) {
                const uint32_t off = member_offset(tid, i, layout);
                end = std::max(end, off + size_of(t.members[i], layout));
            }
            return end;
        }
        default: return 0;
        }
    }
    uint32_t array_stride(uint32_t tid, int layout) const {
        const SType &t = types.at(tid);
        if (layout == 1) {
            if (auto d = deco(tid, D_ArrayStride)) return (*d)[1];
        }
        if (layout == 2) return 16 * ((size_of(t.elem, 2) + 15) / 16);
        const uint32_t s = size_of(t.elem, layout);
        return (s + 3) / 4 * 4;
    }
    uint32_t member_offset(uint32_t sid, uint32_t i, int layout) const {
        if (layout == 1) {
            if (auto d = mdeco(sid, i, D_Offset)) return (*d)[1];
        }
        const SType &t = types.at(sid);
        uint32_t off = 0;
        for (uint32_t k = 0; k < i; ++k) {
            uint32_t s = size_of(t.members[k], layout);
            if (layout == 2) s = (s + 15) / 16 * 16;
            off += s;
        }
        return off;
    }
    uint32_t matrix_stride(uint32_t struct_id, uint32_t member, int layout) const {
        if (layout == 1 && struct_id) {
            if (auto d = mdeco(struct_id, member, D_MatrixStride)) return (*d)[1];
        }
        return 16;
    }
    uint32_t const_u32(uint32_t id) const {
        auto it = consts.find(id);
        if (it == consts.end() || it->second.second.empty()) return 0;
        return it->second.second[0];
    }

    uint32_t emit(Op op, std::vector<uint32_t> args, const Type &rt, std::vector<uint32_t> fields = {},
                  const std::string &name = "") {
        Instruction ins;
        ins.op = op;
        ins.args = std::move(args);
        ins.fields = std::move(fields);
        uint32_t r = 0;
        if (rt.scalar != ir::Scalar::Void) {
            r = m.new_value(rt, name);
            ins.results.push_back(r);
        }
        if (have_loc) {
            ins.has_location = true;
            ins.loc = cur_loc;
        }
        m.main.blocks[cur_block].body.push_back(std::move(ins));
        return r;
    }
    void emit_jump(Op op, uint32_t block) { emit(op, {}, Type{}, {block}); }

    uint32_t ptr_offset(uint32_t base, uint32_t off) {
        if (off == 0) return base;
        return emit(Op::PtrOffset, {base}, Type::ptr(), {off});
    }

    std::string value_name(uint32_t sid) {
        auto it = names.find(sid);
        if (it != names.end() && !it->second.empty()) return it->second;
        return "id_" + std::to_string(sid);
    }

    bool collect();
    bool build_interface(uint32_t entry_fn, const std::vector<uint32_t> &iface, ir::Stage stage);
    bool translate_function(uint32_t fn);
    bool build_cfg(uint32_t fn, FnCfg &c);
    bool emit_from(uint32_t label, bool header_entry);
    bool emit_target(uint32_t label);
    int try_jump(uint32_t label);
    bool emit_if(uint32_t cond, uint32_t t, uint32_t f);
    bool emit_switch(const CfgBlock &b, const std::vector<uint32_t> &o);
    bool emit_phi_stores(const CfgBlock &b);
    bool body_ins(const SpirvInstruction &ins);
    bool call(const std::vector<uint32_t> &o);
    uint32_t local_var(const Type &t, uint32_t bytes, const std::string &name);
    const Val *binary_operands(const std::vector<uint32_t> &o, const Val *&b);
    bool load(uint32_t rtype, uint32_t rid, const Val &p);
    bool store(const Val &p, const Val &v);
    bool access_chain(uint32_t rid, uint32_t ptr_type, const Val &base, const std::vector<uint32_t> &idx);
    const Val *get(uint32_t id);
};

bool Translator::collect() {
    for (const auto &ins : sm.instructions) {
        const auto &o = ins.operands;
        // ids every table below is keyed by must lie in [1, bound)
        // (spirv-id-map/src/lib.rs:94-210: IdOutOfBounds)
        if ((ins.opcode == OpLabel || ins.opcode == OpName || ins.opcode == OpDecorate || ins.opcode == OpString) &&
            !o.empty() && !types.in_bounds(o[0]))
            return fail("SPIR-V id %" + std::to_string(o[0]) + " is out of bounds (header bound " +
                        std::to_string(sm.bound) + ")");
        if ((ins.opcode == OpFunction || ins.opcode == OpVariable) && o.size() > 1 && !types.in_bounds(o[1]))
            return fail("SPIR-V id %" + std::to_string(o[1]) + " is out of bounds (header bound " +
                        std::to_string(sm.bound) + ")");
        switch (ins.opcode) {
        case OpName: names[o[0]] = spirv_string(o, 1); break;
        case OpString: strings[o[0]] = spirv_string(o, 1); break;
        case OpDecorate: {
            std::vector<uint32_t> d(o.begin() + 1, o.end());
            decos[o[0]].push_back(d);
            break;
        }
        case OpMemberDecorate: {
            std::vector<uint32_t> d(o.begin() + 2, o.end());
            mdecos[{o[0], o[1]}].push_back(d);
            break;
        }
        case OpExtInstImport:
            if (spirv_string(o, 1) != "GLSL.std.450") return fail("unsupported extended instruction set");
            glsl_set = o[0];
            break;
        case OpTypeVoid: types[o[0]].k = SType::Void; break;
        case OpTypeBool: types[o[0]].k = SType::Bool; break;
        case OpTypeInt: {
            SType t;
            t.k = SType::Int;
            t.width = o[1];
            t.is_signed = o[2];
            if (t.width != 32) return fail("only 32-bit integer types are supported");
            types[o[0]] = t;
            break;
        }
        case OpTypeFloat: {
            SType t;
            t.k = SType::Float;
            t.width = o[1];
            if (t.width != 32) return fail("only 32-bit float types are supported");
            types[o[0]] = t;
            break;
        }
        case OpTypeVector:
        case OpTypeMatrix: {
            SType t;
            t.k = ins.opcode == OpTypeVector ? SType::Vector : SType::Matrix;
            t.elem = o[1];
            t.count = o[2];
            types[o[0]] = t;
            break;
        }
        case OpTypeArray:
        case OpTypeRuntimeArray: {
            SType t;
            t.k = SType::Array;
            t.elem = o[1];
            t.count = ins.opcode == OpTypeArray ? o[2] : 0;  // constant id; 0: sized at run time (no storage here)
            types[o[0]] = t;
            break;
        }
        case OpTypeStruct: {
            SType t;
            t.k = SType::Struct;
            t.members.assign(o.begin() + 1, o.end());
            types[o[0]] = t;
            break;
        }
        case OpTypePointer: {
            SType t;
            t.k = SType::Pointer;
            t.storage = o[1];
            t.pointee = o[2];
            types[o[0]] = t;
            break;
        }
        case OpTypeFunction: types[o[0]].k = SType::Function; break;
        case OpTypeImage: types[o[0]].k = SType::Image; break;
        case OpTypeSampler: types[o[0]].k = SType::Sampler; break;
        case OpTypeSampledImage: {
            SType t;
            t.k = SType::SampledImage;
            t.elem = o[1];
            types[o[0]] = t;
            break;
        }
        case OpConstant:
        case OpSpecConstant: {
            std::vector<uint32_t> w(o.begin() + 2, o.end());
            if (ins.opcode == OpSpecConstant) {
                if (auto d = deco(o[1], D_SpecId)) {
                    auto it = spec.find((*d)[1]);
                    if (it != spec.end() && it->second.size() >= 4) std::memcpy(&w[0], it->second.data(), 4);
                }
            }
            consts[o[1]] = {o[0], w};
            break;
        }
        case OpConstantTrue:
        case OpConstantFalse:
        case OpSpecConstantTrue:
        case OpSpecConstantFalse:
            consts[o[1]] = {o[0], {(uint32_t)(ins.opcode == OpConstantTrue || ins.opcode == OpSpecConstantTrue)}};
            break;
        case OpConstantComposite: {
            std::vector<uint32_t> w;
            for (size_t i = 2; i < o.size(); ++i) {
                auto it = consts.find(o[i]);
                if (it == consts.end()) return fail("composite constant of non-scalar constants is not supported");
                w.push_back(it->second.second[0]);
            }
            consts[o[1]] = {o[0], w};
            break;
        }
        default: break;
        }
    }
    return true;
}

const Val *Translator::get(uint32_t id) {
    auto it = vals.find(id);
    if (it != vals.end()) return &it->second;
    auto c = consts.find(id);
    if (c == consts.end()) {
        fail("use of undefined or unsupported id %" + std::to_string(id));
        return nullptr;
    }
    const uint32_t tid = c->second.first;
    const Type t = ir_type(tid);
    Val v;
    v.type = tid;
    if (t.scalar == ir::Scalar::Void) {
        fail("unsupported constant type");
        return nullptr;
    }
    const uint32_t vid = m.new_value(t, "");
    ir::Value &iv = m.value(vid);
    iv.is_const = true;
    for (size_t i = 0; i < c->second.second.size() && i < 4; ++i) iv.const_bits[i] = c->second.second[i];
    v.ir.push_back(vid);
    vals[id] = v;
    return &vals[id];
}

bool Translator::build_interface(uint32_t entry_fn, const std::vector<uint32_t> &iface, ir::Stage stage) {
    (void)entry_fn;
    m.user_inputs.pointer = m.new_value(Type::ptr(), "user_inputs_block");
    m.user_outputs.pointer = m.new_value(Type::ptr(), "user_outputs_block");
    m.user_inputs.alignment = m.user_outputs.alignment = 16;
    std::set<uint32_t> in_iface(iface.begin(), iface.end());

    // Walk every module-scope variable in declaration order.
    for (const auto &ins : sm.instructions) {
        if (ins.opcode == OpFunction) break;
        if (ins.opcode != OpVariable) continue;
        const uint32_t ptype = ins.operands[0], id = ins.operands[1], sc = ins.operands[2];
        const SType &pt = types.at(ptype);
        const uint32_t pointee = pt.pointee;
        Val v;
        v.is_ptr = true;
        v.type = pointee;
        v.storage = sc;
        if (sc == SC_Input || sc == SC_Output) {
            if (!in_iface.count(id)) continue;
            const bool out = sc == SC_Output;
            v.layout = 2;
            if (auto bi = deco(id, D_BuiltIn)) {
                const uint32_t b = (*bi)[1];
                ir::InterfaceVariable iv;
                iv.type = ir_type(pointee);
                iv.alignment = 4;
                if (b == BI_Position && out) {
                    iv.kind = ir::BuiltIn::VertexPosition;
                    iv.pointer = m.new_value(Type::ptr(), "vertex_position_out");
                    m.built_in_outputs.push_back(iv);
                } else if (b == BI_VertexIndex && !out) {
                    iv.kind = ir::BuiltIn::VertexIndex;
                    iv.pointer = m.new_value(Type::ptr(), "vertex_index_in");
                    m.built_in_inputs.push_back(iv);
                } else if (b == BI_InstanceIndex && !out) {
                    iv.kind = ir::BuiltIn::InstanceIndex;
                    iv.pointer = m.new_value(Type::ptr(), "instance_index_in");
                    m.built_in_inputs.push_back(iv);
                } else if (b == BI_FragCoord && !out) {
                    iv.kind = ir::BuiltIn::FragCoord;
                    iv.pointer = m.new_value(Type::ptr(), "fragment_coord_in");
                    m.built_in_inputs.push_back(iv);
                } else if (!out && stage == ir::Stage::Compute &&
                           (b == BI_GlobalInvocationId || b == BI_LocalInvocationId || b == BI_WorkgroupId ||
                            b == BI_NumWorkgroups || b == BI_LocalInvocationIndex)) {
                    static const std::map<uint32_t, std::pair<ir::BuiltIn, const char *>> kinds = {
                        {BI_GlobalInvocationId, {ir::BuiltIn::GlobalInvocationId, "global_invocation_id_in"}},
                        {BI_LocalInvocationId, {ir::BuiltIn::LocalInvocationId, "local_invocation_id_in"}},
                        {BI_WorkgroupId, {ir::BuiltIn::WorkgroupId, "workgroup_id_in"}},
                        {BI_NumWorkgroups, {ir::BuiltIn::NumWorkgroups, "num_workgroups_in"}},
                        {BI_LocalInvocationIndex, {ir::BuiltIn::LocalInvocationIndex, "local_invocation_index_in"}}};
                    iv.kind = kinds.at(b).first;
                    iv.pointer = m.new_value(Type::ptr(), kinds.at(b).second);
                    m.built_in_inputs.push_back(iv);
                } else {
                    return fail("unsupported built-in " + std::to_string(b));
                }
                v.ir.push_back(iv.pointer);
                v.layout = 0;
                vals[id] = v;
                continue;
            }
            const SType &st = types.at(pointee);
            if (st.k == SType::Struct && mdeco(pointee, 0, D_BuiltIn)) {
                // gl_PerVertex-style block: each built-in member is its own
                // interface variable.
                for (uint32_t i = 0; i < st.members.size(); ++i) {
                    auto bi = mdeco(pointee, i, D_BuiltIn);
                    if (!bi) return fail("block mixes built-in and user members");
                    const uint32_t b = (*bi)[1];
                    if (b == BI_Position && out) {
                        ir::InterfaceVariable iv;
                        iv.type = ir_type(st.members[i]);
                        iv.alignment = 4;
                        iv.kind = ir::BuiltIn::VertexPosition;
                        iv.pointer = m.new_value(Type::ptr(), "vertex_position_out");
                        m.built_in_outputs.push_back(iv);
                        block_builtin_ptr[{id, i}] = iv.pointer;
                    }
                    // other per-vertex built-ins (point size, clip distances)
                    // are written to a discarded slot
                }
                v.builtin_struct = "per_vertex";
                v.struct_var = id;
                vals[id] = v;
                continue;
            }
            auto loc = deco(id, D_Location);
            if (!loc) return fail("user interface variable without a Location");
            uint32_t comp = 0;
            if (auto c = deco(id, D_Component)) comp = (*c)[1];
            const uint32_t off = (*loc)[1] * 16 + comp * 4;
            ir::InterfaceBlock &blk = out ? m.user_outputs : m.user_inputs;
            // members are recorded per scalar/vector leaf
            std::vector<std::pair<uint32_t, uint32_t>> stack{{pointee, off}};
            while (!stack.empty()) {
                auto [t, o] = stack.back();
                stack.pop_back();
                const SType &tt = types.at(t);
                if (tt.k == SType::Array || tt.k == SType::Matrix) {
                    const uint32_t n = tt.k == SType::Array ? const_u32(tt.count) : tt.count;
                    for (uint32_t i = n; i-- > 0;) stack.push_back({tt.elem, o + 16 * i});
                } else if (tt.k == SType::Struct) {
                    return fail("struct-typed interface variables are not supported");
                } else {
                    ir::BlockMember bm;
                    bm.offset = o;
                    bm.type = ir_type(t);
                    blk.members.push_back(bm);
                    blk.size = std::max(blk.size, (o + 16) / 16 * 16);
                }
            }
            v.ir.push_back(ptr_offset(out ? m.user_outputs.pointer : m.user_inputs.pointer, off));
            vals[id] = v;
        } else if (sc == SC_PushConstant || sc == SC_Uniform || sc == SC_UniformConstant || sc == SC_StorageBuffer) {
            ir::Resource r;
            if (sc == SC_PushConstant) {
                r.kind = ir::ResourceKind::PushConstants;
                r.pointer = m.new_value(Type::ptr(), "push_constants");
            } else {
                auto set = deco(id, D_DescriptorSet), bind = deco(id, D_Binding);
                r.set = set ? (*set)[1] : 0;
                r.binding = bind ? (*bind)[1] : 0;
                // a storage buffer: StorageBuffer storage, or (SPIR-V 1.0
                // style) a Uniform block decorated BufferBlock
                if (sc == SC_StorageBuffer || (sc == SC_Uniform && deco(pointee, D_BufferBlock))) {
                    r.kind = ir::ResourceKind::StorageBuffer;
                    r.pointer = m.new_value(Type::ptr(), value_name(id));
                } else if (sc == SC_Uniform) {
                    r.kind = ir::ResourceKind::UniformBuffer;
                    r.pointer = m.new_value(Type::ptr(), value_name(id));
                } else {
                    if (types.at(pointee).k != SType::SampledImage)
                        return fail("only combined image samplers are supported in UniformConstant storage");
                    r.kind = ir::ResourceKind::CombinedImageSampler;
                    r.pointer = m.new_value(Type::texture(), value_name(id));
                }
            }
            m.resources.push_back(r);
            v.layout = 1;
            v.ir.push_back(r.pointer);
            vals[id] = v;
        } else if (sc == SC_Workgroup) {
            if (stage != ir::Stage::Compute) return fail("Workgroup storage outside a compute shader");
            ir::Variable wv;
            wv.size = size_of(pointee, 0);
            wv.type = ir_type(pointee);
            wv.alignment = 4;
            wv.pointer = m.new_value(Type::ptr(), value_name(id));
            m.workgroup_variables.push_back(wv);
            v.layout = 0;
            v.ir.push_back(wv.pointer);
            vals[id] = v;
        } else if (sc == SC_Private) {
            ir::Variable gv;
            gv.size = size_of(pointee, 0);
            gv.type = ir_type(pointee);
            gv.alignment = 4;
            gv.pointer = m.new_value(Type::ptr(), "global_var" + std::to_string(m.invocation_globals.size() + 1));
            m.invocation_globals.push_back(gv);
            v.layout = 0;
            v.ir.push_back(gv.pointer);
            vals[id] = v;
        } else {
            return fail("unsupported storage class " + std::to_string(sc));
        }
    }
    return true;
}

bool Translator::load(uint32_t rtype, uint32_t rid, const Val &p) {
    const SType &t = types.at(p.type);
    Val v;
    v.type = rtype;
    if (t.k == SType::SampledImage) {
        v.ir = p.ir;
        vals[rid] = v;
        return true;
    }
    if (t.k == SType::Matrix) {
        const Type col = ir_type(t.elem);
        for (uint32_t c = 0; c < t.count; ++c) {
            const uint32_t pc = ptr_offset(p.ir[0], c * p.mat_stride);
            v.ir.push_back(emit(Op::Load, {pc}, col, {4}, value_name(rid) + "_c" + std::to_string(c)));
        }
        vals[rid] = v;
        return true;
    }
    const Type rt = ir_type(rtype);
    if (rt.scalar == ir::Scalar::Void) return fail("load of an unsupported type");
    v.ir.push_back(emit(Op::Load, {p.ir[0]}, rt, {4}, value_name(rid)));
    vals[rid] = v;
    return true;
}

bool Translator::store(const Val &p, const Val &v) {
    const SType &t = types.at(p.type);
    if (t.k == SType::Matrix) {
        for (uint32_t c = 0; c < t.count; ++c) emit(Op::Store, {ptr_offset(p.ir[0], c * p.mat_stride), v.ir[c]}, Type{}, {4});
        return true;
    }
    if (ir_type(p.type).scalar == ir::Scalar::Void) return fail("store of an unsupported type");
    emit(Op::Store, {p.ir[0], v.ir[0]}, Type{}, {4});
    return true;
}

bool Translator::access_chain(uint32_t rid, uint32_t ptr_type, const Val &base, const std::vector<uint32_t> &idx) {
    Val r = base;
    r.type = types.at(ptr_type).pointee;
    uint32_t cur = base.type;
    size_t first = 0;
    uint32_t ptr = base.ir.empty() ? 0 : base.ir[0];
    if (!base.builtin_struct.empty()) {
        if (idx.empty()) return fail("access chain into a per-vertex block needs an index");
        const uint32_t mem = const_u32(idx[0]);
        auto it = block_builtin_ptr.find({base.struct_var, mem});
        if (it == block_builtin_ptr.end()) {
            // an unused per-vertex built-in: give it private storage
            ir::Variable lv;
            lv.type = ir_type(types.at(cur).members[mem]);
            lv.size = 16;
            lv.pointer = m.new_value(Type::ptr(), "discarded_builtin" + std::to_string(m.main.locals.size()));
            m.main.locals.push_back(lv);
            block_builtin_ptr[{base.struct_var, mem}] = lv.pointer;
            it = block_builtin_ptr.find({base.struct_var, mem});
        }
        ptr = it->second;
        cur = types.at(cur).members[mem];
        r.builtin_struct.clear();
        r.layout = 0;
        first = 1;
    }
    uint32_t off = 0;
    for (size_t i = first; i < idx.size(); ++i) {
        const SType &t = types.at(cur);
        if (!consts.count(idx[i])) {
            // a run-time index: the constant offset so far, then index * stride
            uint32_t stride = 0;
            if (t.k == SType::Array) stride = array_stride(cur, r.layout);
            else if (t.k == SType::Matrix) stride = r.mat_stride;
            else if (t.k == SType::Vector) stride = 4;
            else return fail("run-time index into a struct");
            const Val *iv = get(idx[i]);
            if (!iv) return false;
            if (m.value(iv->ir[0]).type != Type::i32()) return fail("access-chain index is not a 32-bit integer");
            ptr = emit(Op::PtrIndex, {ptr_offset(ptr, off), iv->ir[0]}, Type::ptr(), {stride});
            off = 0;
            cur = t.k == SType::Array || t.k == SType::Matrix || t.k == SType::Vector ? t.elem : cur;
            continue;
        }
        const uint32_t k = const_u32(idx[i]);
        switch (t.k) {
        case SType::Struct:
            if (k >= t.members.size()) return fail("struct index out of range");
            off += member_offset(cur, k, r.layout);
            r.mat_stride = matrix_stride(cur, k, r.layout);
            cur = t.members[k];
            break;
        case SType::Array:
            off += k * array_stride(cur, r.layout);
            cur = t.elem;
            break;
        case SType::Matrix:
            off += k * r.mat_stride;
            cur = t.elem;
            break;
        case SType::Vector:
            off += 4 * k;
            cur = t.elem;
            break;
        default: return fail("access chain through a non-composite type");
        }
    }
    r.ir = {ptr_offset(ptr, off)};
    r.type = cur;
    vals[rid] = r;
    return true;
}

bool Translator::build_cfg(uint32_t fn, FnCfg &c) {
    c.fn = fn;
    const auto &ins = sm.instructions;
    size_t i = 0;
    while (i < ins.size() && !(ins[i].opcode == OpFunction && ins[i].operands.size() > 1 && ins[i].operands[1] == fn))
        ++i;
    if (i == ins.size()) return fail("function %" + std::to_string(fn) + " not found");
    c.ret_type = ins[i].operands[0];
    CfgBlock *cur = nullptr;
    for (++i; i < ins.size() && ins[i].opcode != OpFunctionEnd; ++i) {
        const auto &o = ins[i].operands;
        switch (ins[i].opcode) {
        case OpFunctionParameter:
            if (!c.blocks.empty()) return fail("OpFunctionParameter after the first block");
            c.params.push_back(o[1]);
            break;
        case OpLabel:
            if (cur) return fail("block %" + std::to_string(o[0]) + " starts before the previous one ended");
            if (c.blocks.count(o[0])) return fail("duplicate label");
            cur = &c.blocks[o[0]];
            cur->begin = i + 1;
            c.order.push_back(o[0]);
            break;
        case OpSelectionMerge:
            if (!cur) return fail("instruction outside a block");
            cur->selection = true;
            cur->merge = o[0];
            break;
        case OpLoopMerge:
            if (!cur) return fail("instruction outside a block");
            cur->loop = true;
            cur->merge = o[0];
            cur->cont = o[1];
            break;
        case OpBranch:
        case OpBranchConditional:
        case OpSwitch:
        case OpReturn:
        case OpReturnValue:
        case OpKill:
        case OpUnreachable:
            if (!cur) return fail("terminator outside a block");
            cur->end = i;
            cur = nullptr;
            break;
        default:
            if (!cur && ins[i].opcode != OpLine && ins[i].opcode != OpNoLine)
                return fail("instruction outside a block");
            break;
        }
    }
    if (i == ins.size()) return fail("function without OpFunctionEnd");
    if (cur) return fail("last block without a terminator");
    if (c.order.empty()) return fail("function without blocks");
    // phis become function locals: stored on the incoming edges, loaded
    // where the phi stands
    for (auto &kv : c.blocks) {
        for (size_t k = kv.second.begin; k < kv.second.end; ++k) {
            const auto &pi = ins[k];
            if (pi.opcode != OpPhi) continue;
            const uint32_t rt = pi.operands[0], rid = pi.operands[1];
            const Type t = ir_type(rt);
            if (t.scalar == ir::Scalar::Void || t.scalar == ir::Scalar::Texture || t.scalar == ir::Scalar::DataPtr)
                return fail("OpPhi of an unsupported type");
            phi_local[rid] = local_var(t, t.size_bytes(), value_name(rid) + "_phi");
            for (size_t a = 2; a + 1 < pi.operands.size(); a += 2) {
                auto pb = c.blocks.find(pi.operands[a + 1]);
                if (pb == c.blocks.end()) return fail("OpPhi names an unknown predecessor");
                pb->second.phi_out.push_back({rid, pi.operands[a]});
            }
        }
    }
    return true;
}

uint32_t Translator::local_var(const Type &t, uint32_t bytes, const std::string &name) {
    ir::Variable lv;
    lv.type = t;
    lv.size = bytes;
    lv.alignment = 4;
    lv.pointer = m.new_value(Type::ptr(), name);
    m.main.locals.push_back(lv);
    return lv.pointer;
}

bool Translator::emit_phi_stores(const CfgBlock &b) {
    for (const auto &ps : b.phi_out) {
        const Val *v = get(ps.second);
        if (!v) return false;
        emit(Op::Store, {phi_local.at(ps.first), v->ir[0]}, Type{}, {4});
    }
    return true;
}

// 1: `label` is a jump out of (or within) an enclosing construct and the
// jump was emitted; 0: it is ordinary code to emit in place.
int Translator::try_jump(uint32_t label) {
    for (size_t i = constructs.size(); i-- > construct_base;) {
        const Construct &c = constructs[i];
        if (label == c.merge_label) {
            emit_jump(Op::BreakBlock, c.merge_block);
            return 1;
        }
        if (c.kind == Construct::Switch)
            for (const auto &cs : c.cases)
                if (cs.first == label) {
                    emit_jump(Op::BreakBlock, cs.second);
                    return 1;
                }
        if (c.kind == Construct::Loop) {
            if (c.cont_block && label == c.cont_label) {
                emit_jump(Op::BreakBlock, c.cont_block);
                return 1;
            }
            if (label == c.header_label) {
                emit_jump(Op::ContinueLoop, c.loop_block);
                return 1;
            }
        }
    }
    return 0;
}

bool Translator::emit_target(uint32_t label) {
    if (try_jump(label)) return true;
    return emit_from(label, false);
}

// if (cond) t else f, as the reference IR writes it: nested blocks and a
// branch that breaks from the chosen one
//   block f_arm { block t_arm { branch cond {1 -> t_arm, default -> f_arm} } <t> } <f>
bool Translator::emit_if(uint32_t cond, uint32_t t, uint32_t f) {
    const uint32_t fb = m.main.new_block(value_name(f)), tb = m.main.new_block(value_name(t));
    const uint32_t outer = cur_block;
    emit_jump(Op::Block, fb);
    cur_block = fb;
    emit_jump(Op::Block, tb);
    cur_block = tb;
    emit(Op::Switch, {cond}, Type{}, {fb, 1u, tb});
    cur_block = fb;
    if (!emit_target(t)) return false;
    cur_block = outer;
    return emit_target(f);
}

// OpSwitch: the case targets in block order, each case's code after the end
// of its own block, nested so that falling off one case's code reaches the
// next one's (fallthrough), with the dispatch innermost.
bool Translator::emit_switch(const CfgBlock &b, const std::vector<uint32_t> &o) {
    const Val *sel = get(o[0]);
    if (!sel) return false;
    if (m.value(sel->ir[0]).type != Type::i32()) return fail("OpSwitch selector is not a 32-bit integer");
    std::vector<uint32_t> targets;
    auto add = [&](uint32_t l) {
        if (l != b.merge && std::find(targets.begin(), targets.end(), l) == targets.end()) targets.push_back(l);
    };
    add(o[1]);
    for (size_t i = 2; i + 1 < o.size(); i += 2) add(o[i + 1]);
    auto pos = [&](uint32_t l) {
        return std::find(fcfg->order.begin(), fcfg->order.end(), l) - fcfg->order.begin();
    };
    std::sort(targets.begin(), targets.end(), [&](uint32_t x, uint32_t y) { return pos(x) < pos(y); });
    // A case that falls through to another must be laid out right before
    // it. The case construct of target t is what t reaches without passing
    // the merge or another target; the first other target it reaches is the
    // one it falls through to.
    std::map<uint32_t, uint32_t> falls_to;
    for (uint32_t t : targets) {
        std::vector<uint32_t> stack{t};
        std::set<uint32_t> seen{t};
        while (!stack.empty() && !falls_to.count(t)) {
            const uint32_t l = stack.back();
            stack.pop_back();
            auto bi = fcfg->blocks.find(l);
            if (bi == fcfg->blocks.end()) continue;
            const SpirvInstruction &term = sm.instructions[bi->second.end];
            std::vector<uint32_t> succ;
            if (term.opcode == OpBranch) succ = {term.operands[0]};
            else if (term.opcode == OpBranchConditional) succ = {term.operands[1], term.operands[2]};
            else if (term.opcode == OpSwitch)
                for (size_t i = 1; i < term.operands.size(); i += 2) succ.push_back(term.operands[i]);
            if (bi->second.merge) succ.push_back(bi->second.merge);
            for (uint32_t n : succ) {
                if (n == b.merge || seen.count(n)) continue;
                if (std::find(targets.begin(), targets.end(), n) != targets.end()) {
                    falls_to[t] = n;
                    break;
                }
                seen.insert(n);
                stack.push_back(n);
            }
        }
    }
    {
        std::vector<uint32_t> chained;
        std::set<uint32_t> has_pred;
        for (const auto &kv : falls_to) has_pred.insert(kv.second);
        for (uint32_t t : targets) {
            if (has_pred.count(t)) continue;
            for (uint32_t u = t; u && std::find(chained.begin(), chained.end(), u) == chained.end();) {
                chained.push_back(u);
                auto it = falls_to.find(u);
                u = it == falls_to.end() ? 0 : it->second;
            }
        }
        if (chained.size() != targets.size()) return fail("switch cases fall through in a cycle");
        targets.swap(chained);
    }
    const uint32_t mb = m.main.new_block(value_name(b.merge));
    const uint32_t outer = cur_block;
    emit_jump(Op::Block, mb);
    cur_block = mb;
    std::vector<uint32_t> blk(targets.size());
    for (size_t k = 0; k < targets.size(); ++k) blk[k] = m.main.new_block(value_name(targets[k]));
    for (size_t k = targets.size(); k-- > 0;) {
        emit_jump(Op::Block, blk[k]);
        cur_block = blk[k];
    }
    auto block_of = [&](uint32_t l) {
        for (size_t k = 0; k < targets.size(); ++k)
            if (targets[k] == l) return blk[k];
        return mb;
    };
    std::vector<uint32_t> fields{block_of(o[1])};
    for (size_t i = 2; i + 1 < o.size(); i += 2) {
        fields.push_back(o[i]);
        fields.push_back(block_of(o[i + 1]));
    }
    emit(Op::Switch, {sel->ir[0]}, Type{}, fields);
    Construct c;
    c.kind = Construct::Switch;
    c.merge_label = b.merge;
    c.merge_block = mb;
    for (size_t k = 0; k < targets.size(); ++k) c.cases.push_back({targets[k], blk[k]});
    constructs.push_back(c);
    for (size_t k = 0; k < targets.size(); ++k) {
        cur_block = k + 1 < targets.size() ? blk[k + 1] : mb;
        if (!emit_from(targets[k], false)) return false;
    }
    constructs.pop_back();
    cur_block = outer;
    return true;
}

bool Translator::emit_from(uint32_t label, bool header_entry) {
    for (;;) {
        auto it = fcfg->blocks.find(label);
        if (it == fcfg->blocks.end()) return fail("branch to an unknown block %" + std::to_string(label));
        CfgBlock &b = it->second;
        if (b.loop && !header_entry) {
            // block merge { loop header { block cont { <header .. body> } <continue construct> } }
            if (b.emitted) return fail("unstructured control flow: loop header reached twice");
            const uint32_t mb = m.main.new_block(value_name(b.merge));
            const uint32_t lb = m.main.new_block(value_name(label), true);
            const uint32_t cb = b.cont != label ? m.main.new_block(value_name(b.cont)) : 0;
            const uint32_t outer = cur_block;
            emit_jump(Op::Block, mb);
            cur_block = mb;
            emit_jump(Op::Loop, lb);
            cur_block = lb;
            if (cb) {
                emit_jump(Op::Block, cb);
                cur_block = cb;
            }
            Construct c;
            c.kind = Construct::Loop;
            c.merge_label = b.merge;
            c.merge_block = mb;
            c.cont_label = b.cont;
            c.cont_block = cb;
            c.header_label = label;
            c.loop_block = lb;
            constructs.push_back(c);
            if (!emit_from(label, true)) return false;
            if (cb) {
                // the continue construct: runs after the body's continues,
                // ends with the back edge (or a break)
                cur_block = lb;
                constructs.back().cont_block = 0;
                if (!emit_from(b.cont, false)) return false;
            }
            constructs.pop_back();
            cur_block = outer;
            if (try_jump(b.merge)) return true;
            label = b.merge;
            continue;
        }
        header_entry = false;
        if (b.emitted) return fail("unstructured control flow: block %" + std::to_string(label) + " is reached twice");
        b.emitted = true;
        for (size_t i = b.begin; i < b.end; ++i)
            if (!body_ins(sm.instructions[i])) return false;
        const SpirvInstruction &t = sm.instructions[b.end];
        const auto &o = t.operands;
        switch (t.opcode) {
        case OpReturn:
            emit_jump(Op::BreakBlock, fcfg->ret_block);
            return true;
        case OpReturnValue: {
            const Val *v = get(o[0]);
            if (!v) return false;
            if (!fcfg->ret_local) return fail("OpReturnValue in a function without a result");
            emit(Op::Store, {fcfg->ret_local, v->ir[0]}, Type{}, {4});
            emit_jump(Op::BreakBlock, fcfg->ret_block);
            return true;
        }
        case OpKill:
            if (m.stage != ir::Stage::Fragment) return fail("OpKill outside a fragment shader");
            emit(Op::Kill, {}, Type{});
            return true;
        case OpUnreachable:
            emit(Op::Unreachable, {}, Type{});
            return true;
        case OpBranch:
            if (!emit_phi_stores(b)) return false;
            if (try_jump(o[0])) return true;
            label = o[0];
            continue;
        case OpBranchConditional: {
            if (!emit_phi_stores(b)) return false;
            const Val *c = get(o[0]);
            if (!c) return false;
            if (m.value(c->ir[0]).type != Type::boolean()) return fail("branch condition is not a bool");
            if (!b.selection) return emit_if(c->ir[0], o[1], o[2]);  // a loop's exit test, a conditional break
            const uint32_t mb = m.main.new_block(value_name(b.merge));
            const uint32_t outer = cur_block;
            emit_jump(Op::Block, mb);
            cur_block = mb;
            Construct sc;
            sc.kind = Construct::Selection;
            sc.merge_label = b.merge;
            sc.merge_block = mb;
            constructs.push_back(sc);
            if (!emit_if(c->ir[0], o[1], o[2])) return false;
            constructs.pop_back();
            cur_block = outer;
            if (try_jump(b.merge)) return true;
            label = b.merge;
            continue;
        }
        case OpSwitch:
            if (!emit_phi_stores(b)) return false;
            if (!b.selection) return fail("OpSwitch without OpSelectionMerge");
            if (!emit_switch(b, o)) return false;
            if (try_jump(b.merge)) return true;
            label = b.merge;
            continue;
        default: return fail("unsupported block terminator " + std::to_string(t.opcode));
        }
    }
}

// OpFunctionCall: the callee inlined into a block that its returns break from
bool Translator::call(const std::vector<uint32_t> &o) {
    if (++inline_depth > 32) return fail("function calls nested too deeply (recursion?)");
    FnCfg cc;
    if (!build_cfg(o[2], cc)) return false;
    if (cc.params.size() != o.size() - 3) return fail("OpFunctionCall argument count differs from the callee's");
    std::set<uint32_t> before;
    for (const auto &kv : vals) before.insert(kv.first);
    for (size_t i = 0; i < cc.params.size(); ++i) {
        const Val *a = get(o[3 + i]);
        if (!a) return false;
        vals[cc.params[i]] = *a;
    }
    const SType &rt = types.at(o[0]);
    Type rty;
    if (rt.k != SType::Void) {
        rty = ir_type(o[0]);
        if (rty.scalar == ir::Scalar::Void) return fail("function result of an unsupported type");
        cc.ret_local = local_var(rty, rty.size_bytes(), value_name(o[1]) + "_ret");
    }
    cc.ret_block = m.main.new_block("inline_" + value_name(o[2]));
    const uint32_t outer = cur_block;
    emit_jump(Op::Block, cc.ret_block);
    cur_block = cc.ret_block;
    FnCfg *saved = fcfg;
    const size_t saved_base = construct_base;
    fcfg = &cc;
    construct_base = constructs.size();
    const bool ok = emit_from(cc.order[0], false);
    fcfg = saved;
    construct_base = saved_base;
    cur_block = outer;
    if (!ok) return false;
    // the callee's ids may be defined again by another call
    std::vector<uint32_t> callee_ids;
    for (const auto &kv : vals)
        if (!before.count(kv.first)) callee_ids.push_back(kv.first);
    for (uint32_t id : callee_ids) vals.erase(id);
    if (rt.k != SType::Void) {
        Val r;
        r.type = o[0];
        r.ir.push_back(emit(Op::Load, {cc.ret_local}, rty, {4}, value_name(o[1])));
        vals[o[1]] = r;
    }
    --inline_depth;
    return true;
}

bool Translator::translate_function(uint32_t fn) {
    m.main.name = "main";
    FnCfg c;
    if (!build_cfg(fn, c)) return false;
    m.main.body().name = value_name(c.order[0]);
    c.ret_block = 0;
    fcfg = &c;
    cur_block = 0;
    const bool ok = emit_from(c.order[0], false);
    fcfg = nullptr;
    return ok;
}

const Val *Translator::binary_operands(const std::vector<uint32_t> &o, const Val *&b) {
    const Val *a = get(o[2]);
    b = o.size() > 3 ? get(o[3]) : nullptr;
    return a;
}

bool Translator::body_ins(const SpirvInstruction &ins) {
    const auto &o = ins.operands;
    auto V = [&](uint32_t id) -> const Val * { return get(id); };
    auto define = [&](uint32_t rtype, uint32_t rid, uint32_t ir_id) {
        Val r;
        r.type = rtype;
        r.ir.push_back(ir_id);
        vals[rid] = r;
        return true;
    };
    switch (ins.opcode) {
    case OpLine:
        cur_loc.file = strings.count(o[0]) ? strings[o[0]] : value_name(o[0]);
        cur_loc.line = o[1];
        cur_loc.column = o[2];
        have_loc = true;
        return true;
    case OpNoLine: have_loc = false; return true;
    case OpNop:
    case OpSelectionMerge:
    case OpLoopMerge:
        return true;
    case OpPhi: {
        const Type t = ir_type(o[0]);
        return define(o[0], o[1], emit(Op::Load, {phi_local.at(o[1])}, t, {4}, value_name(o[1])));
    }
    case OpControlBarrier:
        if (m.stage != ir::Stage::Compute) return fail("OpControlBarrier outside a compute shader");
        emit(Op::Barrier, {}, Type{});
        return true;
    case OpMemoryBarrier:  // (one workgroup runs on one CU: its memory is ordered at the control barrier)
        return true;
    case OpAtomicLoad: {
        const Val *p = V(o[2]);
        return p && load(o[0], o[1], *p);
    }
    case OpAtomicStore: {
        const Val *p = V(o[0]);
        const Val *v = V(o[3]);
        return p && v && store(*p, *v);
    }
    case OpAtomicExchange:
    case OpAtomicIIncrement:
    case OpAtomicIDecrement:
    case OpAtomicIAdd:
    case OpAtomicISub:
    case OpAtomicSMin:
    case OpAtomicUMin:
    case OpAtomicSMax:
    case OpAtomicUMax:
    case OpAtomicAnd:
    case OpAtomicOr:
    case OpAtomicXor: {
        const Val *p = V(o[2]);
        if (!p) return false;
        const Type t = ir_type(o[0]);
        if (t != Type::i32()) return fail("atomics on 32-bit integers only");
        uint32_t value = 0;
        if (ins.opcode == OpAtomicIIncrement || ins.opcode == OpAtomicIDecrement) {
            value = m.new_const_i32(1);
        } else {
            const Val *v = V(o[5]);
            if (!v) return false;
            value = v->ir[0];
        }
        static const std::map<uint32_t, uint32_t> kind = {
            {OpAtomicExchange, ir::AExchange}, {OpAtomicIIncrement, ir::AAdd}, {OpAtomicIDecrement, ir::ASub},
            {OpAtomicIAdd, ir::AAdd}, {OpAtomicISub, ir::ASub}, {OpAtomicSMin, ir::ASMin}, {OpAtomicUMin, ir::AUMin},
            {OpAtomicSMax, ir::ASMax}, {OpAtomicUMax, ir::AUMax}, {OpAtomicAnd, ir::AAnd}, {OpAtomicOr, ir::AOr},
            {OpAtomicXor, ir::AXor}};
        return define(o[0], o[1], emit(Op::Atomic, {p->ir[0], value}, t, {kind.at(ins.opcode)}, value_name(o[1])));
    }
    case OpFunctionCall: return call(o);
    case OpCopyObject: {
        const Val *a = V(o[2]);
        if (!a) return false;
        vals[o[1]] = *a;
        return true;
    }
    case OpUndef: {
        const Type t = ir_type(o[0]);
        if (t.scalar == ir::Scalar::Void || t.scalar == ir::Scalar::Texture) return fail("OpUndef of an unsupported type");
        const uint32_t vid = m.new_value(t, "");
        m.value(vid).is_const = true;  // an undefined value may be any value: zero
        return define(o[0], o[1], vid);
    }
    case OpCompositeInsert: {
        const Val *obj = V(o[2]);
        const Val *comp = V(o[3]);
        if (!obj || !comp) return false;
        const SType &ct = types.at(o[0]);
        if (ct.k != SType::Vector || o.size() != 5) return fail("composite insert into an unsupported type");
        const Type vt = ir_type(o[0]);
        std::vector<uint32_t> parts;
        for (uint32_t l = 0; l < vt.lanes; ++l)
            parts.push_back(l == o[4] ? comp->ir[0] : emit(Op::Extract, {obj->ir[0]}, Type{vt.scalar, 1}, {l}));
        return define(o[0], o[1], emit(Op::MakeVector, parts, vt, {}, value_name(o[1])));
    }
    case OpSNegate:
    case OpNot:
    case OpLogicalNot:
    case OpConvertUToF:
    case OpConvertFToU:
    case OpBitcast:
    case OpAny:
    case OpAll:
    case OpIsNan:
    case OpIsInf:
    case OpBitReverse:
    case OpBitCount: {
        const Val *a = V(o[2]);
        if (!a) return false;
        Op op;
        switch (ins.opcode) {
        case OpBitReverse: op = Op::BitReverse; break;
        case OpBitCount: op = Op::BitCount; break;
        case OpSNegate: op = Op::INeg; break;
        case OpNot:
        case OpLogicalNot: op = Op::Not; break;
        case OpConvertUToF: op = Op::ConvUToF; break;
        case OpConvertFToU: op = Op::ConvFToU; break;
        case OpBitcast: op = Op::Bitcast; break;
        case OpAny: op = Op::Any; break;
        case OpAll: op = Op::All; break;
        case OpIsNan: op = Op::IsNan; break;
        default: op = Op::IsInf; break;
        }
        const Type rt = ir_type(o[0]);
        if (rt.scalar == ir::Scalar::Void) return fail("unsupported result type");
        if (op == Op::Bitcast && rt.lanes != m.value(a->ir[0]).type.lanes) return fail("bitcast between vector sizes");
        return define(o[0], o[1], emit(op, {a->ir[0]}, rt, {}, value_name(o[1])));
    }
    case OpISub:
    case OpIMul:
    case OpUDiv:
    case OpSDiv:
    case OpUMod:
    case OpSRem:
    case OpSMod:
    case OpFRem:
    case OpFMod:
    case OpShiftLeftLogical:
    case OpShiftRightLogical:
    case OpShiftRightArithmetic:
    case OpBitwiseAnd:
    case OpBitwiseOr:
    case OpBitwiseXor:
    case OpLogicalAnd:
    case OpLogicalOr: {
        const Val *b = nullptr;
        const Val *a = binary_operands(o, b);
        if (!a || !b) return false;
        Op op;
        switch (ins.opcode) {
        case OpISub: op = Op::ISub; break;
        case OpIMul: op = Op::IMul; break;
        case OpUDiv: op = Op::UDiv; break;
        case OpSDiv: op = Op::SDiv; break;
        case OpUMod: op = Op::UMod; break;
        case OpSRem: op = Op::SRem; break;
        case OpSMod: op = Op::SMod; break;
        case OpFRem: op = Op::FRem; break;
        case OpFMod: op = Op::FMod; break;
        case OpShiftLeftLogical: op = Op::Shl; break;
        case OpShiftRightLogical: op = Op::ShrL; break;
        case OpShiftRightArithmetic: op = Op::ShrA; break;
        case OpBitwiseAnd:
        case OpLogicalAnd: op = Op::And; break;
        case OpBitwiseOr:
        case OpLogicalOr: op = Op::Or; break;
        default: op = Op::Xor; break;
        }
        return define(o[0], o[1], emit(op, {a->ir[0], b->ir[0]}, ir_type(o[0]), {}, value_name(o[1])));
    }
    case OpIEqual:
    case OpINotEqual:
    case OpLogicalEqual:
    case OpLogicalNotEqual:
    case OpUGreaterThan:
    case OpSGreaterThan:
    case OpUGreaterThanEqual:
    case OpSGreaterThanEqual:
    case OpULessThan:
    case OpSLessThan:
    case OpULessThanEqual:
    case OpSLessThanEqual:
    case OpFOrdEqual:
    case OpFUnordEqual:
    case OpFOrdNotEqual:
    case OpFUnordNotEqual:
    case OpFUnordLessThan:
    case OpFUnordGreaterThan:
    case OpFOrdLessThanEqual:
    case OpFUnordLessThanEqual:
    case OpFOrdGreaterThanEqual:
    case OpFUnordGreaterThanEqual: {
        const Val *b = nullptr;
        const Val *a = binary_operands(o, b);
        if (!a || !b) return false;
        Op op = Op::ICmp;
        uint32_t pred = 0;
        switch (ins.opcode) {
        case OpIEqual:
        case OpLogicalEqual: pred = ir::IEq; break;
        case OpINotEqual:
        case OpLogicalNotEqual: pred = ir::INe; break;
        case OpUGreaterThan: pred = ir::IUGt; break;
        case OpSGreaterThan: pred = ir::ISGt; break;
        case OpUGreaterThanEqual: pred = ir::IUGe; break;
        case OpSGreaterThanEqual: pred = ir::ISGe; break;
        case OpULessThan: pred = ir::IULt; break;
        case OpSLessThan: pred = ir::ISLt; break;
        case OpULessThanEqual: pred = ir::IULe; break;
        case OpSLessThanEqual: pred = ir::ISLe; break;
        default:
            op = Op::FCmp;
            switch (ins.opcode) {
            case OpFOrdEqual: pred = ir::FOEq; break;
            case OpFUnordEqual: pred = ir::FUEq; break;
            case OpFOrdNotEqual: pred = ir::FONe; break;
            case OpFUnordNotEqual: pred = ir::FUNe; break;
            case OpFUnordLessThan: pred = ir::FULt; break;
            case OpFUnordGreaterThan: pred = ir::FUGt; break;
            case OpFOrdLessThanEqual: pred = ir::FOLe; break;
            case OpFUnordLessThanEqual: pred = ir::FULe; break;
            case OpFOrdGreaterThanEqual: pred = ir::FOGe; break;
            default: pred = ir::FUGe; break;
            }
        }
        return define(o[0], o[1], emit(op, {a->ir[0], b->ir[0]}, ir_type(o[0]), {pred}, value_name(o[1])));
    }
        case OpVariable: {
            if (o[2] != SC_Function) return fail("non-function variable inside a function");
            const uint32_t pointee = types.at(o[0]).pointee;
            ir::Variable lv;
            lv.type = ir_type(pointee);
            lv.size = size_of(pointee, 0);
            lv.alignment = 4;
            lv.pointer = m.new_value(Type::ptr(), value_name(o[1]));
            m.main.locals.push_back(lv);
            Val v;
            v.is_ptr = true;
            v.type = pointee;
            v.storage = SC_Function;
            v.ir.push_back(lv.pointer);
            vals[o[1]] = v;
            if (o.size() > 3) {
                const Val *init = V(o[3]);
                if (!init || !store(v, *init)) return false;
            }
            break;
        }
        case OpLoad: {
            const Val *p = V(o[2]);
            if (!p || !load(o[0], o[1], *p)) return false;
            break;
        }
        case OpStore: {
            const Val *p = V(o[0]);
            const Val *v = V(o[1]);
            if (!p || !v || !store(*p, *v)) return false;
            break;
        }
        case OpAccessChain:
        case OpInBoundsAccessChain: {
            const Val *b = V(o[2]);
            if (!b) return false;
            std::vector<uint32_t> idx(o.begin() + 3, o.end());
            if (!access_chain(o[1], o[0], *b, idx)) return false;
            break;
        }
        case OpCompositeExtract: {
            const Val *c = V(o[2]);
            if (!c) return false;
            const SType &ct = types.at(c->type);
            Val r;
            r.type = o[0];
            if (ct.k == SType::Matrix) {
                const uint32_t col = c->ir.at(o[3]);
                if (o.size() == 4) {
                    r.ir.push_back(col);
                } else {
                    r.ir.push_back(emit(Op::Extract, {col}, ir_type(o[0]), {o[4]}, value_name(o[1])));
                }
            } else if (ct.k == SType::Struct && c->ir.size() == ct.members.size()) {
                // a struct value (ModfStruct, FrexpStruct) is its members
                const uint32_t mem = c->ir.at(o[3]);
                if (o.size() == 4) {
                    r.ir.push_back(mem);
                } else {
                    if (o.size() != 5) return fail("nested extract from a struct value");
                    r.ir.push_back(emit(Op::Extract, {mem}, ir_type(o[0]), {o[4]}, value_name(o[1])));
                }
            } else if (ct.k == SType::Vector) {
                if (o.size() != 4) return fail("nested extract from a vector");
                r.ir.push_back(emit(Op::Extract, {c->ir[0]}, ir_type(o[0]), {o[3]}, value_name(o[1])));
            } else {
                return fail("composite extract from an unsupported type");
            }
            vals[o[1]] = r;
            break;
        }
        case OpCompositeConstruct: {
            const SType &rt = types.at(o[0]);
            Val r;
            r.type = o[0];
            if (rt.k == SType::Vector) {
                std::vector<uint32_t> parts;
                for (size_t i = 2; i < o.size(); ++i) {
                    const Val *p = V(o[i]);
                    if (!p) return false;
                    const Type pt = m.value(p->ir[0]).type;
                    if (pt.lanes == 1) {
                        parts.push_back(p->ir[0]);
                    } else {
                        for (int l = 0; l < pt.lanes; ++l)
                            parts.push_back(emit(Op::Extract, {p->ir[0]}, Type{pt.scalar, 1}, {(uint32_t)l}));
                    }
                }
                r.ir.push_back(emit(Op::MakeVector, parts, ir_type(o[0]), {}, value_name(o[1])));
            } else if (rt.k == SType::Matrix) {
                for (size_t i = 2; i < o.size(); ++i) {
                    const Val *p = V(o[i]);
                    if (!p) return false;
                    r.ir.push_back(p->ir[0]);
                }
            } else {
                return fail("composite construct of an unsupported type");
            }
            vals[o[1]] = r;
            break;
        }
        case OpVectorShuffle: {
            const Val *a = V(o[2]);
            const Val *b = V(o[3]);
            if (!a || !b) return false;
            std::vector<uint32_t> lanes(o.begin() + 4, o.end());
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(Op::Shuffle, {a->ir[0], b->ir[0]}, ir_type(o[0]), lanes, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
        case OpFNegate:
        case OpConvertSToF:
        case OpConvertFToS: {
            const Val *a = V(o[2]);
            if (!a) return false;
            const Op op = ins.opcode == OpFNegate ? Op::FNeg : ins.opcode == OpConvertSToF ? Op::ConvSToF : Op::ConvFToS;
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(op, {a->ir[0]}, ir_type(o[0]), {}, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
        case OpIAdd:
        case OpFAdd:
        case OpFSub:
        case OpFMul:
        case OpFDiv:
        case OpDot:
        case OpFOrdLessThan:
        case OpFOrdGreaterThan: {
            const Val *a = V(o[2]);
            const Val *b = V(o[3]);
            if (!a || !b) return false;
            Op op;
            switch (ins.opcode) {
            case OpIAdd: op = Op::Add; break;
            case OpFAdd: op = Op::FAdd; break;
            case OpFSub: op = Op::FSub; break;
            case OpFMul: op = Op::FMul; break;
            case OpFDiv: op = Op::FDiv; break;
            case OpDot: op = Op::Dot; break;
            case OpFOrdLessThan: op = Op::FOrdLt; break;
            default: op = Op::FOrdGt; break;
            }
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(op, {a->ir[0], b->ir[0]}, ir_type(o[0]), {}, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
        case OpSelect: {
            const Val *c = V(o[2]);
            const Val *a = V(o[3]);
            const Val *b = V(o[4]);
            if (!c || !a || !b) return false;
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(Op::Select, {c->ir[0], a->ir[0], b->ir[0]}, ir_type(o[0]), {}, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
        case OpVectorTimesScalar: {
            const Val *v = V(o[2]);
            const Val *s = V(o[3]);
            if (!v || !s) return false;
            const Type vt = ir_type(o[0]);
            const uint32_t sp = emit(Op::Splat, {s->ir[0]}, vt);
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(Op::FMul, {v->ir[0], sp}, vt, {}, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
        case OpMatrixTimesVector: {
            const Val *mm = V(o[2]);
            const Val *v = V(o[3]);
            if (!mm || !v) return false;
            const Type vt = ir_type(o[0]);
            uint32_t acc = 0;
            for (size_t c = 0; c < mm->ir.size(); ++c) {
                const uint32_t vc = emit(Op::Extract, {v->ir[0]}, Type::f32(), {(uint32_t)c});
                const uint32_t prod = emit(Op::FMul, {mm->ir[c], emit(Op::Splat, {vc}, vt)}, vt);
                acc = c == 0 ? prod : emit(Op::FAdd, {acc, prod}, vt);
            }
            Val r;
            r.type = o[0];
            r.ir.push_back(acc);
            m.value(acc).name = value_name(o[1]);
            vals[o[1]] = r;
            break;
        }
        case OpBitFieldInsert:
        case OpBitFieldSExtract:
        case OpBitFieldUExtract: {
            // offset and count are scalars that apply to every component
            const Type rt = ir_type(o[0]);
            if (rt.scalar == ir::Scalar::Void) return fail("unsupported result type");
            std::vector<uint32_t> args;
            for (size_t i = 2; i < o.size(); ++i) {
                const Val *p = V(o[i]);
                if (!p) return false;
                args.push_back(p->ir[0]);
            }
            const size_t want = ins.opcode == OpBitFieldInsert ? 4 : 3;
            if (args.size() != want) return fail("bit field instruction with " + std::to_string(args.size()) + " operands");
            const Op op = ins.opcode == OpBitFieldInsert    ? Op::BitFieldInsert
                          : ins.opcode == OpBitFieldSExtract ? Op::BitFieldSExtract
                                                             : Op::BitFieldUExtract;
            return define(o[0], o[1], emit(op, args, rt, {}, value_name(o[1])));
        }
        case OpMatrixTimesScalar:
        case OpMatrixTimesMatrix:
        case OpOuterProduct:
        case OpTranspose: {
            // a matrix is its columns; the result's column type is its element type
            const SType &rs = types.at(o[0]);
            if (rs.k != SType::Matrix) return fail("matrix instruction with a non-matrix result");
            const Type ct = ir_type(rs.elem);
            const Val *a = V(o[2]);
            if (!a) return false;
            Val r;
            r.type = o[0];
            if (ins.opcode == OpTranspose) {
                // column c of the transpose holds row c: lane r of a's column r
                for (uint32_t c = 0; c < rs.count; ++c) {
                    std::vector<uint32_t> lanes;
                    for (size_t k = 0; k < a->ir.size(); ++k)
                        lanes.push_back(emit(Op::Extract, {a->ir[k]}, Type::f32(), {c}));
                    r.ir.push_back(emit(Op::MakeVector, lanes, ct));
                }
            } else {
                const Val *b = V(o[3]);
                if (!b) return false;
                for (uint32_t c = 0; c < rs.count; ++c) {
                    if (ins.opcode == OpMatrixTimesScalar) {
                        r.ir.push_back(emit(Op::FMul, {a->ir.at(c), emit(Op::Splat, {b->ir[0]}, ct)}, ct));
                    } else if (ins.opcode == OpOuterProduct) {
                        const uint32_t s = emit(Op::Extract, {b->ir[0]}, Type::f32(), {c});
                        r.ir.push_back(emit(Op::FMul, {a->ir[0], emit(Op::Splat, {s}, ct)}, ct));
                    } else {
                        // column c = a * (column c of b)
                        uint32_t acc = 0;
                        for (size_t k = 0; k < a->ir.size(); ++k) {
                            const uint32_t s = emit(Op::Extract, {b->ir.at(c)}, Type::f32(), {(uint32_t)k});
                            const uint32_t prod = emit(Op::FMul, {a->ir[k], emit(Op::Splat, {s}, ct)}, ct);
                            acc = k == 0 ? prod : emit(Op::FAdd, {acc, prod}, ct);
                        }
                        r.ir.push_back(acc);
                    }
                }
            }
            vals[o[1]] = r;
            break;
        }
        case OpArrayLength: {
            // elements of the run-time array that ends the storage buffer's
            // block: (bound bytes - the member's offset) / its stride, 0 when
            // the range ends before the member
            const Val *b = V(o[2]);
            if (!b) return false;
            if (!b->is_ptr || b->layout != 1) return fail("OpArrayLength of a non-buffer pointer");
            const SType &st = types.at(b->type);
            if (st.k != SType::Struct || o[3] >= st.members.size() || types.at(st.members[o[3]]).k != SType::Array ||
                types.at(st.members[o[3]]).count != 0)
                return fail("OpArrayLength of a member that is not a run-time array");
            const uint32_t off = member_offset(b->type, o[3], 1);
            const uint32_t stride = array_stride(st.members[o[3]], 1);
            if (stride == 0) return fail("run-time array without a stride");
            const Type i = Type::i32();
            const uint32_t bytes = emit(Op::BufferSize, {b->ir[0]}, i);
            const uint32_t over = emit(Op::UMax, {bytes, m.new_const_i32((int32_t)off)}, i);
            const uint32_t n = emit(Op::UDiv, {emit(Op::ISub, {over, m.new_const_i32((int32_t)off)}, i),
                                               m.new_const_i32((int32_t)stride)}, i, {}, value_name(o[1]));
            return define(o[0], o[1], n);
        }
        case OpVectorTimesMatrix: {
            // lane c = dot(v, column c)
            const Val *v = V(o[2]);
            const Val *mm = V(o[3]);
            if (!v || !mm) return false;
            std::vector<uint32_t> lanes;
            for (uint32_t col : mm->ir) lanes.push_back(emit(Op::Dot, {v->ir[0], col}, Type::f32()));
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(Op::MakeVector, lanes, ir_type(o[0]), {}, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
        case OpExtInst: {
            if (o[2] != glsl_set) return fail("unknown extended instruction set");
            const uint32_t inst = o[3];
            std::vector<const Val *> a;
            for (size_t i = 4; i < o.size(); ++i) {
                const Val *p = V(o[i]);
                if (!p) return false;
                a.push_back(p);
            }
            const Type rt = ir_type(o[0]);
            uint32_t res = 0;
            auto one = [&]() { return m.new_const_f32(1.0f); };
            // operand count of each GLSL.std.450 instruction this translator takes
            static const std::map<uint32_t, uint32_t> arity = {
                {1, 1}, {2, 1}, {3, 1}, {4, 1}, {5, 1}, {6, 1}, {8, 1}, {9, 1}, {10, 1}, {13, 1}, {14, 1},
                {26, 2}, {27, 1}, {28, 1}, {29, 1}, {30, 1}, {31, 1}, {32, 1}, {37, 2}, {38, 2}, {39, 2},
                {40, 2}, {41, 2}, {42, 2}, {43, 3}, {44, 3}, {45, 3}, {46, 3}, {48, 2}, {49, 3}, {66, 1},
                {67, 2}, {68, 2}, {69, 1}, {71, 2}, {7, 1}, {11, 1}, {12, 1}, {15, 1}, {16, 1}, {17, 1}, {18, 1},
                {19, 1}, {20, 1}, {21, 1}, {22, 1}, {23, 1}, {24, 1}, {25, 2}, {33, 1}, {34, 1}, {50, 3}, {70, 3},
                {72, 3}, {73, 1}, {74, 1}, {75, 1}, {79, 2}, {80, 2}, {81, 3}, {54, 1}, {55, 1}, {56, 1}, {57, 1},
                {58, 1}, {59, 1}, {60, 1}, {61, 1}, {62, 1}, {63, 1}, {35, 2}, {51, 2}, {53, 2}, {36, 1}, {52, 1}};
            auto ar = arity.find(inst);
            if (ar == arity.end()) return fail("unsupported GLSL.std.450 instruction " + std::to_string(inst));
            if (a.size() != ar->second)
                return fail("GLSL.std.450 instruction " + std::to_string(inst) + " takes " +
                            std::to_string(ar->second) + " operands");
            if (inst == 36 || inst == 52) {
                // ModfStruct {x - trunc(x), trunc(x)}, FrexpStruct {mantissa, exponent}
                const SType &st = types.at(o[0]);
                if (st.k != SType::Struct || st.members.size() != 2) return fail("ModfStruct / FrexpStruct result type");
                const Type ft = ir_type(st.members[0]), it = ir_type(st.members[1]);
                if (ft.scalar != ir::Scalar::F32) return fail("ModfStruct / FrexpStruct result type");
                Val r;
                r.type = o[0];
                if (inst == 36) {
                    const uint32_t w = emit(Op::Trunc, {a[0]->ir[0]}, ft);
                    r.ir = {emit(Op::FSub, {a[0]->ir[0], w}, ft), w};
                } else {
                    r.ir = {emit(Op::FrexpMant, {a[0]->ir[0]}, ft), emit(Op::FrexpExp, {a[0]->ir[0]}, it)};
                }
                vals[o[1]] = r;
                break;
            }
            if (inst == 33 || inst == 34) {
                // Determinant, MatrixInverse: cofactor expansion over the columns' lanes
                // (element (row r, column c) is lane r of column c)
                const size_t n = a[0]->ir.size();
                if (n < 2 || n > 4 || types.at(a[0]->type).k != SType::Matrix)
                    return fail("Determinant / MatrixInverse of a non-square or unsupported matrix");
                uint32_t e[4][4];
                for (size_t c = 0; c < n; ++c)
                    for (size_t r = 0; r < n; ++r) e[r][c] = emit(Op::Extract, {a[0]->ir[c]}, Type::f32(), {(uint32_t)r});
                const Type f = Type::f32();
                std::function<uint32_t(const std::vector<int> &, const std::vector<int> &)> det =
                    [&](const std::vector<int> &rows, const std::vector<int> &cols) -> uint32_t {
                    if (rows.size() == 1) return e[rows[0]][cols[0]];
                    if (rows.size() == 2)
                        return emit(Op::FSub, {emit(Op::FMul, {e[rows[0]][cols[0]], e[rows[1]][cols[1]]}, f),
                                               emit(Op::FMul, {e[rows[1]][cols[0]], e[rows[0]][cols[1]]}, f)}, f);
                    const std::vector<int> rest(cols.begin() + 1, cols.end());
                    uint32_t acc = 0;
                    for (size_t i = 0; i < rows.size(); ++i) {
                        std::vector<int> sub;
                        for (size_t j = 0; j < rows.size(); ++j)
                            if (j != i) sub.push_back(rows[j]);
                        const uint32_t t = emit(Op::FMul, {e[rows[i]][cols[0]], det(sub, rest)}, f);
                        acc = i == 0 ? t : emit(i % 2 ? Op::FSub : Op::FAdd, {acc, t}, f);
                    }
                    return acc;
                };
                std::vector<int> all;
                for (size_t i = 0; i < n; ++i) all.push_back((int)i);
                const uint32_t d = det(all, all);
                Val r;
                r.type = o[0];
                if (inst == 33) {
                    m.value(d).name = value_name(o[1]);
                    r.ir.push_back(d);
                } else {
                    // inverse (row r, column c) = (-1)^(r+c) minor(row c, column r) / det
                    const uint32_t rd = emit(Op::FDiv, {one(), d}, f);
                    const Type ct = ir_type(types.at(o[0]).elem);
                    for (size_t c = 0; c < n; ++c) {
                        std::vector<uint32_t> lanes;
                        for (size_t rr = 0; rr < n; ++rr) {
                            std::vector<int> rows, cols;
                            for (size_t k = 0; k < n; ++k) {
                                if (k != c) rows.push_back((int)k);
                                if (k != rr) cols.push_back((int)k);
                            }
                            uint32_t mnr = emit(Op::FMul, {det(rows, cols), rd}, f);
                            if ((rr + c) % 2) mnr = emit(Op::FNeg, {mnr}, f);
                            lanes.push_back(mnr);
                        }
                        r.ir.push_back(emit(Op::MakeVector, lanes, ct));
                    }
                }
                vals[o[1]] = r;
                break;
            }
            if (rt.scalar == ir::Scalar::Void) return fail("GLSL.std.450 result of an unsupported type");
            auto splat = [&](uint32_t scalar) { return rt.lanes > 1 ? emit(Op::Splat, {scalar}, rt) : scalar; };
            auto unary = [&](Op op) { return emit(op, {a[0]->ir[0]}, rt); };
            auto binary = [&](Op op) { return emit(op, {a[0]->ir[0], a[1]->ir[0]}, rt); };
            switch (inst) {
            case 1: {                                                                                   // Round: trunc(x + copysign(0.5, x))
                const uint32_t half = splat(m.new_const_f32(0.5f));
                const uint32_t neg = emit(Op::FCmp, {a[0]->ir[0], splat(m.new_const_f32(0.0f))},
                                          Type::boolean(rt.lanes), {ir::FOLt});
                const uint32_t h = emit(Op::Select, {neg, emit(Op::FNeg, {half}, rt), half}, rt);
                res = emit(Op::Trunc, {emit(Op::FAdd, {a[0]->ir[0], h}, rt)}, rt);
                break;
            }
            // The forms below are those whose precision the Vulkan specification states
            // for these functions ("inherited from" them), except where noted.
            case 7: {                                                                                   // SSign: clamp(x, -1, 1)
                const uint32_t lo = splat(m.new_const_i32(-1)), hi = splat(m.new_const_i32(1));
                res = emit(Op::SClamp, {a[0]->ir[0], lo, hi}, rt);
                break;
            }
            case 11: res = emit(Op::FMul, {a[0]->ir[0], splat(m.new_const_f32(0.017453292519943295f))}, rt); break;  // Radians
            case 12: res = emit(Op::FMul, {a[0]->ir[0], splat(m.new_const_f32(57.29577951308232f))}, rt); break;     // Degrees
            case 15: res = emit(Op::FDiv, {unary(Op::Sin), unary(Op::Cos)}, rt); break;              // Tan: sin / cos
            case 16:                                                                                    // Asin: atan2(x, sqrt(1 - x x))
            case 17: {                                                                                  // Acos: atan2(sqrt(1 - x x), x)
                const uint32_t x = a[0]->ir[0];
                const uint32_t c = emit(Op::Sqrt, {emit(Op::FSub, {splat(one()), emit(Op::FMul, {x, x}, rt)}, rt)}, rt);
                res = inst == 16 ? emit(Op::Atan2, {x, c}, rt) : emit(Op::Atan2, {c, x}, rt);
                break;
            }
            case 18: res = emit(Op::Atan2, {a[0]->ir[0], splat(one())}, rt); break;                  // Atan: atan2(x, 1)
            case 25: res = binary(Op::Atan2); break;                                                   // Atan2
            case 19:                                                                                    // Sinh: (e^x - e^-x) / 2
            case 20: {                                                                                  // Cosh: (e^x + e^-x) / 2
                const uint32_t ep = unary(Op::Exp), en = emit(Op::Exp, {emit(Op::FNeg, {a[0]->ir[0]}, rt)}, rt);
                res = emit(Op::FMul, {emit(inst == 19 ? Op::FSub : Op::FAdd, {ep, en}, rt), splat(m.new_const_f32(0.5f))},
                           rt);
                break;
            }
            case 21: {                                                                                  // Tanh: 1 - 2 / (e^2x + 1)
                // (sinh / cosh would give inf / inf for |x| > 89; this form saturates to +-1)
                const uint32_t e2 = emit(Op::Exp, {emit(Op::FAdd, {a[0]->ir[0], a[0]->ir[0]}, rt)}, rt);
                const uint32_t q = emit(Op::FDiv, {splat(m.new_const_f32(2.0f)), emit(Op::FAdd, {e2, splat(one())}, rt)}, rt);
                res = emit(Op::FSub, {splat(one()), q}, rt);
                break;
            }
            case 22: {                                                                                  // Asinh: sign(x) log(|x| + sqrt(x x + 1))
                // (the odd symmetry keeps large negative x from cancelling)
                const uint32_t x = a[0]->ir[0];
                const uint32_t ax = emit(Op::FMax, {x, emit(Op::FNeg, {x}, rt)}, rt);
                const uint32_t r = emit(Op::Sqrt, {emit(Op::FAdd, {emit(Op::FMul, {x, x}, rt), splat(one())}, rt)}, rt);
                res = emit(Op::FMul, {unary(Op::FSign), emit(Op::Log, {emit(Op::FAdd, {ax, r}, rt)}, rt)}, rt);
                break;
            }
            case 23: {                                                                                  // Acosh: log(x + sqrt(x x - 1))
                const uint32_t x = a[0]->ir[0];
                const uint32_t r = emit(Op::Sqrt, {emit(Op::FSub, {emit(Op::FMul, {x, x}, rt), splat(one())}, rt)}, rt);
                res = emit(Op::Log, {emit(Op::FAdd, {x, r}, rt)}, rt);
                break;
            }
            case 24: {                                                                                  // Atanh: log((1 + x) / (1 - x)) / 2
                const uint32_t x = a[0]->ir[0];
                const uint32_t q = emit(Op::FDiv, {emit(Op::FAdd, {splat(one()), x}, rt), emit(Op::FSub, {splat(one()), x}, rt)},
                                        rt);
                res = emit(Op::FMul, {emit(Op::Log, {q}, rt), splat(m.new_const_f32(0.5f))}, rt);
                break;
            }
            case 50:                                                                                    // Fma: a b + c (OpFMul, OpFAdd)
                res = emit(Op::FAdd, {emit(Op::FMul, {a[0]->ir[0], a[1]->ir[0]}, rt), a[2]->ir[0]}, rt);
                break;
            case 70: {                                                                                  // FaceForward: dot(Nref, I) < 0 ? N : -N
                const uint32_t d = rt.lanes > 1 ? emit(Op::Dot, {a[2]->ir[0], a[1]->ir[0]}, Type::f32())
                                                : emit(Op::FMul, {a[2]->ir[0], a[1]->ir[0]}, rt);
                const uint32_t lt = emit(Op::FCmp, {d, m.new_const_f32(0.0f)}, Type::boolean(1), {ir::FOLt});
                res = emit(Op::Select, {lt, a[0]->ir[0], emit(Op::FNeg, {a[0]->ir[0]}, rt)}, rt);
                break;
            }
            case 72: {                                                                                  // Refract(I, N, eta)
                // k = 1 - eta^2 (1 - dot(N, I)^2); k < 0 ? 0 : eta I - (eta dot(N, I) + sqrt(k)) N
                const uint32_t i = a[0]->ir[0], nn = a[1]->ir[0], eta = a[2]->ir[0];
                const Type f = Type::f32();
                const uint32_t d = rt.lanes > 1 ? emit(Op::Dot, {nn, i}, f) : emit(Op::FMul, {nn, i}, f);
                const uint32_t k = emit(Op::FSub, {one(), emit(Op::FMul, {emit(Op::FMul, {eta, eta}, f),
                                                                          emit(Op::FSub, {one(), emit(Op::FMul, {d, d}, f)}, f)}, f)},
                                        f);
                const uint32_t t = emit(Op::FAdd, {emit(Op::FMul, {eta, d}, f), emit(Op::Sqrt, {k}, f)}, f);
                const uint32_t v = emit(Op::FSub, {emit(Op::FMul, {splat(eta), i}, rt), emit(Op::FMul, {splat(t), nn}, rt)}, rt);
                const uint32_t neg = emit(Op::FCmp, {k, m.new_const_f32(0.0f)}, Type::boolean(1), {ir::FOLt});
                res = emit(Op::Select, {neg, splat(m.new_const_f32(0.0f)), v}, rt);
                break;
            }
            case 54:                                                                                    // PackSnorm4x8
            case 55:                                                                                    // PackUnorm4x8
            case 56:                                                                                    // PackSnorm2x16
            case 57:                                                                                    // PackUnorm2x16
            case 58: {                                                                                  // PackHalf2x16
                // field c = round(clamp(v[c], lo, 1) * scale) (snorm: two's complement),
                // at bit c * width; half: the binary16 bits
                const bool half = inst == 58, snorm = inst == 54 || inst == 56, four = inst <= 55;
                const uint32_t width = four ? 8 : 16, n = four ? 4 : 2;
                const float scale = four ? (snorm ? 127.0f : 255.0f) : (snorm ? 32767.0f : 65535.0f);
                const Type i = Type::i32(), f = Type::f32();
                uint32_t acc = 0;
                for (uint32_t c = 0; c < n; ++c) {
                    const uint32_t x = emit(Op::Extract, {a[0]->ir[0]}, f, {c});
                    uint32_t q;
                    if (half) {
                        q = emit(Op::F32ToF16, {x}, i);
                    } else {
                        const uint32_t cl = emit(Op::FClamp, {x, m.new_const_f32(snorm ? -1.0f : 0.0f), one()}, f);
                        const uint32_t r = emit(Op::RoundEven, {emit(Op::FMul, {cl, m.new_const_f32(scale)}, f)}, f);
                        q = emit(snorm ? Op::ConvFToS : Op::ConvFToU, {r}, i);
                        q = emit(Op::And, {q, m.new_const_i32((int32_t)((1u << width) - 1u))}, i);
                    }
                    if (c) q = emit(Op::Shl, {q, m.new_const_i32((int32_t)(c * width))}, i);
                    acc = c ? emit(Op::Or, {acc, q}, i) : q;
                }
                res = acc;
                break;
            }
            case 59:                                                                                    // UnpackHalf2x16
            case 60:                                                                                    // UnpackSnorm2x16
            case 61:                                                                                    // UnpackUnorm2x16
            case 62:                                                                                    // UnpackSnorm4x8
            case 63: {                                                                                  // UnpackUnorm4x8
                // component c from the field at bit c * width: unorm f / scale,
                // snorm clamp(f / scale, -1, 1) of the sign-extended field
                const bool half = inst == 59, snorm = inst == 60 || inst == 62, four = inst >= 62;
                const uint32_t width = four ? 8 : 16, n = four ? 4 : 2;
                const float scale = four ? (snorm ? 127.0f : 255.0f) : (snorm ? 32767.0f : 65535.0f);
                const Type i = Type::i32(), f = Type::f32();
                const uint32_t u = a[0]->ir[0];
                std::vector<uint32_t> lanes;
                for (uint32_t c = 0; c < n; ++c) {
                    uint32_t x;
                    if (half) {
                        const uint32_t bits = c ? emit(Op::ShrL, {u, m.new_const_i32(16)}, i) : u;
                        x = emit(Op::F16ToF32, {emit(Op::And, {bits, m.new_const_i32(0xFFFF)}, i)}, f);
                    } else if (snorm) {
                        const uint32_t up = emit(Op::Shl, {u, m.new_const_i32((int32_t)(32 - (c + 1) * width))}, i);
                        const uint32_t v = emit(Op::ShrA, {up, m.new_const_i32((int32_t)(32 - width))}, i);
                        x = emit(Op::FDiv, {emit(Op::ConvSToF, {v}, f), m.new_const_f32(scale)}, f);
                        x = emit(Op::FClamp, {x, m.new_const_f32(-1.0f), one()}, f);
                    } else {
                        const uint32_t sh = c ? emit(Op::ShrL, {u, m.new_const_i32((int32_t)(c * width))}, i) : u;
                        const uint32_t v = emit(Op::And, {sh, m.new_const_i32((int32_t)((1u << width) - 1u))}, i);
                        x = emit(Op::FDiv, {emit(Op::ConvUToF, {v}, f), m.new_const_f32(scale)}, f);
                    }
                    lanes.push_back(x);
                }
                res = emit(Op::MakeVector, lanes, rt);
                break;
            }
            case 35:                                                                                    // Modf(x, *whole): x - trunc(x)
            case 51: {                                                                                  // Frexp(x, *exponent): mantissa
                if (!a[1]->is_ptr) return fail("Modf / Frexp without a pointer operand");
                Val second;
                second.type = a[1]->type;
                second.ir.push_back(inst == 35 ? unary(Op::Trunc) : emit(Op::FrexpExp, {a[0]->ir[0]}, Type::i32(rt.lanes)));
                if (!store(*a[1], second)) return false;
                res = inst == 35 ? emit(Op::FSub, {a[0]->ir[0], second.ir[0]}, rt) : unary(Op::FrexpMant);
                break;
            }
            case 53: res = binary(Op::Ldexp); break;                                                   // Ldexp
            case 73: res = unary(Op::FindILsb); break;
            case 74: res = unary(Op::FindSMsb); break;
            case 75: res = unary(Op::FindUMsb); break;
            // NMin, NMax, NClamp: fminf / fmaxf already return the operand that is not a NaN
            case 79: res = binary(Op::FMin); break;
            case 80: res = binary(Op::FMax); break;
            case 81: res = emit(Op::FClamp, {a[0]->ir[0], a[1]->ir[0], a[2]->ir[0]}, rt); break;
            case 2: res = unary(Op::RoundEven); break;
            case 3: res = unary(Op::Trunc); break;
            case 5: res = unary(Op::SAbs); break;
            case 6: res = unary(Op::FSign); break;
            case 13: res = unary(Op::Sin); break;
            case 14: res = unary(Op::Cos); break;
            case 26: res = binary(Op::Pow); break;
            case 27: res = unary(Op::Exp); break;
            case 28: res = unary(Op::Log); break;
            case 29: res = unary(Op::Exp2); break;
            case 30: res = unary(Op::Log2); break;
            case 38: res = binary(Op::UMin); break;
            case 39: res = binary(Op::SMin); break;
            case 41: res = binary(Op::UMax); break;
            case 42: res = binary(Op::SMax); break;
            case 44: res = emit(Op::UClamp, {a[0]->ir[0], a[1]->ir[0], a[2]->ir[0]}, rt); break;
            case 45: res = emit(Op::SClamp, {a[0]->ir[0], a[1]->ir[0], a[2]->ir[0]}, rt); break;
            case 48: {                                                                                  // Step: x < edge ? 0 : 1
                const uint32_t lt = emit(Op::FCmp, {a[1]->ir[0], a[0]->ir[0]}, Type::boolean(rt.lanes), {ir::FOLt});
                res = emit(Op::Select, {lt, splat(m.new_const_f32(0.0f)), splat(one())}, rt);
                break;
            }
            case 49: {                                                                                  // SmoothStep: t*t*(3 - 2t), t = clamp((x-e0)/(e1-e0), 0, 1)
                const uint32_t d = emit(Op::FDiv, {emit(Op::FSub, {a[2]->ir[0], a[0]->ir[0]}, rt),
                                                   emit(Op::FSub, {a[1]->ir[0], a[0]->ir[0]}, rt)}, rt);
                const uint32_t t = emit(Op::FClamp, {d, splat(m.new_const_f32(0.0f)), splat(one())}, rt);
                const uint32_t k = emit(Op::FSub, {splat(m.new_const_f32(3.0f)),
                                                   emit(Op::FMul, {splat(m.new_const_f32(2.0f)), t}, rt)}, rt);
                res = emit(Op::FMul, {emit(Op::FMul, {t, t}, rt), k}, rt);
                break;
            }
            case 31: res = emit(Op::Sqrt, {a[0]->ir[0]}, rt); break;                                    // Sqrt
            case 32: res = emit(Op::FDiv, {one(), emit(Op::Sqrt, {a[0]->ir[0]}, rt)}, rt); break;      // InverseSqrt
            case 37: res = emit(Op::FMin, {a[0]->ir[0], a[1]->ir[0]}, rt); break;                      // FMin
            case 40: res = emit(Op::FMax, {a[0]->ir[0], a[1]->ir[0]}, rt); break;                      // FMax
            case 43: res = emit(Op::FClamp, {a[0]->ir[0], a[1]->ir[0], a[2]->ir[0]}, rt); break;       // FClamp
            case 66: {                                                                                  // Length
                const uint32_t d = emit(Op::Dot, {a[0]->ir[0], a[0]->ir[0]}, Type::f32());
                res = emit(Op::Sqrt, {d}, rt);
                break;
            }
            case 69: {                                                                                  // Normalize
                const uint32_t d = emit(Op::Dot, {a[0]->ir[0], a[0]->ir[0]}, Type::f32());
                const uint32_t rl = emit(Op::FDiv, {one(), emit(Op::Sqrt, {d}, Type::f32())}, Type::f32());
                res = emit(Op::FMul, {a[0]->ir[0], emit(Op::Splat, {rl}, rt)}, rt);
                break;
            }
            case 8: res = emit(Op::Floor, {a[0]->ir[0]}, rt); break;                                   // Floor
            case 9: res = emit(Op::Ceil, {a[0]->ir[0]}, rt); break;                                    // Ceil
            case 10:                                                                                    // Fract: x - floor(x)
                res = emit(Op::FSub, {a[0]->ir[0], emit(Op::Floor, {a[0]->ir[0]}, rt)}, rt);
                break;
            case 4: res = emit(Op::FMax, {a[0]->ir[0], emit(Op::FNeg, {a[0]->ir[0]}, rt)}, rt); break;    // FAbs
            case 46: {                                                                                  // FMix: x * (1 - a) + y * a
                // the form whose precision Vulkan states; exact at a = 0 and a = 1
                const uint32_t om = emit(Op::FSub, {splat(one()), a[2]->ir[0]}, rt);
                res = emit(Op::FAdd, {emit(Op::FMul, {a[0]->ir[0], om}, rt), emit(Op::FMul, {a[1]->ir[0], a[2]->ir[0]}, rt)},
                           rt);
                break;
            }
            case 67: {                                                                                  // Distance
                const Type vt = ir_type(a[0]->type);
                const uint32_t d = emit(Op::FSub, {a[0]->ir[0], a[1]->ir[0]}, vt);
                const uint32_t dd = vt.lanes > 1 ? emit(Op::Dot, {d, d}, Type::f32()) : emit(Op::FMul, {d, d}, vt);
                res = emit(Op::Sqrt, {dd}, rt);
                break;
            }
            case 68: {                                                                                  // Cross: x.yzx*y.zxy - x.zxy*y.yzx
                const uint32_t x = a[0]->ir[0], y = a[1]->ir[0];
                auto sw = [&](uint32_t v, uint32_t i, uint32_t j, uint32_t k) { return emit(Op::Shuffle, {v, v}, rt, {i, j, k}); };
                const uint32_t p = emit(Op::FMul, {sw(x, 1, 2, 0), sw(y, 2, 0, 1)}, rt);
                const uint32_t q = emit(Op::FMul, {sw(x, 2, 0, 1), sw(y, 1, 2, 0)}, rt);
                res = emit(Op::FSub, {p, q}, rt);
                break;
            }
            case 71: {                                                                                  // Reflect: I - 2 dot(N, I) N
                const uint32_t i = a[0]->ir[0], n = a[1]->ir[0];
                const uint32_t d = rt.lanes > 1 ? emit(Op::Dot, {n, i}, Type::f32()) : emit(Op::FMul, {n, i}, rt);
                const uint32_t d2 = emit(Op::FMul, {m.new_const_f32(2.0f), d}, Type::f32());
                const uint32_t s2 = rt.lanes > 1 ? emit(Op::Splat, {d2}, rt) : d2;
                res = emit(Op::FSub, {i, emit(Op::FMul, {s2, n}, rt)}, rt);
                break;
            }
            default: return fail("unsupported GLSL.std.450 instruction " + std::to_string(inst));
            }
            m.value(res).name = value_name(o[1]);
            Val r;
            r.type = o[0];
            r.ir.push_back(res);
            vals[o[1]] = r;
            break;
        }
        case OpImageSampleImplicitLod:
        case OpImageSampleExplicitLod: {
            const Val *tex = V(o[2]);
            const Val *coord = V(o[3]);
            if (!tex || !coord) return false;
            Val r;
            r.type = o[0];
            r.ir.push_back(emit(Op::Sample, {tex->ir[0], coord->ir[0]}, ir_type(o[0]), {}, value_name(o[1])));
            vals[o[1]] = r;
            break;
        }
    default:
        return fail("unsupported SPIR-V opcode " + std::to_string(ins.opcode) + " in function body");
    }
    return true;
}

}  // namespace

bool translate_spirv(const uint32_t *words, size_t nwords, const std::string &entry_name, ir::Stage stage,
                     const std::vector<SpecConstant> &spec, ir::Module &out, std::string &err) {
    SpirvModule sm;
    if (!spirv_parse(words, nwords, sm, err)) return false;
    out = ir::Module();
    out.stage = stage;
    Translator tr(sm, out, err);
    for (const auto &s : spec) tr.spec[s.id] = s.bytes;
    if (!tr.collect()) return false;
    if (const uint32_t id = tr.out_of_bounds())
        return tr.fail("SPIR-V id %" + std::to_string(id) + " is out of bounds (header bound " + std::to_string(sm.bound) +
                       ")");
    const uint32_t model = stage == ir::Stage::Vertex ? 0u : stage == ir::Stage::Fragment ? 4u : 5u;
    uint32_t fn = 0;
    std::vector<uint32_t> iface;
    for (const auto &ins : sm.instructions) {
        if (ins.opcode == OpCapability) {
            const uint32_t cap = ins.operands[0];
            if (cap != 0 && cap != 1 && cap != 11) return tr.fail("unsupported capability " + std::to_string(cap));
        } else if (ins.opcode == OpMemoryModel) {
            if (ins.operands[0] != 0) return tr.fail("only the Logical addressing model is supported");
        } else if (ins.opcode == OpEntryPoint && ins.operands[0] == model) {
            size_t next;
            if (spirv_string(ins.operands, 2, &next) == entry_name) {
                fn = ins.operands[1];
                iface.assign(ins.operands.begin() + next, ins.operands.end());
            }
        }
    }
    if (!fn) return tr.fail("entry point \"" + entry_name + "\" not found for this stage");
    bool upper_left = stage != ir::Stage::Fragment;
    for (const auto &ins : sm.instructions) {
        if (ins.opcode != OpExecutionMode || ins.operands[0] != fn) continue;
        const uint32_t mode = ins.operands[1];
        if (mode == 17 && ins.operands.size() >= 5) {  // LocalSize x y z
            for (int k = 0; k < 3; ++k) out.local_size[k] = ins.operands[2 + k];
        }
        if (mode == 7) upper_left = true;
        else if (mode == 8) return tr.fail("OriginLowerLeft is not supported (Vulkan requires OriginUpperLeft)");
    }
    if (!upper_left) return tr.fail("fragment shader without OriginUpperLeft");
    if (stage == ir::Stage::Compute) {
        const uint64_t n = (uint64_t)out.local_size[0] * out.local_size[1] * out.local_size[2];
        if (n == 0) return tr.fail("compute shader without a LocalSize");
        if (n > 1024) return tr.fail("workgroup of " + std::to_string(n) + " invocations (at most 1024)");
    }
    if (!tr.build_interface(fn, iface, stage)) return false;
    if (!tr.translate_function(fn)) return false;
    if (const uint32_t id = tr.out_of_bounds())
        return tr.fail("SPIR-V id %" + std::to_string(id) + " is out of bounds (header bound " + std::to_string(sm.bound) +
                       ")");
    // KAZAN_IR_ROUND_TRIP=1: hand on the module as its text form parsed back
    // (what a pipeline built from IR text would get; tests render with it)
    const char *rt = std::getenv("KAZAN_IR_ROUND_TRIP");
    if (rt && std::strcmp(rt, "0") != 0) {
        ir::Module back;
        if (!ir::parse_module_text(ir::module_text(out), back, err)) {
            err = "IR text round trip: " + err;
            return false;
        }
        back.stage = stage;
        for (int k = 0; k < 3; ++k) back.local_size[k] = out.local_size[k];
        out = std::move(back);
    }
    return true;
}

}  // namespace kz
