// spirv_reader.cpp — SPIR-V binary decoding and a textual dump.
#include "spirv_reader.h"

#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>

namespace kz {

bool spirv_words_from_bytes(const uint8_t *bytes, size_t n, std::vector<uint32_t> &words, std::string &err) {
    if (n % 4 != 0) {
        err = "SPIR-V byte length is not a multiple of 4";
        return false;
    }
    words.resize(n / 4);
    std::memcpy(words.data(), bytes, n);
    if (words.empty()) {
        err = "empty SPIR-V module";
        return false;
    }
    if (words[0] == kSpirvMagic) return true;
    if (__builtin_bswap32(words[0]) == kSpirvMagic) {
        for (auto &w : words) w = __builtin_bswap32(w);
        return true;
    }
    err = "bad SPIR-V magic number";
    return false;
}

bool spirv_parse(const uint32_t *words, size_t n, SpirvModule &out, std::string &err) {
    if (n < 5) {
        err = "SPIR-V module shorter than its header";
        return false;
    }
    if (words[0] != kSpirvMagic) {
        err = "bad SPIR-V magic number";
        return false;
    }
    out.version = words[1];
    out.generator = words[2];
    out.bound = words[3];
    out.schema = words[4];
    out.instructions.clear();
    size_t i = 5;
    while (i < n) {
        const uint32_t w = words[i];
        const uint32_t wc = w >> 16;
        if (wc == 0 || i + wc > n) {
            err = "SPIR-V instruction word count out of range at word " + std::to_string(i);
            return false;
        }
        SpirvInstruction ins;
        ins.opcode = (uint16_t)(w & 0xFFFF);
        ins.word_offset = (uint32_t)i;
        ins.operands.assign(words + i + 1, words + i + wc);
        out.instructions.push_back(std::move(ins));
        i += wc;
    }
    return true;
}

std::string spirv_string(const std::vector<uint32_t> &ops, size_t start, size_t *next) {
    std::string s;
    size_t i = start;
    for (; i < ops.size(); ++i) {
        const uint32_t w = ops[i];
        bool end = false;
        for (int b = 0; b < 4; ++b) {
            const char c = (char)((w >> (8 * b)) & 0xFF);
            if (c == 0) {
                end = true;
                break;
            }
            s.push_back(c);
        }
        if (end) {
            ++i;
            break;
        }
    }
    if (next) *next = i;
    return s;
}

// ---------------------------------------------------------------------------
// disassembly

namespace {

enum Kind : uint8_t {
    K_ID,        // %n
    K_RTYPE,     // result type id
    K_RESULT,    // result id
    K_LIT,       // decimal literal
    K_STR,       // literal string
    K_CONST,     // context-dependent number, hex
    K_CAP,
    K_ADDR,
    K_MEM,
    K_EXECMODEL,
    K_EXECMODE,  // mode + decimal literals
    K_DECO,      // decoration + its literal parameters
    K_STORAGE,
    K_FNCTL,
    K_SELCTL,
    K_LOOPCTL,
    K_DIM,
    K_IMGFMT,
    K_SRCLANG,
    K_IMGOPS,    // image operand mask + ids
    K_IDS,       // trailing ids
    K_EXTINST,   // ext-inst set id + instruction name + ids
    K_SPECOP,    // opcode name + ids
    K_SWITCH,    // selector, default, (literal, label)*
    K_LITS,      // trailing decimal literals
};

struct OpInfo {
    const char *name;
    std::vector<Kind> kinds;
};

const std::map<uint16_t, OpInfo> &op_table() {
    static const std::map<uint16_t, OpInfo> t = {
        {1, {"OpUndef", {K_RTYPE, K_RESULT}}},
        {3, {"OpSource", {K_SRCLANG, K_LIT}}},
        {5, {"OpName", {K_ID, K_STR}}},
        {6, {"OpMemberName", {K_ID, K_LIT, K_STR}}},
        {7, {"OpString", {K_RESULT, K_STR}}},
        {8, {"OpLine", {K_ID, K_LIT, K_LIT}}},
        {11, {"OpExtInstImport", {K_RESULT, K_STR}}},
        {12, {"OpExtInst", {K_RTYPE, K_RESULT, K_EXTINST}}},
        {14, {"OpMemoryModel", {K_ADDR, K_MEM}}},
        {15, {"OpEntryPoint", {K_EXECMODEL, K_ID, K_STR, K_IDS}}},
        {16, {"OpExecutionMode", {K_ID, K_EXECMODE}}},
        {17, {"OpCapability", {K_CAP}}},
        {19, {"OpTypeVoid", {K_RESULT}}},
        {20, {"OpTypeBool", {K_RESULT}}},
        {21, {"OpTypeInt", {K_RESULT, K_LIT, K_LIT}}},
        {22, {"OpTypeFloat", {K_RESULT, K_LIT}}},
        {23, {"OpTypeVector", {K_RESULT, K_ID, K_LIT}}},
        {24, {"OpTypeMatrix", {K_RESULT, K_ID, K_LIT}}},
        {25, {"OpTypeImage", {K_RESULT, K_ID, K_DIM, K_LIT, K_LIT, K_LIT, K_LIT, K_IMGFMT}}},
        {26, {"OpTypeSampler", {K_RESULT}}},
        {27, {"OpTypeSampledImage", {K_RESULT, K_ID}}},
        {28, {"OpTypeArray", {K_RESULT, K_ID, K_ID}}},
        {30, {"OpTypeStruct", {K_RESULT, K_IDS}}},
        {32, {"OpTypePointer", {K_RESULT, K_STORAGE, K_ID}}},
        {33, {"OpTypeFunction", {K_RESULT, K_ID, K_IDS}}},
        {41, {"OpConstantTrue", {K_RTYPE, K_RESULT}}},
        {42, {"OpConstantFalse", {K_RTYPE, K_RESULT}}},
        {43, {"OpConstant", {K_RTYPE, K_RESULT, K_CONST}}},
        {44, {"OpConstantComposite", {K_RTYPE, K_RESULT, K_IDS}}},
        {50, {"OpSpecConstant", {K_RTYPE, K_RESULT, K_CONST}}},
        {52, {"OpSpecConstantOp", {K_RTYPE, K_RESULT, K_SPECOP}}},
        {54, {"OpFunction", {K_RTYPE, K_RESULT, K_FNCTL, K_ID}}},
        {55, {"OpFunctionParameter", {K_RTYPE, K_RESULT}}},
        {56, {"OpFunctionEnd", {}}},
        {57, {"OpFunctionCall", {K_RTYPE, K_RESULT, K_ID, K_IDS}}},
        {59, {"OpVariable", {K_RTYPE, K_RESULT, K_STORAGE, K_IDS}}},
        {61, {"OpLoad", {K_RTYPE, K_RESULT, K_ID, K_LITS}}},
        {62, {"OpStore", {K_ID, K_ID, K_LITS}}},
        {65, {"OpAccessChain", {K_RTYPE, K_RESULT, K_ID, K_IDS}}},
        {66, {"OpInBoundsAccessChain", {K_RTYPE, K_RESULT, K_ID, K_IDS}}},
        {71, {"OpDecorate", {K_ID, K_DECO}}},
        {72, {"OpMemberDecorate", {K_ID, K_LIT, K_DECO}}},
        {79, {"OpVectorShuffle", {K_RTYPE, K_RESULT, K_ID, K_ID, K_LITS}}},
        {80, {"OpCompositeConstruct", {K_RTYPE, K_RESULT, K_IDS}}},
        {81, {"OpCompositeExtract", {K_RTYPE, K_RESULT, K_ID, K_LITS}}},
        {87, {"OpImageSampleImplicitLod", {K_RTYPE, K_RESULT, K_ID, K_ID, K_IMGOPS}}},
        {88, {"OpImageSampleExplicitLod", {K_RTYPE, K_RESULT, K_ID, K_ID, K_IMGOPS}}},
        {109, {"OpConvertFToU", {K_RTYPE, K_RESULT, K_ID}}},
        {110, {"OpConvertFToS", {K_RTYPE, K_RESULT, K_ID}}},
        {111, {"OpConvertSToF", {K_RTYPE, K_RESULT, K_ID}}},
        {112, {"OpConvertUToF", {K_RTYPE, K_RESULT, K_ID}}},
        {113, {"OpUConvert", {K_RTYPE, K_RESULT, K_ID}}},
        {114, {"OpSConvert", {K_RTYPE, K_RESULT, K_ID}}},
        {124, {"OpBitcast", {K_RTYPE, K_RESULT, K_ID}}},
        {126, {"OpSNegate", {K_RTYPE, K_RESULT, K_ID}}},
        {127, {"OpFNegate", {K_RTYPE, K_RESULT, K_ID}}},
        {128, {"OpIAdd", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {129, {"OpFAdd", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {130, {"OpISub", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {131, {"OpFSub", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {132, {"OpIMul", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {133, {"OpFMul", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {135, {"OpSDiv", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {136, {"OpFDiv", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {142, {"OpVectorTimesScalar", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {145, {"OpMatrixTimesVector", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {148, {"OpDot", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {169, {"OpSelect", {K_RTYPE, K_RESULT, K_ID, K_ID, K_ID}}},
        {184, {"OpFOrdLessThan", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {186, {"OpFOrdGreaterThan", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {197, {"OpBitwiseOr", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {198, {"OpBitwiseXor", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {199, {"OpBitwiseAnd", {K_RTYPE, K_RESULT, K_ID, K_ID}}},
        {245, {"OpPhi", {K_RTYPE, K_RESULT, K_IDS}}},
        {246, {"OpLoopMerge", {K_ID, K_ID, K_LOOPCTL}}},
        {247, {"OpSelectionMerge", {K_ID, K_SELCTL}}},
        {248, {"OpLabel", {K_RESULT}}},
        {249, {"OpBranch", {K_ID}}},
        {250, {"OpBranchConditional", {K_ID, K_ID, K_ID, K_LITS}}},
        {251, {"OpSwitch", {K_SWITCH}}},
        {252, {"OpKill", {}}},
        {253, {"OpReturn", {}}},
        {254, {"OpReturnValue", {K_ID}}},
    };
    return t;
}

const char *spec_op_name(uint32_t op) {
    switch (op) {
    case 128: return "IAdd";
    case 130: return "ISub";
    case 132: return "IMul";
    case 135: return "SDiv";
    case 197: return "BitwiseOr";
    case 198: return "BitwiseXor";
    case 199: return "BitwiseAnd";
    default: return nullptr;
    }
}

const char *glsl_name(uint32_t i) {
    static const std::map<uint32_t, const char *> t = {
        {1, "Round"}, {4, "FAbs"}, {8, "Floor"}, {13, "Sin"}, {14, "Cos"}, {26, "Pow"}, {31, "Sqrt"},
        {32, "InverseSqrt"}, {37, "FMin"}, {40, "FMax"}, {43, "FClamp"}, {46, "FMix"}, {66, "Length"},
        {67, "Distance"}, {68, "Cross"}, {69, "Normalize"}, {71, "Reflect"}};
    auto it = t.find(i);
    return it == t.end() ? nullptr : it->second;
}

const char *enum_name(Kind k, uint32_t v) {
    switch (k) {
    case K_CAP:
        switch (v) {
        case 0: return "Matrix";
        case 1: return "Shader";
        case 11: return "Int64";
        case 12: return "Int16";
        }
        break;
    case K_ADDR:
        if (v == 0) return "Logical";
        if (v == 1) return "Physical32";
        if (v == 2) return "Physical64";
        break;
    case K_MEM:
        if (v == 0) return "Simple";
        if (v == 1) return "GLSL450";
        break;
    case K_EXECMODEL:
        switch (v) {
        case 0: return "Vertex";
        case 4: return "Fragment";
        case 5: return "GLCompute";
        }
        break;
    case K_STORAGE:
        switch (v) {
        case 0: return "UniformConstant";
        case 1: return "Input";
        case 2: return "Uniform";
        case 3: return "Output";
        case 4: return "Workgroup";
        case 6: return "Private";
        case 7: return "Function";
        case 9: return "PushConstant";
        case 12: return "StorageBuffer";
        }
        break;
    case K_FNCTL:
        if (v == 0) return "None";
        if (v == 1) return "Inline";
        if (v == 2) return "DontInline";
        break;
    case K_SELCTL:
    case K_LOOPCTL:
        if (v == 0) return "None";
        break;
    case K_DIM:
        switch (v) {
        case 0: return "1D";
        case 1: return "2D";
        case 2: return "3D";
        case 3: return "Cube";
        }
        break;
    case K_IMGFMT:
        if (v == 0) return "Unknown";
        break;
    case K_SRCLANG:
        if (v == 2) return "GLSL";
        if (v == 1) return "ESSL";
        break;
    default:
        break;
    }
    return nullptr;
}

const char *builtin_name(uint32_t v) {
    switch (v) {
    case 0: return "Position";
    case 15: return "FragCoord";
    case 42: return "VertexIndex";
    case 43: return "InstanceIndex";
    }
    return nullptr;
}

// Decoration name and its number of literal parameters (-1: builtin enum).
const char *deco_name(uint32_t v, int *nparams) {
    *nparams = 0;
    switch (v) {
    case 1: *nparams = 1; return "SpecId";
    case 2: return "Block";
    case 3: return "BufferBlock";
    case 4: return "RowMajor";
    case 5: return "ColMajor";
    case 6: *nparams = 1; return "ArrayStride";
    case 7: *nparams = 1; return "MatrixStride";
    case 11: *nparams = -1; return "BuiltIn";
    case 14: return "Flat";
    case 30: *nparams = 1; return "Location";
    case 31: *nparams = 1; return "Component";
    case 33: *nparams = 1; return "Binding";
    case 34: *nparams = 1; return "DescriptorSet";
    case 35: *nparams = 1; return "Offset";
    }
    return nullptr;
}

const char *exec_mode_name(uint32_t v) {
    switch (v) {
    case 7: return "OriginUpperLeft";
    case 8: return "OriginLowerLeft";
    case 9: return "EarlyFragmentTests";
    case 17: return "LocalSize";
    }
    return nullptr;
}

std::string num_or_enum(Kind k, uint32_t v) {
    const char *n = enum_name(k, v);
    return n ? std::string(n) : std::to_string(v);
}

std::string image_operands(uint32_t mask) {
    static const char *names[] = {"Bias", "Lod", "Grad", "ConstOffset", "Offset", "ConstOffsets", "Sample", "MinLod"};
    if (mask == 0) return "None";
    std::string s;
    for (int b = 0; b < 8; ++b)
        if (mask >> b & 1u) {
            if (!s.empty()) s += "|";
            s += names[b];
        }
    return s;
}

}  // namespace

std::string spirv_disassemble(const SpirvModule &m) {
    std::string out;
    char buf[128];
    std::snprintf(buf, sizeof buf, "; SPIR-V\n; Version: %u.%u\n; Generator: 0x%x\n; Bound: %u\n; Schema: %u\n",
                  (m.version >> 16) & 0xFF, (m.version >> 8) & 0xFF, m.generator, m.bound, m.schema);
    out += buf;
    const auto &tab = op_table();
    std::unordered_map<uint32_t, uint32_t> int_width, type_of;  // integer type -> bits; value -> type
    for (const auto &ins : m.instructions) {
        auto it = tab.find(ins.opcode);
        std::string result, text;
        const auto &ops = ins.operands;
        if (ins.opcode == 21 && ops.size() >= 2) int_width[ops[0]] = ops[1];  // OpTypeInt
        if (it != tab.end() && it->second.kinds.size() >= 2 && it->second.kinds[0] == K_RTYPE &&
            it->second.kinds[1] == K_RESULT && ops.size() >= 2)
            type_of[ops[1]] = ops[0];
        size_t i = 0;
        auto id = [&](uint32_t v) { return "%" + std::to_string(v); };
        if (it == tab.end()) {
            text = "Op" + std::to_string(ins.opcode);
            for (; i < ops.size(); ++i) text += " " + std::to_string(ops[i]);
        } else {
            text = it->second.name;
            for (Kind k : it->second.kinds) {
                if (i >= ops.size() && k != K_IDS && k != K_LITS) break;
                switch (k) {
                case K_RESULT: result = id(ops[i++]); break;
                case K_RTYPE:
                case K_ID: text += " " + id(ops[i++]); break;
                case K_LIT: text += " " + std::to_string(ops[i++]); break;
                case K_STR: {
                    size_t nx;
                    text += " \"" + spirv_string(ops, i, &nx) + "\"";
                    i = nx;
                    break;
                }
                case K_CONST:
                    for (; i < ops.size(); ++i) {
                        std::snprintf(buf, sizeof buf, " 0x%08X", ops[i]);
                        text += buf;
                    }
                    break;
                case K_EXECMODE: {
                    const char *n = exec_mode_name(ops[i]);
                    text += " " + (n ? std::string(n) : std::to_string(ops[i]));
                    ++i;
                    for (; i < ops.size(); ++i) text += " " + std::to_string(ops[i]);
                    break;
                }
                case K_DECO: {
                    int np;
                    const char *n = deco_name(ops[i], &np);
                    text += " " + (n ? std::string(n) : std::to_string(ops[i]));
                    ++i;
                    if (np == -1 && i < ops.size()) {
                        const char *b = builtin_name(ops[i]);
                        text += " " + (b ? std::string(b) : std::to_string(ops[i]));
                        ++i;
                    }
                    for (; i < ops.size(); ++i) text += " " + std::to_string(ops[i]);
                    break;
                }
                case K_IMGOPS:
                    if (i < ops.size()) {
                        text += " " + image_operands(ops[i++]);
                        for (; i < ops.size(); ++i) text += " " + id(ops[i]);
                    }
                    break;
                case K_IDS:
                    for (; i < ops.size(); ++i) text += " " + id(ops[i]);
                    break;
                case K_LITS:
                    for (; i < ops.size(); ++i) text += " " + std::to_string(ops[i]);
                    break;
                case K_EXTINST: {
                    text += " " + id(ops[i++]);
                    if (i < ops.size()) {
                        const char *n = glsl_name(ops[i]);
                        text += " " + (n ? std::string(n) : std::to_string(ops[i]));
                        ++i;
                    }
                    for (; i < ops.size(); ++i) text += " " + id(ops[i]);
                    break;
                }
                case K_SPECOP: {
                    const char *n = spec_op_name(ops[i]);
                    text += " " + (n ? std::string(n) : std::to_string(ops[i]));
                    ++i;
                    for (; i < ops.size(); ++i) text += " " + id(ops[i]);
                    break;
                }
                case K_SWITCH: {
                    // case literals have the selector's width (one or two words)
                    const uint32_t sel = ops[i];
                    auto ty = type_of.find(sel);
                    const bool wide = ty != type_of.end() && int_width.count(ty->second) &&
                                      int_width[ty->second] == 64;
                    text += " " + id(ops[i++]);
                    if (i < ops.size()) text += " " + id(ops[i++]);
                    const size_t step = wide ? 3 : 2;
                    for (; i + step - 1 < ops.size(); i += step) {
                        const uint64_t v = wide ? ((uint64_t)ops[i + 1] << 32) | ops[i] : ops[i];
                        text += " " + std::to_string(v) + " " + id(ops[i + step - 1]);
                    }
                    break;
                }
                default:
                    text += " " + num_or_enum(k, ops[i++]);
                    break;
                }
            }
        }
        std::string line;
        if (!result.empty()) {
            std::string r = result + " = ";
            if (r.size() < 15) line.assign(15 - r.size(), ' ');
            line += r;
        } else {
            line.assign(15, ' ');
        }
        line += text;
        std::snprintf(buf, sizeof buf, " ; 0x%08x\n", ins.word_offset * 4);
        line += buf;
        out += line;
    }
    return out;
}

}  // namespace kz
