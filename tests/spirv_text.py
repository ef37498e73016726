"""Reassembles SPIR-V words from the parser-dump text format.

The reference's parser tests hold the expected dump of three .spv files
that its tree does not contain. Each dump line carries the instruction's
byte offset, so reassembling the words from the text and checking every
offset (and the total size) pins the reassembly; the dump of the
reassembled words must then reproduce the text exactly.

Only the opcodes, operand kinds and enumerants those dumps (and this
project's tests) use are known here; values follow the SPIR-V 1.3 spec.
"""
import re
import struct

MAGIC = 0x07230203

# opcode: (number, has result type, has result id, operand kinds after them)
# kinds: id, ids (rest), lit, lits (rest), str, enum:<table>, mask:<table>,
# spec_op, ext_inst, const (literal sized by the result type), switch
OPS = {
    "OpSource": (3, False, False, ["enum:SourceLanguage", "lit"]),
    "OpName": (5, False, False, ["id", "str"]),
    "OpString": (7, False, True, ["str"]),
    "OpLine": (8, False, False, ["id", "lit", "lit"]),
    "OpExtInstImport": (11, False, True, ["str"]),
    "OpExtInst": (12, True, True, ["id", "ext_inst", "ids"]),
    "OpMemoryModel": (14, False, False, ["enum:AddressingModel", "enum:MemoryModel"]),
    "OpEntryPoint": (15, False, False, ["enum:ExecutionModel", "id", "str", "ids"]),
    "OpExecutionMode": (16, False, False, ["id", "enum:ExecutionMode", "lits"]),
    "OpCapability": (17, False, False, ["enum:Capability"]),
    "OpTypeVoid": (19, False, True, []),
    "OpTypeInt": (21, False, True, ["lit", "lit"]),
    "OpTypeFloat": (22, False, True, ["lit"]),
    "OpTypeVector": (23, False, True, ["id", "lit"]),
    "OpTypeStruct": (30, False, True, ["ids"]),
    "OpTypePointer": (32, False, True, ["enum:StorageClass", "id"]),
    "OpTypeFunction": (33, False, True, ["id", "ids"]),
    "OpConstant": (43, True, True, ["const"]),
    "OpSpecConstant": (50, True, True, ["const"]),
    "OpSpecConstantOp": (52, True, True, ["spec_op", "ids"]),
    "OpFunction": (54, True, True, ["mask:FunctionControl", "id"]),
    "OpFunctionEnd": (56, False, False, []),
    "OpFunctionCall": (57, True, True, ["id", "ids"]),
    "OpVariable": (59, True, True, ["enum:StorageClass", "ids"]),
    "OpLoad": (61, True, True, ["id", "lits"]),
    "OpStore": (62, False, False, ["id", "id", "lits"]),
    "OpAccessChain": (65, True, True, ["id", "ids"]),
    "OpDecorate": (71, False, False, ["id", "enum:Decoration", "lits"]),
    "OpMemberDecorate": (72, False, False, ["id", "lit", "enum:Decoration", "lits"]),
    "OpCompositeConstruct": (80, True, True, ["ids"]),
    "OpCompositeExtract": (81, True, True, ["id", "lits"]),
    "OpImageSampleImplicitLod": (87, True, True, ["id", "id", "mask:ImageOperands", "ids"]),
    "OpConvertFToU": (109, True, True, ["id"]),
    "OpConvertSToF": (111, True, True, ["id"]),
    "OpUConvert": (113, True, True, ["id"]),
    "OpBitcast": (124, True, True, ["id"]),
    "OpSelectionMerge": (247, False, False, ["id", "mask:SelectionControl"]),
    "OpLabel": (248, False, True, []),
    "OpBranch": (249, False, False, ["id"]),
    "OpSwitch": (251, False, False, ["switch"]),
    "OpReturn": (253, False, False, []),
    "OpReturnValue": (254, False, False, ["id"]),
}

ENUMS = {
    "SourceLanguage": {"GLSL": 2},
    "AddressingModel": {"Logical": 0},
    "MemoryModel": {"GLSL450": 1},
    "ExecutionModel": {"Vertex": 0, "Fragment": 4, "GLCompute": 5},
    "ExecutionMode": {"LocalSize": 17, "OriginUpperLeft": 7},
    "Capability": {"Shader": 1, "Int64": 11},
    "StorageClass": {"UniformConstant": 0, "Input": 1, "Uniform": 2, "Output": 3, "Private": 6, "Function": 7,
                     "PushConstant": 9},
    "Decoration": {"SpecId": 1, "Block": 2, "BuiltIn": 11, "Location": 30},
    "FunctionControl": {"None": 0},
    "SelectionControl": {"None": 0},
    "ImageOperands": {"None": 0, "Bias": 0x1, "Lod": 0x2, "MinLod": 0x80},
}
DECORATION_ARGS = {"BuiltIn": {"Position": 0}}
SPEC_OPS = {"IMul": 132, "SDiv": 135, "BitwiseXor": 198, "BitwiseAnd": 199}
GLSL_STD_450 = {"Cos": 14, "Length": 66}

_LINE = re.compile(r"^\s*(?:%(\d+) = )?(Op\w+)(.*?)(?: ; 0x([0-9a-f]{8}))?$")


def _tokens(s):
    return re.findall(r'"[^"]*"|\S+', s)


def _lit(tok):
    return int(tok, 16) if tok.startswith("0x") else int(tok)


def _string_words(s):
    b = s.encode() + b"\0"
    b += b"\0" * (-len(b) % 4)
    return list(struct.unpack("<%dI" % (len(b) // 4), b))


def assemble(text):
    """(words, [(offset from the dump, offset of the reassembly)])."""
    lines = text.splitlines()
    hdr = {}
    for ln in lines:
        m = re.match(r"; (\w+): (.*)", ln)
        if m:
            hdr[m.group(1)] = m.group(2)
    major, minor = (int(v) for v in hdr["Version"].split("."))
    words = [MAGIC, (major << 16) | (minor << 8), int(hdr["Generator"], 16), int(hdr["Bound"]), int(hdr["Schema"])]
    width = {}    # type id -> scalar width (for literals sized by type)
    type_of = {}  # value id -> type id
    offsets = []
    for ln in lines:
        if not ln.strip() or ln.startswith(";"):
            continue
        m = _LINE.match(ln)
        if not m:
            raise ValueError("cannot parse: %r" % ln)
        rid, name, rest, off = m.group(1), m.group(2), m.group(3), m.group(4)
        num, has_type, has_res, kinds = OPS[name]
        toks = _tokens(rest)
        out = []
        prev_enum = None  # the last enumerant read (decoration arguments depend on it)
        if has_type:
            rtype = int(toks.pop(0)[1:])
            out.append(rtype)
        if has_res:
            out.append(int(rid))
        if name in ("OpTypeInt", "OpTypeFloat"):
            width[int(rid)] = _lit(toks[0])
        for kind in kinds:
            if kind == "id":
                out.append(int(toks.pop(0)[1:]))
            elif kind == "ids":
                out += [int(t[1:]) for t in toks]
                toks = []
            elif kind == "lit":
                out.append(_lit(toks.pop(0)))
            elif kind == "lits":
                # decoration arguments may be enumerants (BuiltIn Position)
                for t in toks:
                    if t in DECORATION_ARGS.get(prev_enum, {}):
                        out.append(DECORATION_ARGS[prev_enum][t])
                    else:
                        out.append(_lit(t))
                toks = []
            elif kind == "str":
                out += _string_words(toks.pop(0)[1:-1])
            elif kind.startswith("enum:"):
                prev_enum = toks.pop(0)
                out.append(ENUMS[kind[5:]][prev_enum])
            elif kind.startswith("mask:"):
                v = 0
                for part in toks.pop(0).split("|"):
                    v |= ENUMS[kind[5:]][part]
                out.append(v)
            elif kind == "spec_op":
                out.append(SPEC_OPS[toks.pop(0)])
            elif kind == "ext_inst":
                out.append(GLSL_STD_450[toks.pop(0)])
            elif kind == "const":
                v = _lit(toks.pop(0))
                out.append(v & 0xFFFFFFFF)
                if width.get(rtype, 32) == 64:
                    out.append(v >> 32)
            elif kind == "switch":
                sel = int(toks.pop(0)[1:])
                out += [sel, int(toks.pop(0)[1:])]
                wide = width.get(type_of.get(sel), 32) == 64
                while toks:
                    v = _lit(toks.pop(0))
                    out += [v & 0xFFFFFFFF, v >> 32] if wide else [v]
                    out.append(int(toks.pop(0)[1:]))
        if has_type and has_res:
            type_of[int(rid)] = rtype
        if off is not None:
            offsets.append((int(off, 16), 4 * len(words)))
        words.append(((len(out) + 1) << 16) | num)
        words += out
    return words, offsets
