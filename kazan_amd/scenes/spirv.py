"""A small SPIR-V 1.0 module assembler.

The image has no glslang or spirv-as, so the scene library writes its shader
modules word by word with this builder. Opcode and enumerant values are the
ones the SPIR-V 1.0 specification assigns (the reference decodes the same
binary form in spirv-parser/src/generated_parser.rs; MAGIC_NUMBER at :240).
"""
import struct

MAGIC = 0x07230203
VERSION_1_0 = 0x00010000

# opcodes
OpName = 5
OpMemberName = 6
OpExtInstImport = 11
OpExtInst = 12
OpMemoryModel = 14
OpEntryPoint = 15
OpExecutionMode = 16
OpCapability = 17
OpTypeVoid = 19
OpTypeBool = 20
OpTypeInt = 21
OpTypeFloat = 22
OpTypeVector = 23
OpTypeMatrix = 24
OpTypeImage = 25
OpTypeSampler = 26
OpTypeSampledImage = 27
OpTypeArray = 28
OpTypeStruct = 30
OpTypePointer = 32
OpTypeFunction = 33
OpConstantTrue = 41
OpConstantFalse = 42
OpConstant = 43
OpConstantComposite = 44
OpFunction = 54
OpFunctionParameter = 55
OpFunctionEnd = 56
OpFunctionCall = 57
OpVariable = 59
OpLoad = 61
OpStore = 62
OpAccessChain = 65
OpDecorate = 71
OpMemberDecorate = 72
OpVectorShuffle = 79
OpCompositeConstruct = 80
OpCompositeExtract = 81
OpImageSampleImplicitLod = 87
OpConvertFToS = 110
OpConvertSToF = 111
OpFNegate = 127
OpFAdd = 129
OpFSub = 131
OpFMul = 133
OpFDiv = 136
OpVectorTimesScalar = 142
OpMatrixTimesVector = 145
OpDot = 148
OpFOrdLessThan = 184
OpFOrdGreaterThan = 186
OpSelect = 169
OpPhi = 245
OpLoopMerge = 246
OpSelectionMerge = 247
OpLabel = 248
OpBranch = 249
OpBranchConditional = 250
OpSwitch = 251
OpKill = 252
OpReturn = 253
OpReturnValue = 254
OpIAdd = 128
OpISub = 130
OpIMul = 132
OpUDiv = 134
OpUMod = 137
OpSDiv = 135
OpSMod = 139
OpFMod = 141
OpLogicalAnd = 167
OpLogicalNot = 168
OpIEqual = 170
OpSGreaterThan = 173
OpSLessThan = 177
OpFOrdLessThanEqual = 188
OpFOrdGreaterThanEqual = 190
OpBitwiseAnd = 199
OpShiftRightLogical = 194
OpConvertFToU = 109
OpConvertUToF = 112
OpTypeRuntimeArray = 29
OpControlBarrier = 224
OpAtomicIAdd = 234
OpULessThan = 176
OpUGreaterThan = 172
OpShiftLeftLogical = 196
OpBitwiseXor = 198
OpBitcast = 124
OpTranspose = 84
OpMatrixTimesScalar = 143
OpVectorTimesMatrix = 144
OpMatrixTimesMatrix = 146
OpOuterProduct = 147
OpBitFieldInsert = 201
OpBitFieldSExtract = 202
OpBitFieldUExtract = 203
OpBitReverse = 204
OpBitCount = 205
OpArrayLength = 68

# enumerants
CapabilityShader = 1
AddressingLogical = 0
MemoryGLSL450 = 1
ExecutionModelVertex = 0
ExecutionModelFragment = 4
ExecutionModeOriginUpperLeft = 7
ExecutionModelGLCompute = 5
ExecutionModeLocalSize = 17
StorageWorkgroup = 4
DecorationBufferBlock = 3
BuiltInNumWorkgroups = 24
BuiltInWorkgroupId = 26
BuiltInLocalInvocationId = 27
BuiltInGlobalInvocationId = 28
BuiltInLocalInvocationIndex = 29
ScopeDevice = 1
ScopeWorkgroup = 2
SemanticsAcquireRelease = 0x8
SemanticsUniformMemory = 0x40
SemanticsWorkgroupMemory = 0x100
StorageUniformConstant = 0
StorageInput = 1
StorageUniform = 2
StorageOutput = 3
StorageFunction = 7
StoragePushConstant = 9
DecorationBlock = 2
DecorationColMajor = 5
DecorationArrayStride = 6
DecorationMatrixStride = 7
DecorationBuiltIn = 11
DecorationLocation = 30
DecorationBinding = 33
DecorationDescriptorSet = 34
DecorationOffset = 35
BuiltInPosition = 0
BuiltInFragCoord = 15
BuiltInVertexIndex = 42
BuiltInInstanceIndex = 43
Dim2D = 1
FunctionControlNone = 0

# GLSL.std.450 extended instructions
GLSL_FAbs = 4
GLSL_Floor = 8
GLSL_Ceil = 9
GLSL_Fract = 10
GLSL_FMix = 46
GLSL_Distance = 67
GLSL_Cross = 68
GLSL_Reflect = 71
GLSL_FMin = 37
GLSL_FMax = 40
GLSL_FClamp = 43
GLSL_Sqrt = 31
GLSL_InverseSqrt = 32
GLSL_Length = 66
GLSL_Normalize = 69
GLSL_SSign = 7
GLSL_Radians = 11
GLSL_Degrees = 12
GLSL_Tan = 15
GLSL_Asin = 16
GLSL_Acos = 17
GLSL_Atan = 18
GLSL_Sinh = 19
GLSL_Cosh = 20
GLSL_Tanh = 21
GLSL_Asinh = 22
GLSL_Acosh = 23
GLSL_Atanh = 24
GLSL_Atan2 = 25
GLSL_Determinant = 33
GLSL_MatrixInverse = 34
GLSL_Fma = 50
GLSL_FaceForward = 70
GLSL_Refract = 72
GLSL_FindILsb = 73
GLSL_FindSMsb = 74
GLSL_FindUMsb = 75
GLSL_NMin = 79
GLSL_NMax = 80
GLSL_NClamp = 81
GLSL_RoundEven = 2
GLSL_Modf = 35
GLSL_ModfStruct = 36
GLSL_FrexpStruct = 52
GLSL_Frexp = 51
GLSL_Ldexp = 53
GLSL_PackSnorm4x8 = 54
GLSL_PackUnorm4x8 = 55
GLSL_PackSnorm2x16 = 56
GLSL_PackUnorm2x16 = 57
GLSL_PackHalf2x16 = 58
GLSL_UnpackHalf2x16 = 59
GLSL_UnpackSnorm2x16 = 60
GLSL_UnpackUnorm2x16 = 61
GLSL_UnpackSnorm4x8 = 62
GLSL_UnpackUnorm4x8 = 63


def _string_words(s):
    b = s.encode("utf-8") + b"\0"
    b += b"\0" * ((4 - len(b) % 4) % 4)
    return list(struct.unpack("<%dI" % (len(b) // 4), b))


def _f32_bits(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


class Module:
    """Collects the logical sections of a module and emits its words."""

    def __init__(self):
        self.bound = 1
        self.capabilities = []
        self.ext_imports = []
        self.memory_model = []
        self.entry_points = []
        self.exec_modes = []
        self.debug = []
        self.annotations = []
        self.globals = []  # types, constants, global variables
        self.functions = []
        self._types = {}
        self._consts = {}
        self._glsl = None

    # ids and raw instructions -------------------------------------------
    def id(self):
        i = self.bound
        self.bound += 1
        return i

    @staticmethod
    def _inst(op, *operands):
        words = []
        for o in operands:
            if isinstance(o, (list, tuple)):
                words.extend(o)
            else:
                words.append(o)
        return [((len(words) + 1) << 16) | op] + words

    # module-level --------------------------------------------------------
    def capability(self, cap):
        self.capabilities += self._inst(OpCapability, cap)

    def glsl_std_450(self):
        if self._glsl is None:
            self._glsl = self.id()
            self.ext_imports += self._inst(OpExtInstImport, self._glsl, _string_words("GLSL.std.450"))
        return self._glsl

    def set_memory_model(self):
        self.memory_model = self._inst(OpMemoryModel, AddressingLogical, MemoryGLSL450)

    def entry_point(self, model, fn, name, interface):
        self.entry_points += self._inst(OpEntryPoint, model, fn, _string_words(name), list(interface))

    def execution_mode(self, fn, mode, *lits):
        self.exec_modes += self._inst(OpExecutionMode, fn, mode, list(lits))

    def name(self, target, s):
        self.debug += self._inst(OpName, target, _string_words(s))

    def decorate(self, target, deco, *lits):
        self.annotations += self._inst(OpDecorate, target, deco, list(lits))

    def member_decorate(self, st, member, deco, *lits):
        self.annotations += self._inst(OpMemberDecorate, st, member, deco, list(lits))

    # types (deduplicated) ------------------------------------------------
    def _type(self, key, op, *operands):
        if key not in self._types:
            i = self.id()
            self.globals += self._inst(op, i, *operands)
            self._types[key] = i
        return self._types[key]

    def t_void(self):
        return self._type(("void",), OpTypeVoid)

    def t_bool(self):
        return self._type(("bool",), OpTypeBool)

    def t_int(self, signed=True):
        return self._type(("int", signed), OpTypeInt, 32, 1 if signed else 0)

    def t_float(self):
        return self._type(("float",), OpTypeFloat, 32)

    def t_vec(self, n):
        return self._type(("vec", n), OpTypeVector, self.t_float(), n)

    def t_uvec(self, n):
        return self._type(("uvec", n), OpTypeVector, self.t_int(False), n)

    def t_runtime_array(self, elem):
        """A new runtime array type per call (its ArrayStride is decorated per use)."""
        i = self.id()
        self.globals += self._inst(OpTypeRuntimeArray, i, elem)
        return i

    def t_mat4(self):
        return self._type(("mat4",), OpTypeMatrix, self.t_vec(4), 4)

    def t_mat(self, n):
        """An n x n float matrix (n columns of vec n)."""
        return self.t_mat4() if n == 4 else self._type(("mat", n), OpTypeMatrix, self.t_vec(n), n)

    def t_ptr(self, storage, pointee):
        return self._type(("ptr", storage, pointee), OpTypePointer, storage, pointee)

    def t_fn(self, ret, *params):
        return self._type(("fn", ret) + params, OpTypeFunction, ret, list(params))

    def t_array(self, elem, length_const):
        return self._type(("array", elem, length_const), OpTypeArray, elem, length_const)

    def t_struct(self, *members):
        """A new struct type per call (decorations differ per use)."""
        i = self.id()
        self.globals += self._inst(OpTypeStruct, i, list(members))
        return i

    def t_sampled_image_2d(self):
        img = self._type(("image2d",), OpTypeImage, self.t_float(), Dim2D, 0, 0, 0, 1, 0)
        return self._type(("sampledimage", img), OpTypeSampledImage, img)

    # constants -----------------------------------------------------------
    def c_float(self, v):
        key = ("f", _f32_bits(v))
        if key not in self._consts:
            i = self.id()
            self.globals += self._inst(OpConstant, self.t_float(), i, _f32_bits(v))
            self._consts[key] = i
        return self._consts[key]

    def c_int(self, v, signed=True):
        key = ("i", signed, v)
        if key not in self._consts:
            i = self.id()
            self.globals += self._inst(OpConstant, self.t_int(signed), i, v & 0xFFFFFFFF)
            self._consts[key] = i
        return self._consts[key]

    def c_vec(self, *vals):
        comps = [self.c_float(v) for v in vals]
        key = ("v",) + tuple(comps)
        if key not in self._consts:
            i = self.id()
            self.globals += self._inst(OpConstantComposite, self.t_vec(len(vals)), i, comps)
            self._consts[key] = i
        return self._consts[key]

    # global variables ----------------------------------------------------
    def variable(self, storage, pointee, name=None):
        i = self.id()
        self.globals += self._inst(OpVariable, self.t_ptr(storage, pointee), i, storage)
        if name:
            self.name(i, name)
        return i

    def words(self):
        body = (self.capabilities + self.ext_imports + self.memory_model + self.entry_points +
                self.exec_modes + self.debug + self.annotations + self.globals + self.functions)
        return [MAGIC, VERSION_1_0, 0, self.bound, 0] + body

    def to_bytes(self):
        w = self.words()
        return struct.pack("<%dI" % len(w), *w)


class Function:
    """Emits one function body into the module, in order. `ret` and `params`
    are type ids (default: void main()); self.params holds the parameter
    ids."""

    def __init__(self, module, name="main", ret=None, params=()):
        self.m = module
        self.id = module.id()
        ret = module.t_void() if ret is None else ret
        module.functions += module._inst(OpFunction, ret, self.id, FunctionControlNone, module.t_fn(ret, *params))
        module.name(self.id, name)
        self.params = []
        for p in params:
            pid = module.id()
            module.functions += module._inst(OpFunctionParameter, p, pid)
            self.params.append(pid)
        self.entry_label = self.label()

    def _emit(self, op, *operands):
        self.m.functions += self.m._inst(op, *operands)

    def label(self, i=None):
        """Starts a block: a new label, or the id reserved by new_label()."""
        i = self.m.id() if i is None else i
        self._emit(OpLabel, i)
        return i

    def new_label(self):
        return self.m.id()

    def result_into(self, rid, op, type_id, *operands):
        """An instruction whose result id was reserved earlier (a value a
        phi names before it is defined)."""
        self._emit(op, type_id, rid, *operands)
        return rid

    # structured control flow ---------------------------------------------
    def var(self, pointee):
        """A Function-storage variable (emit these in the first block)."""
        return self._result(OpVariable, self.m.t_ptr(StorageFunction, pointee), StorageFunction)

    def branch(self, target):
        self._emit(OpBranch, target)

    def branch_cond(self, cond, t, f):
        self._emit(OpBranchConditional, cond, t, f)

    def selection_merge(self, merge):
        self._emit(OpSelectionMerge, merge, 0)

    def loop_merge(self, merge, cont):
        self._emit(OpLoopMerge, merge, cont, 0)

    def switch(self, selector, default, cases):
        """cases: [(literal, label)]"""
        self._emit(OpSwitch, selector, default, [w for c in cases for w in c])

    def phi(self, type_id, *pairs):
        """pairs: (value, predecessor label)"""
        return self._result(OpPhi, type_id, [w for p in pairs for w in p])

    def kill(self):
        self._emit(OpKill)

    def call(self, ret_type, fn, *args):
        return self._result(OpFunctionCall, ret_type, fn, list(args))

    def ret_only(self):
        """OpReturn without ending the function (a return inside the body)."""
        self._emit(OpReturn)

    def end(self):
        self._emit(OpFunctionEnd)

    def ret_value(self, v):
        self._emit(OpReturnValue, v)
        self._emit(OpFunctionEnd)

    def _result(self, op, type_id, *operands):
        r = self.m.id()
        self._emit(op, type_id, r, *operands)
        return r

    def load(self, type_id, ptr):
        return self._result(OpLoad, type_id, ptr)

    def store(self, ptr, value):
        self._emit(OpStore, ptr, value)

    def access_chain(self, ptr_type, base, *indices):
        return self._result(OpAccessChain, ptr_type, base, list(indices))

    def extract(self, type_id, composite, *idx):
        return self._result(OpCompositeExtract, type_id, composite, list(idx))

    def construct(self, type_id, *parts):
        return self._result(OpCompositeConstruct, type_id, list(parts))

    def shuffle(self, type_id, a, b, *comps):
        return self._result(OpVectorShuffle, type_id, a, b, list(comps))

    def binop(self, op, type_id, a, b):
        return self._result(op, type_id, a, b)

    def op(self, op, type_id, *operands):
        """Any instruction with a result: op result_type result operands..."""
        return self._result(op, type_id, *operands)

    def fadd(self, t, a, b):
        return self.binop(OpFAdd, t, a, b)

    def fsub(self, t, a, b):
        return self.binop(OpFSub, t, a, b)

    def fmul(self, t, a, b):
        return self.binop(OpFMul, t, a, b)

    def fdiv(self, t, a, b):
        return self.binop(OpFDiv, t, a, b)

    def vec_times_scalar(self, t, v, s):
        return self.binop(OpVectorTimesScalar, t, v, s)

    def mat_times_vec(self, t, m, v):
        return self.binop(OpMatrixTimesVector, t, m, v)

    def dot(self, a, b):
        return self.binop(OpDot, self.m.t_float(), a, b)

    def ext(self, type_id, inst, *args):
        return self._result(OpExtInst, type_id, self.m.glsl_std_450(), inst, list(args))

    def sample(self, type_id, sampled_image, coord):
        return self._result(OpImageSampleImplicitLod, type_id, sampled_image, coord)

    def ret(self):
        self._emit(OpReturn)
        self._emit(OpFunctionEnd)

    def barrier(self, exec_scope, mem_scope, semantics):
        """OpControlBarrier; the operands are constant ids."""
        self._emit(OpControlBarrier, exec_scope, mem_scope, semantics)

    def atomic_iadd(self, type_id, ptr, scope, semantics, value):
        return self._result(OpAtomicIAdd, type_id, ptr, scope, semantics, value)
