"""GLSL.std.450 functions and SPIR-V instructions beyond the ones the reference's
translator takes (it lists them as not yet translated,
vulkan-shader-compiler/src/parse/unimplemented_instructions.rs): the
trigonometric and hyperbolic family, Fma, FaceForward, Refract, NMin/NMax/NClamp,
SSign, FindILsb/FindSMsb/FindUMsb, Determinant, MatrixInverse, the
Pack/Unpack {Unorm4x8, Snorm4x8, Unorm2x16, Snorm2x16, Half2x16} family,
Modf, Frexp and Ldexp (their second results through Function variables),
ModfStruct and FrexpStruct (struct-valued results), and
OpBitCount, OpBitReverse, OpBitField{Insert,SExtract,UExtract}, OpTranspose,
OpMatrixTimesScalar, OpVectorTimesMatrix, OpMatrixTimesMatrix, OpOuterProduct.

Each fragment shader is translated, compiled offline (also through the IR text
round trip), and rendered over the oracle's random-triangle scene. The float
shaders are checked against the same formulas applied in numpy (float64) to
the oracle's interpolated colour, on covered pixels, within 2 LSB (the colour
is quantised to 8 bits before the formula: <= 1/2 LSB through slopes of at
most ~1.2, interpolation <= 1 LSB, rounding). The integer shader works on the
exact pixel coordinate (gl_FragCoord) and is checked bit for bit against the
same bit operations restated in Python; so are the packing shader and the
Modf/Frexp/Ldexp shader, whose inputs are chosen so that every result is exact."""
import math

import numpy as np
import pytest

from kazan_amd.scenes import generators as g
from kazan_amd.scenes import renderer as rd
from kazan_amd.scenes import shaders
from kazan_amd.scenes import spirv as sp


def _color_fs(build):
    """A fragment shader o = build(f, m, x, y, z, w) of the interpolated colour."""

    def body(m, var):
        f32, v4 = m.t_float(), m.t_vec(4)
        col_in = var(sp.StorageInput, v4, "v_color", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        c = f.load(v4, col_in)
        x, y, z, w = (f.extract(f32, c, i) for i in range(4))
        f.store(out, f.construct(v4, *build(f, m, x, y, z, w)))
        f.ret()
        return f

    return shaders._fragment_module(body).to_bytes()


def _ops(f, m):
    f32 = m.t_float()
    k = m.c_float
    add = lambda a, b: f.fadd(f32, a, b)
    mul = lambda a, b: f.fmul(f32, a, b)
    e = lambda inst, *a: f.ext(f32, inst, *a)
    return f32, k, add, mul, e


def fs_trig():
    # o = (tan(0.4x), asin(y/2) + acos(y/2)/4, 0.8 atan(z) + 0.2 atan2(z, w + 1),
    #      tanh(degrees(radians(w)))/2 + (cosh(w) - 1)/4 + sinh(w)/10)
    def build(f, m, x, y, z, w):
        f32, k, add, mul, e = _ops(f, m)
        tx = e(sp.GLSL_Tan, mul(k(0.4), x))
        hy = mul(k(0.5), y)
        ty = add(e(sp.GLSL_Asin, hy), mul(k(0.25), e(sp.GLSL_Acos, hy)))
        tz = add(mul(k(0.8), e(sp.GLSL_Atan, z)), mul(k(0.2), e(sp.GLSL_Atan2, z, add(w, k(1.0)))))
        dw = e(sp.GLSL_Degrees, e(sp.GLSL_Radians, w))
        tw = add(add(mul(k(0.5), e(sp.GLSL_Tanh, dw)), mul(k(0.25), f.fsub(f32, e(sp.GLSL_Cosh, w), k(1.0)))),
                 mul(k(0.1), e(sp.GLSL_Sinh, w)))
        return tx, ty, tz, tw

    return _color_fs(build)


def _trig_expected(x, y, z, w):
    return np.stack([np.tan(0.4 * x), np.arcsin(y / 2) + np.arccos(y / 2) / 4,
                     0.8 * np.arctan(z) + 0.2 * np.arctan2(z, w + 1),
                     np.tanh(w) / 2 + (np.cosh(w) - 1) / 4 + np.sinh(w) / 10], -1)


def fs_misc():
    # x: asinh(x - 0.5)/2 + 0.5
    # y: acosh(y + 1.5)/4 + atanh((y - 0.5)/2)/2 + 0.3
    # z: fma(z, 0.5, 0.1) + nclamp(w, 0.2, 0.8)/4 + nmax(z, 0.5)/10 + nmin(x, y)/20
    # w: 0.5 + refract(I, N, 0.5 + w/4).x / 2 + faceforward(N, I, (0, 0, z - 0.5)).z / 5
    #      + refract(I2, N, 2 + w).x + refract(I2, N, 2 + w).z
    #    I = (x - 0.5, y - 0.5, -1), I2 = (x - 0.5, y - 0.5, -0.5), N = (0, 0, 1):
    #    the second refraction is total (k = 1 - eta^2 * 0.75 < 0), so it is 0
    def build(f, m, x, y, z, w):
        f32, k, add, mul, e = _ops(f, m)
        v3 = m.t_vec(3)
        bx = add(mul(k(0.5), e(sp.GLSL_Asinh, f.fsub(f32, x, k(0.5)))), k(0.5))
        by = add(add(mul(k(0.25), e(sp.GLSL_Acosh, add(y, k(1.5)))),
                     mul(k(0.5), e(sp.GLSL_Atanh, mul(k(0.5), f.fsub(f32, y, k(0.5)))))), k(0.3))
        bz = add(add(e(sp.GLSL_Fma, z, k(0.5), k(0.1)), mul(k(0.25), e(sp.GLSL_NClamp, w, k(0.2), k(0.8)))),
                 add(mul(k(0.1), e(sp.GLSL_NMax, z, k(0.5))), mul(k(0.05), e(sp.GLSL_NMin, x, y))))
        cx, cy = f.fsub(f32, x, k(0.5)), f.fsub(f32, y, k(0.5))
        i1 = f.construct(v3, cx, cy, k(-1.0))
        i2 = f.construct(v3, cx, cy, k(-0.5))
        n = m.c_vec(0.0, 0.0, 1.0)
        r1 = f.ext(v3, sp.GLSL_Refract, i1, n, add(k(0.5), mul(k(0.25), w)))
        r2 = f.ext(v3, sp.GLSL_Refract, i2, n, add(k(2.0), w))
        ff = f.ext(v3, sp.GLSL_FaceForward, n, i1, f.construct(v3, k(0.0), k(0.0), f.fsub(f32, z, k(0.5))))
        bw = add(add(k(0.5), mul(k(0.5), f.extract(f32, r1, 0))), mul(k(0.2), f.extract(f32, ff, 2)))
        bw = add(bw, add(f.extract(f32, r2, 0), f.extract(f32, r2, 2)))
        return bx, by, bz, bw

    return _color_fs(build)


def _misc_expected(x, y, z, w):
    bx = np.arcsinh(x - 0.5) / 2 + 0.5
    by = np.arccosh(y + 1.5) / 4 + np.arctanh((y - 0.5) / 2) / 2 + 0.3
    bz = z * 0.5 + 0.1 + np.clip(w, 0.2, 0.8) / 4 + np.maximum(z, 0.5) / 10 + np.minimum(x, y) / 20
    # refract with N = (0,0,1), I.z = -1: dot = -1, k = 1, result.x = eta * I.x
    bw = 0.5 + (0.5 + w / 4) * (x - 0.5) / 2 + np.where(z > 0.5, 1.0, -1.0) / 5
    return np.stack([bx, by, bz, bw], -1)


def fs_bits():
    # from the pixel's integer coordinate (ix, iy) = floor(gl_FragCoord.xy):
    #   h = ix * 0x9E3779B1 ^ iy * 0x85EBCA77 (uint), s = int(h)
    #   R = bitCount(h) + 7 * (findLSB(h) & 15)
    #   G = findMSB(h >> (ix & 15)) + 1 + 8 * ((findMSB(s) + 1) & 15)
    #   B = bitfieldExtract(h, ix & 7, 4) + 16 * (bitfieldExtract(s, iy & 15, 4) + 8)
    #   A = (bitfieldInsert(bitfieldReverse(h) >> 24, ix, 2, 3) & 0xFC) ^ (sign(s) + 1)
    #   o = vec4(R, G, B, A) / 255
    def body(m, var):
        f32, v4, u32, i32 = m.t_float(), m.t_vec(4), m.t_int(False), m.t_int(True)
        fc_in = var(sp.StorageInput, v4, "fc", builtin=sp.BuiltInFragCoord)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        fc = f.load(v4, fc_in)
        ix = f.op(sp.OpConvertFToU, u32, f.extract(f32, fc, 0))
        iy = f.op(sp.OpConvertFToU, u32, f.extract(f32, fc, 1))
        cu = lambda v: m.c_int(v, signed=False)
        ci = lambda v: m.c_int(v, signed=True)
        h = f.binop(sp.OpBitwiseXor, u32, f.binop(sp.OpIMul, u32, ix, cu(0x9E3779B1)),
                    f.binop(sp.OpIMul, u32, iy, cu(0x85EBCA77)))
        s = f.op(sp.OpBitcast, i32, h)
        band = lambda a, b: f.binop(sp.OpBitwiseAnd, u32, a, b)
        iadd = lambda a, b: f.binop(sp.OpIAdd, u32, a, b)
        imul = lambda a, b: f.binop(sp.OpIMul, u32, a, b)
        r = iadd(f.op(sp.OpBitCount, u32, h), imul(cu(7), band(f.ext(u32, sp.GLSL_FindILsb, h), cu(15))))
        sh = f.binop(sp.OpShiftRightLogical, u32, h, band(ix, cu(15)))
        gch = iadd(iadd(f.ext(u32, sp.GLSL_FindUMsb, sh), cu(1)),
                   imul(cu(8), band(iadd(f.ext(i32, sp.GLSL_FindSMsb, s), cu(1)), cu(15))))
        ue = f.op(sp.OpBitFieldUExtract, u32, h, band(ix, cu(7)), cu(4))
        se = f.op(sp.OpBitFieldSExtract, i32, s, band(iy, cu(15)), ci(4))
        b = iadd(ue, imul(cu(16), iadd(se, cu(8))))
        rev = f.binop(sp.OpShiftRightLogical, u32, f.op(sp.OpBitReverse, u32, h), cu(24))
        ins = f.op(sp.OpBitFieldInsert, u32, rev, ix, cu(2), cu(3))
        a = f.binop(sp.OpBitwiseXor, u32, band(ins, cu(0xFC)), iadd(f.ext(i32, sp.GLSL_SSign, s), ci(1)))
        inv = m.c_float(1.0 / 255.0)
        chans = [f.fmul(f32, f.op(sp.OpConvertUToF, f32, v), inv) for v in (r, gch, b, a)]
        f.store(out, f.construct(v4, *chans))
        f.ret()
        return f

    return shaders._fragment_module(body).to_bytes()


def _lsb(v):
    return -1 if v == 0 else (v & -v).bit_length() - 1


def _umsb(v):
    return v.bit_length() - 1  # -1 for 0


def _smsb(v):  # v as a signed 32-bit value
    return _umsb(~v & 0xFFFFFFFF if v & 0x80000000 else v)


def _sext(v, off, count):
    u = (v >> off) & ((1 << count) - 1)
    return u - (1 << count) if u >> (count - 1) else u


def _bits_expected(ix, iy):
    h = (ix * 0x9E3779B1 ^ iy * 0x85EBCA77) & 0xFFFFFFFF
    r = bin(h).count("1") + 7 * (_lsb(h) & 15)
    gch = _umsb(h >> (ix & 15)) + 1 + 8 * ((_smsb(h) + 1) & 15)
    b = ((h >> (ix & 7)) & 15) + 16 * (_sext(h, iy & 15, 4) + 8)
    rev = int(f"{h:032b}"[::-1], 2) >> 24
    ins = (rev & ~(7 << 2)) | ((ix & 7) << 2)
    sign = 0 if h == 0 else (-1 if h & 0x80000000 else 1)
    a = (ins & 0xFC) ^ (sign + 1)
    return r, gch, b, a


def fs_matrix():
    # A = mat3((1 + x/5, y/10, z/10), (w/10, 1 + y/5, x/10), (z/10, w/10, 1 + z/5)) (columns)
    # R = (det(A) - 1) / 2
    # G = inverse(A)[0][0] / 2 + transpose(A)[1][0] + inverse(A)[1][0] / 5 + 0.25
    # B = (vec3(x, y, z) * A).y / 4 + (A * A)[2][2] / 10 + (A * 0.5)[0][1] * 0.3
    # W = outerProduct(vec3(x, y, z), vec3(w, 0.5, 1))[1][2] / 2 + (det(M4) - 1) / 4 + det(M2) / 10
    #   M4 = mat4((1 + x/10, y/20, 0, z/20), (w/20, 1 + y/10, x/20, 0), (0, z/20, 1 + z/10, w/20),
    #             (x/20, 0, y/20, 1 + w/10)), M2 = mat2((1 + x, y/2), (z/2, 1 + w))
    def build(f, m, x, y, z, w):
        f32, k, add, mul, e = _ops(f, m)
        v2, v3, v4 = m.t_vec(2), m.t_vec(3), m.t_vec(4)
        m2, m3, m4 = m.t_mat(2), m.t_mat(3), m.t_mat(4)
        s = lambda c, v: mul(k(c), v)
        one = lambda c, v: add(k(1.0), mul(k(c), v))
        a = f.construct(m3, f.construct(v3, one(0.2, x), s(0.1, y), s(0.1, z)),
                        f.construct(v3, s(0.1, w), one(0.2, y), s(0.1, x)),
                        f.construct(v3, s(0.1, z), s(0.1, w), one(0.2, z)))
        r = mul(k(0.5), f.fsub(f32, e(sp.GLSL_Determinant, a), k(1.0)))
        ai = f.ext(m3, sp.GLSL_MatrixInverse, a)
        at = f.op(sp.OpTranspose, m3, a)
        gch = add(add(mul(k(0.5), f.extract(f32, ai, 0, 0)), f.extract(f32, at, 1, 0)),
                  add(mul(k(0.2), f.extract(f32, ai, 1, 0)), k(0.25)))
        vm = f.op(sp.OpVectorTimesMatrix, v3, f.construct(v3, x, y, z), a)
        aa = f.op(sp.OpMatrixTimesMatrix, m3, a, a)
        ah = f.op(sp.OpMatrixTimesScalar, m3, a, k(0.5))
        b = add(add(mul(k(0.25), f.extract(f32, vm, 1)), mul(k(0.1), f.extract(f32, aa, 2, 2))),
                mul(k(0.3), f.extract(f32, ah, 0, 1)))
        op = f.op(sp.OpOuterProduct, m3, f.construct(v3, x, y, z), f.construct(v3, w, k(0.5), k(1.0)))
        z0 = k(0.0)
        q = f.construct(m4, f.construct(v4, one(0.1, x), s(0.05, y), z0, s(0.05, z)),
                        f.construct(v4, s(0.05, w), one(0.1, y), s(0.05, x), z0),
                        f.construct(v4, z0, s(0.05, z), one(0.1, z), s(0.05, w)),
                        f.construct(v4, s(0.05, x), z0, s(0.05, y), one(0.1, w)))
        p = f.construct(m2, f.construct(v2, add(k(1.0), x), s(0.5, y)), f.construct(v2, s(0.5, z), add(k(1.0), w)))
        wch = add(add(mul(k(0.5), f.extract(f32, op, 1, 2)), mul(k(0.25), f.fsub(f32, e(sp.GLSL_Determinant, q), k(1.0)))),
                  mul(k(0.1), e(sp.GLSL_Determinant, p)))
        return r, gch, b, wch

    return _color_fs(build)


def _matrix_expected(x, y, z, w):
    def mat(cols):  # columns -> row-major array [..., r, c]
        return np.stack([np.stack(c, -1) for c in cols], -1)

    zero = np.zeros_like(x)
    a = mat([(1 + x / 5, y / 10, z / 10), (w / 10, 1 + y / 5, x / 10), (z / 10, w / 10, 1 + z / 5)])
    ai = np.linalg.inv(a)
    r = (np.linalg.det(a) - 1) / 2
    gch = ai[..., 0, 0] / 2 + a[..., 1, 0] + ai[..., 0, 1] / 5 + 0.25  # transpose(A) col 1 row 0 = A row 1 col 0
    v = np.stack([x, y, z], -1)
    vm = np.einsum("...r,...rc->...c", v, a)
    aa = a @ a
    b = vm[..., 1] / 4 + aa[..., 2, 2] / 10 + (a[..., 1, 0] * 0.5) * 0.3
    q = mat([(1 + x / 10, y / 20, zero, z / 20), (w / 20, 1 + y / 10, x / 20, zero), (zero, z / 20, 1 + z / 10, w / 20),
             (x / 20, zero, y / 20, 1 + w / 10)])
    p = mat([(1 + x, y / 2), (z / 2, 1 + w)])
    wch = (z * 0.5) / 2 + (np.linalg.det(q) - 1) / 4 + np.linalg.det(p) / 10
    return np.stack([r, gch, b, wch], -1)


def fs_pack():
    # (ix, iy) = floor(gl_FragCoord.xy); a = ix & 255, b = iy & 255, c = (7 ix + iy) & 255
    #   p  = packUnorm4x8(vec4(a / 255, b / 255, c / 255, 0.5)) -> bytes a, b, c, 128 (127.5 to even)
    #   q  = packSnorm4x8(vec4(-1, -0.5, 0.25, 1.5))        -> 0x81, 0xC0, 0x20, 0x7F
    #   ps = packSnorm2x16(vec2(-0.5, 0.5))                 -> 0x4000C000
    #   h  = unpackHalf2x16(packHalf2x16(vec2(a / 8, -b)))  -> (a / 8, -b), exact
    #   u2 = unpackUnorm2x16(packUnorm2x16(vec2(b / 255, 0.25)))
    #   R = (p >> 8 (ix & 3)) & 255
    #   G = ((q >> 8 (iy & 3)) & 255) ^ (ps >> 24) ^ ((ps >> 8) & 255)
    #   B = (uint(h.x * 8) ^ uint(-h.y)) & 255
    #   A = (uint(roundEven(unpackUnorm4x8(p).z * 255)) + uint(roundEven(u2.x * 255))
    #        + uint(roundEven(unpackSnorm4x8(q).y * 127) + 64) + uint(roundEven(unpackSnorm2x16(ps).y * 2))) & 255
    #   o = vec4(R, G, B, A) / 255
    def body(m, var):
        f32, v2, v4, u32, i32 = m.t_float(), m.t_vec(2), m.t_vec(4), m.t_int(False), m.t_int(True)
        fc_in = var(sp.StorageInput, v4, "fc", builtin=sp.BuiltInFragCoord)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        fc = f.load(v4, fc_in)
        ix = f.op(sp.OpConvertFToU, u32, f.extract(f32, fc, 0))
        iy = f.op(sp.OpConvertFToU, u32, f.extract(f32, fc, 1))
        cu, k = (lambda v: m.c_int(v, signed=False)), m.c_float
        band = lambda x, y: f.binop(sp.OpBitwiseAnd, u32, x, y)
        iadd = lambda x, y: f.binop(sp.OpIAdd, u32, x, y)
        shr = lambda x, y: f.binop(sp.OpShiftRightLogical, u32, x, y)
        xor = lambda x, y: f.binop(sp.OpBitwiseXor, u32, x, y)
        tof = lambda x: f.op(sp.OpConvertUToF, f32, x)
        tou = lambda x: f.op(sp.OpConvertFToU, u32, x)
        rnd = lambda x: f.ext(f32, sp.GLSL_RoundEven, x)
        a, b = band(ix, cu(255)), band(iy, cu(255))
        c = band(iadd(f.binop(sp.OpIMul, u32, ix, cu(7)), iy), cu(255))
        inv = k(1.0 / 255.0)
        p = f.ext(u32, sp.GLSL_PackUnorm4x8, f.construct(v4, f.fdiv(f32, tof(a), k(255.0)), f.fdiv(f32, tof(b), k(255.0)),
                                                         f.fdiv(f32, tof(c), k(255.0)), k(0.5)))
        q = f.ext(u32, sp.GLSL_PackSnorm4x8, m.c_vec(-1.0, -0.5, 0.25, 1.5))
        ps = f.ext(u32, sp.GLSL_PackSnorm2x16, m.c_vec(-0.5, 0.5))
        h = f.ext(v2, sp.GLSL_UnpackHalf2x16, f.ext(u32, sp.GLSL_PackHalf2x16, f.construct(
            v2, f.fmul(f32, tof(a), k(0.125)), f.op(sp.OpFNegate, f32, tof(b)))))
        u2 = f.ext(v2, sp.GLSL_UnpackUnorm2x16, f.ext(u32, sp.GLSL_PackUnorm2x16, f.construct(
            v2, f.fdiv(f32, tof(b), k(255.0)), k(0.25))))
        eight = lambda x: f.binop(sp.OpShiftLeftLogical, u32, x, cu(3))
        r = band(shr(p, eight(band(ix, cu(3)))), cu(255))
        g_ = xor(xor(band(shr(q, eight(band(iy, cu(3)))), cu(255)), shr(ps, cu(24))), band(shr(ps, cu(8)), cu(255)))
        bl = band(xor(tou(f.fmul(f32, f.extract(f32, h, 0), k(8.0))), tou(f.op(sp.OpFNegate, f32, f.extract(f32, h, 1)))),
                  cu(255))
        t1 = tou(rnd(f.fmul(f32, f.extract(f32, f.ext(v4, sp.GLSL_UnpackUnorm4x8, p), 2), k(255.0))))
        t2 = tou(rnd(f.fmul(f32, f.extract(f32, u2, 0), k(255.0))))
        t3 = tou(f.fadd(f32, rnd(f.fmul(f32, f.extract(f32, f.ext(v4, sp.GLSL_UnpackSnorm4x8, q), 1), k(127.0))), k(64.0)))
        t4 = tou(rnd(f.fmul(f32, f.extract(f32, f.ext(v2, sp.GLSL_UnpackSnorm2x16, ps), 1), k(2.0))))
        al = band(iadd(iadd(t1, t2), iadd(t3, t4)), cu(255))
        chans = [f.fmul(f32, tof(v), inv) for v in (r, g_, bl, al)]
        f.store(out, f.construct(v4, *chans))
        f.ret()
        return f

    return shaders._fragment_module(body).to_bytes()


def _pack_expected(ix, iy):
    a, b, c = ix & 255, iy & 255, (7 * ix + iy) & 255
    r = (a, b, c, 128)[ix & 3]
    g_ = (0x81, 0xC0, 0x20, 0x7F)[iy & 3] ^ 0x40 ^ 0xC0
    bl = (a ^ b) & 255
    # unpackUnorm4x8(p).z * 255 -> c; u2.x * 255 -> b (16-bit field round(b / 255 * 65535) / 65535 * 255);
    # unpackSnorm4x8(q).y * 127 = -64 -> +64 = 0; unpackSnorm2x16(ps).y * 2 = 16384 / 32767 * 2 -> 1
    al = (c + b + 0 + 1) & 255
    return r, g_, bl, al


def fs_frexp():
    # (ix, iy) = floor(gl_FragCoord.xy); a = ix & 255, b = iy & 7; v = a / 16 + 1/8 (exact)
    #   fr = modf(v, wh); m = frexp(v, e)
    #   R = uint(fr * 16) + 16 * (uint(wh) & 15)
    #   G = uint(e + 2) + 8 * (uint(m * 16) - 8)
    #   B = uint(ldexp(m, e) * 16) & 255          (= a + 2)
    #   A = uint(ldexp(1.0, b) + ldexp(-0.5, -1)) (= 2^b - 0.25, truncated)
    #   R, G and B also add the differences (0) between modfStruct(v) / frexpStruct(v) and the above
    #   o = vec4(R, G, B, A) / 255
    def body(m, var):
        f32, v4, u32, i32 = m.t_float(), m.t_vec(4), m.t_int(False), m.t_int(True)
        fc_in = var(sp.StorageInput, v4, "fc", builtin=sp.BuiltInFragCoord)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        wh_var, e_var = f.var(f32), f.var(i32)
        fc = f.load(v4, fc_in)
        ix = f.op(sp.OpConvertFToU, u32, f.extract(f32, fc, 0))
        iy = f.op(sp.OpConvertFToU, u32, f.extract(f32, fc, 1))
        cu, k = (lambda v: m.c_int(v, signed=False)), m.c_float
        band = lambda x, y: f.binop(sp.OpBitwiseAnd, u32, x, y)
        iadd = lambda x, y: f.binop(sp.OpIAdd, u32, x, y)
        imul = lambda x, y: f.binop(sp.OpIMul, u32, x, y)
        tou = lambda x: f.op(sp.OpConvertFToU, u32, x)
        a, b = band(ix, cu(255)), band(iy, cu(7))
        v = f.fadd(f32, f.fmul(f32, f.op(sp.OpConvertUToF, f32, a), k(0.0625)), k(0.125))
        fr = f.ext(f32, sp.GLSL_Modf, v, wh_var)
        mt = f.ext(f32, sp.GLSL_Frexp, v, e_var)
        wh, e = f.load(f32, wh_var), f.load(i32, e_var)
        ms = f.ext(m.t_struct(f32, f32), sp.GLSL_ModfStruct, v)
        fs = f.ext(m.t_struct(f32, i32), sp.GLSL_FrexpStruct, v)
        isub = lambda x, y: f.binop(sp.OpISub, u32, x, y)
        x16 = lambda x: tou(f.fmul(f32, x, k(16.0)))
        d_modf = iadd(isub(x16(f.extract(f32, ms, 0)), x16(fr)), isub(tou(f.extract(f32, ms, 1)), tou(wh)))
        d_exp = isub(f.extract(i32, fs, 1), e)
        d_mant = isub(x16(f.extract(f32, fs, 0)), x16(mt))
        r = iadd(iadd(x16(fr), imul(cu(16), band(tou(wh), cu(15)))), d_modf)
        g_ = iadd(iadd(f.binop(sp.OpIAdd, u32, e, m.c_int(2)), imul(cu(8), isub(x16(mt), cu(8)))), d_exp)
        bl = band(iadd(tou(f.fmul(f32, f.ext(f32, sp.GLSL_Ldexp, mt, e), k(16.0))), d_mant), cu(255))
        al = tou(f.fadd(f32, f.ext(f32, sp.GLSL_Ldexp, k(1.0), b), f.ext(f32, sp.GLSL_Ldexp, k(-0.5), m.c_int(-1))))
        inv = k(1.0 / 255.0)
        chans = [f.fmul(f32, f.op(sp.OpConvertUToF, f32, x), inv) for x in (r, g_, bl, al)]
        f.store(out, f.construct(v4, *chans))
        f.ret()
        return f

    return shaders._fragment_module(body).to_bytes()


def _frexp_expected(ix, iy):
    a, b = ix & 255, iy & 7
    v = a / 16 + 0.125
    wh = math.floor(v)
    mt, e = math.frexp(v)
    return (int((v - wh) * 16) + 16 * (wh & 15), (e + 2) + 8 * (int(mt * 16) - 8), (a + 2) & 255,
            int(2.0 ** b - 0.25))


SHADERS = {"trig": fs_trig, "misc": fs_misc, "bits": fs_bits, "matrix": fs_matrix, "pack": fs_pack,
           "frexp": fs_frexp}


def test_expected_forms_pin_the_numpy_restatements():
    """The numpy restatements against closed forms at points where the
    functions are known (so a wrong sign or index in them fails here)."""
    one = np.ones(1)
    t = _trig_expected(0 * one, 0 * one, 0 * one, 0 * one)[0]
    np.testing.assert_allclose(t, [0.0, math.pi / 8, 0.0, 0.0], atol=1e-12)
    mx = _matrix_expected(0 * one, 0 * one, 0 * one, 0 * one)[0]
    np.testing.assert_allclose(mx, [0.0, 0.75, 0.1, 0.1], atol=1e-12)  # A = I: det 1, inverse I
    # x = 1 only: A = diag(1.2, 1, 1) + 0.1 e(1,2): det 1.2; inverse(A)[0][0] = 1/1.2
    mx = _matrix_expected(one, 0 * one, 0 * one, 0 * one)[0]
    np.testing.assert_allclose(mx[:2], [0.1, 0.5 / 1.2 + 0.25], atol=1e-12)
    assert _bits_expected(0, 0) == (0 + 7 * 15, 0 + 8 * 0, 8 * 16, 1)
    assert _lsb(8) == 3 and _umsb(8) == 3 and _smsb(0xFFFFFFFF) == -1 and _smsb(0xFFFFFFFE) == 0
    assert _sext(0b1000, 0, 4) == -8 and _sext(0b0111, 0, 4) == 7


@pytest.mark.parametrize("round_trip", [False, True])
@pytest.mark.parametrize("name", sorted(SHADERS))
def test_glsl_functions_translate_and_compile(driver, monkeypatch, name, round_trip):
    if round_trip:
        monkeypatch.setenv("KAZAN_IR_ROUND_TRIP", "1")
    monkeypatch.setitem(rd._FS, g.FS_COLOR, SHADERS[name])
    rc, txt = rd.compile_offline(driver, g.random_triangles(64, 32, 32, seed=5), 2)
    assert rc == 0, txt


def _render(ctx, oracle, monkeypatch, fs, seed):
    s = g.random_triangles(300, 96, 64, seed=seed)
    oc, od, _, _ = oracle.render(s)
    monkeypatch.setitem(rd._FS, g.FS_COLOR, fs)
    r = rd.SceneRenderer(ctx, s)
    r.render(1)
    c, d, _ = r.read_back()
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    covered = od < 1.0
    assert covered.sum() > 1000
    return c, oc, covered


def _check_float(c, oc, covered, expected, mask=None):
    inp = oc[covered].astype(np.float64) / 255.0
    want = np.clip(np.rint(expected(*(inp[:, i] for i in range(4))) * 255.0), 0, 255).astype(np.int32)
    diff = np.abs(c[covered].astype(np.int32) - want)
    if mask is not None:
        diff = diff[mask(inp)]
    assert diff.max() <= 2, (diff.max(axis=0), np.argmax(diff, axis=0))


@pytest.mark.gpu
def test_trig_render(ctx, oracle, monkeypatch):
    c, oc, covered = _render(ctx, oracle, monkeypatch, fs_trig, 41)
    _check_float(c, oc, covered, _trig_expected)


@pytest.mark.gpu
def test_misc_render(ctx, oracle, monkeypatch):
    c, oc, covered = _render(ctx, oracle, monkeypatch, fs_misc, 42)
    # FaceForward's choice flips at z = 0.5: leave out pixels whose quantised z is within 3 LSB of it
    _check_float(c, oc, covered, _misc_expected, mask=lambda inp: np.abs(inp[:, 2] - 0.5) > 3 / 255)


@pytest.mark.gpu
def test_bits_render(ctx, oracle, monkeypatch):
    c, oc, covered = _render(ctx, oracle, monkeypatch, fs_bits, 43)
    ys, xs, _ = np.nonzero(covered)  # (H, W, samples)
    want = np.array([_bits_expected(int(x), int(y)) for y, x in zip(ys, xs)], dtype=np.int32)
    np.testing.assert_array_equal(c[covered].astype(np.int32), want)


@pytest.mark.gpu
def test_matrix_render(ctx, oracle, monkeypatch):
    c, oc, covered = _render(ctx, oracle, monkeypatch, fs_matrix, 44)
    _check_float(c, oc, covered, _matrix_expected)


@pytest.mark.gpu
def test_pack_render(ctx, oracle, monkeypatch):
    c, oc, covered = _render(ctx, oracle, monkeypatch, fs_pack, 45)
    ys, xs, _ = np.nonzero(covered)
    want = np.array([_pack_expected(int(x), int(y)) for y, x in zip(ys, xs)], dtype=np.int32)
    np.testing.assert_array_equal(c[covered].astype(np.int32), want)


@pytest.mark.gpu
def test_frexp_render(ctx, oracle, monkeypatch):
    c, oc, covered = _render(ctx, oracle, monkeypatch, fs_frexp, 46)
    ys, xs, _ = np.nonzero(covered)
    want = np.array([_frexp_expected(int(x), int(y)) for y, x in zip(ys, xs)], dtype=np.int32)
    np.testing.assert_array_equal(c[covered].astype(np.int32), want)
