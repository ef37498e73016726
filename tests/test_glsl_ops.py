"""GLSL.std.450 FAbs, FMix, Distance, Cross, Reflect, Floor, Ceil and Fract (listed as not yet
translated in the reference, vulkan-shader-compiler/src/parse/unimplemented_instructions.rs)
through SPIR-V -> IR -> HIP: a fragment shader
    c = v_color;
    o = vec4(mix(c.x, c.y, 0.5), 0.5 * distance(c.xy, c.zw),
             0.5 + cross(c.xyz, vec3(0.5)).x, 0.5 * abs(reflect(c.xyz, vec3(0, 0, 1)).z - c.w));
checked against the same formula applied in numpy to the oracle's
interpolated colour (the pass-through shader's output), on covered pixels;
and a second one with floor, ceil and fract away from their steps:
    o = vec4(fract(c.x + 2), floor(c.y + 2) * 0.25, ceil(c.z + 2) - 2.5, 1)."""
import numpy as np
import pytest

from kazan_amd.scenes import generators as g
from kazan_amd.scenes import renderer as rd
from kazan_amd.scenes import shaders
from kazan_amd.scenes import spirv as sp


def fs_glsl_ops():
    def body(m, var):
        f32, v2, v3, v4 = m.t_float(), m.t_vec(2), m.t_vec(3), m.t_vec(4)
        col_in = var(sp.StorageInput, v4, "v_color", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        c = f.load(v4, col_in)
        x, y, z, w = (f.extract(f32, c, i) for i in range(4))
        half = m.c_float(0.5)
        r = f.ext(f32, sp.GLSL_FMix, x, y, half)
        gd = f.fmul(f32, half, f.ext(f32, sp.GLSL_Distance, f.shuffle(v2, c, c, 0, 1), f.shuffle(v2, c, c, 2, 3)))
        c3 = f.shuffle(v3, c, c, 0, 1, 2)
        cr = f.ext(v3, sp.GLSL_Cross, c3, m.c_vec(0.5, 0.5, 0.5))
        b = f.fadd(f32, half, f.extract(f32, cr, 0))
        rf = f.ext(v3, sp.GLSL_Reflect, c3, m.c_vec(0.0, 0.0, 1.0))
        a = f.fmul(f32, half, f.ext(f32, sp.GLSL_FAbs, f.fsub(f32, f.extract(f32, rf, 2), w)))
        f.store(out, f.construct(v4, r, gd, b, a))
        f.ret()
        return f

    return shaders._fragment_module(body).to_bytes()


def fs_round_ops():
    def body(m, var):
        f32, v4 = m.t_float(), m.t_vec(4)
        col_in = var(sp.StorageInput, v4, "v_color", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        f = sp.Function(m)
        c = f.load(v4, col_in)
        x, y, z = (f.extract(f32, c, i) for i in range(3))
        two = m.c_float(2.0)
        r = f.ext(f32, sp.GLSL_Fract, f.fadd(f32, x, two))
        gq = f.fmul(f32, f.ext(f32, sp.GLSL_Floor, f.fadd(f32, y, two)), m.c_float(0.25))
        b = f.fsub(f32, f.ext(f32, sp.GLSL_Ceil, f.fadd(f32, z, two)), m.c_float(2.5))
        f.store(out, f.construct(v4, r, gq, b, m.c_float(1.0)))
        f.ret()
        return f

    return shaders._fragment_module(body).to_bytes()


def _expected(c):
    x, y, z, w = (c[..., i].astype(np.float64) for i in range(4))
    cross_x = y * 0.5 - z * 0.5
    return np.stack([x + (y - x) * 0.5, 0.5 * np.hypot(x - z, y - w), 0.5 + cross_x, 0.5 * np.abs(-z - w)], -1)


@pytest.mark.parametrize("fs", [fs_glsl_ops, fs_round_ops])
def test_glsl_ops_translate_and_compile(driver, monkeypatch, fs):
    monkeypatch.setitem(rd._FS, g.FS_COLOR, fs)
    rc, txt = rd.compile_offline(driver, g.random_triangles(64, 32, 32, seed=5), 2)
    assert rc == 0, txt


@pytest.mark.gpu
def test_round_ops_render(ctx, oracle, monkeypatch):
    s = g.random_triangles(300, 96, 64, seed=34)
    oc, od, _, _ = oracle.render(s)
    monkeypatch.setitem(rd._FS, g.FS_COLOR, fs_round_ops)
    r = rd.SceneRenderer(ctx, s)
    r.render(1)
    c, d, _ = r.read_back()
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    covered = od < 1.0
    got, ref = c[covered].astype(np.int32), oc[covered].astype(np.int32)
    # fract(x + 2) = x except for x within an ulp of 2 below 1; floor(y + 2) = 2; ceil(z + 2) = 3 for z > 0
    assert np.abs(got[:, 0] - ref[:, 0]).max() <= 1
    assert (got[:, 1] == 128).all() and (got[:, 2] == 128).all() and (got[:, 3] == 255).all()


@pytest.mark.gpu
def test_glsl_ops_render(ctx, oracle, monkeypatch):
    s = g.random_triangles(300, 96, 64, seed=33)
    oc, od, _, _ = oracle.render(s)
    monkeypatch.setitem(rd._FS, g.FS_COLOR, fs_glsl_ops)
    r = rd.SceneRenderer(ctx, s)
    r.render(1)
    c, d, _ = r.read_back()
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    covered = od < 1.0
    assert covered.sum() > 1000
    want = np.clip(np.rint(_expected(oc[covered] / 255.0) * 255.0), 0, 255)
    diff = np.abs(c[covered].astype(np.int32) - want.astype(np.int32))
    # input quantisation (<= 1/2 LSB through slopes summing to <= 1) + interpolation (<= 1 LSB) + rounding
    assert diff.max() <= 2, diff.max()
