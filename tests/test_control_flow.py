"""Structured control flow through SPIR-V -> IR -> HIP.

The reference's structuriser is pinned by 17 control-flow-graph templates
(shader-compiler/src/cfg.rs:543-1900: returns, a discard, if/else with and
without arms, switches with breaks, returns and fallthrough in both block
orders, while loops). Each template is restated here as a fragment shader
with the same blocks and branches, whose blocks each append a mark to a path
code (code = code * 7 + k, 32-bit wrapping), taking its condition from the
input colour: cond = v.r < 0.5, selector = int(v.g * 4), loop trip count
from n = int(v.b * 4). A Python model of each template gives the code every
input must produce. On the CPU every template translates and compiles; on
the GPU each renders triangles of flat colours and every pixel must carry
the code its triangle's colour gives (bytes: code & 255, code >> 8 & 255,
code >> 16 & 255). Beyond the 17: a for loop with its own continue target,
break and continue from inside selections, nested loops, a call with a
return value, and the loop Phong shader's phis and run-time indices.
"""
import ctypes

import numpy as np
import pytest

from kazan_amd import vk
from kazan_amd.scenes import generators as g
from kazan_amd.scenes import renderer as rd
from kazan_amd.scenes import shaders
from kazan_amd.scenes import spirv as sp


def wrap32(x):
    return (x + 2**31) % 2**32 - 2**31


class Builder:
    """Blocks of a path-code shader: mark(k), cond(), sel(), step() and
    finish() emit the pieces the templates are made of."""

    def __init__(self, m, f, vin, out):
        self.m, self.f = m, f
        self.f32, self.v4, self.i32, self.b = m.t_float(), m.t_vec(4), m.t_int(), m.t_bool()
        self.out = out
        self.code = f.var(self.i32)
        self.count = f.var(self.i32)
        f.store(self.code, m.c_int(1))
        f.store(self.count, m.c_int(0))
        self.v = f.load(self.v4, vin)

    def L(self):
        return self.f.new_label()

    def mark(self, k):
        f, m = self.f, self.m
        c = f.load(self.i32, self.code)
        f.store(self.code, f.binop(sp.OpIAdd, self.i32, f.binop(sp.OpIMul, self.i32, c, m.c_int(7)), m.c_int(k)))

    def cond(self):
        return self.f.binop(sp.OpFOrdLessThan, self.b, self.f.extract(self.f32, self.v, 0), self.m.c_float(0.5))

    def _quarter(self, lane):
        f = self.f
        x = f.fmul(self.f32, f.extract(self.f32, self.v, lane), self.m.c_float(4.0))
        return f._result(sp.OpConvertFToS, self.i32, x)

    def sel(self):
        return self._quarter(1)

    def step(self):
        """count++ < int(v.b * 4): true for the first n calls."""
        f, m = self.f, self.m
        i = f.load(self.i32, self.count)
        f.store(self.count, f.binop(sp.OpIAdd, self.i32, i, m.c_int(1)))
        return f.binop(sp.OpSLessThan, self.b, i, self._quarter(2))

    def finish(self):
        f, m = self.f, self.m
        c = f.load(self.i32, self.code)
        parts = []
        for sh in (0, 8, 16):
            x = f.binop(sp.OpShiftRightLogical, self.i32, c, m.c_int(sh)) if sh else c
            x = f.binop(sp.OpBitwiseAnd, self.i32, x, m.c_int(255))
            parts.append(f.fdiv(self.f32, f._result(sp.OpConvertSToF, self.f32, x), m.c_float(255.0)))
        f.store(self.out, f.construct(self.v4, *parts, m.c_float(1.0)))
        f.ret_only()


def path_shader(build, helpers=None):
    def body(m, var):
        v4 = m.t_vec(4)
        vin = var(sp.StorageInput, v4, "v", location=0)
        out = var(sp.StorageOutput, v4, "o", location=0)
        extra = helpers(m) if helpers else None
        f = sp.Function(m)
        b = Builder(m, f, vin, out)
        build(b, extra)
        f.end()
        return f

    return shaders._fragment_module(body).to_bytes()


# --- the reference's templates, restated ------------------------------------
# each: (build(b, extra), model(cond, sel, n) -> code or None (discarded))

def t_return(b, _):
    b.mark(1)
    b.finish()


def t_return_value_helpers(m):
    # int f(int x) { return x * 7 + 5; }  (one block ending in OpReturnValue)
    i32 = m.t_int()
    fn = sp.Function(m, "f", ret=i32, params=(i32,))
    fn.ret_value(fn.binop(sp.OpIAdd, i32, fn.binop(sp.OpIMul, i32, fn.params[0], m.c_int(7)), m.c_int(5)))
    return fn.id


def t_return_value(b, callee):
    c = b.f.load(b.i32, b.code)
    b.f.store(b.code, b.f.call(b.i32, callee, c))
    b.finish()


def t_simple_discard(b, _):
    l2 = b.L()
    b.mark(1)
    b.f.branch(l2)
    b.f.label(l2)
    b.f.kill()


def t_conditional_none_none(b, _):
    endif = b.L()
    b.mark(1)
    c = b.cond()
    b.f.selection_merge(endif)
    b.f.branch_cond(c, endif, endif)
    b.f.label(endif)
    b.mark(2)
    b.finish()


def t_conditional_merge_none(b, _):
    then, endif = b.L(), b.L()
    b.mark(1)
    c = b.cond()
    b.f.selection_merge(endif)
    b.f.branch_cond(c, then, endif)
    b.f.label(then)
    b.mark(2)
    b.f.branch(endif)
    b.f.label(endif)
    b.mark(3)
    b.finish()


def t_conditional_none_merge(b, _):
    els, endif = b.L(), b.L()
    b.mark(1)
    c = b.cond()
    b.f.selection_merge(endif)
    b.f.branch_cond(c, endif, els)
    b.f.label(els)
    b.mark(2)
    b.f.branch(endif)
    b.f.label(endif)
    b.mark(3)
    b.finish()


def t_conditional_return_merge(b, _):
    then, els, endif = b.L(), b.L(), b.L()
    b.mark(1)
    c = b.cond()
    b.f.selection_merge(endif)
    b.f.branch_cond(c, then, els)
    b.f.label(then)
    b.mark(2)
    b.finish()
    b.f.label(els)
    b.mark(3)
    b.f.branch(endif)
    b.f.label(endif)
    b.mark(4)
    b.finish()


def t_conditional_return_return(b, _):
    then, els, endif = b.L(), b.L(), b.L()
    b.mark(1)
    c = b.cond()
    b.f.selection_merge(endif)
    b.f.branch_cond(c, then, els)
    b.f.label(then)
    b.mark(2)
    b.finish()
    b.f.label(els)
    b.mark(3)
    b.finish()
    b.f.label(endif)
    b.mark(4)
    b.finish()


def _switch(b, default, cases, order, bodies, merge):
    """order: labels in block order; bodies[label] = (mark, next label or 'ret')."""
    b.mark(1)
    s = b.sel()
    b.f.selection_merge(merge)
    b.f.switch(s, default, cases)
    for lab in order:
        k, nxt = bodies[lab]
        b.f.label(lab)
        b.mark(k)
        if nxt == "ret":
            b.finish()
        else:
            b.f.branch(nxt)
    b.f.label(merge)
    b.mark(9)
    b.finish()


def t_switch_default_break(b, _):
    default, merge = b.L(), b.L()
    _switch(b, default, [], [default], {default: (2, merge)}, merge)


def t_switch_return_default_break(b, _):
    case1, default, merge = b.L(), b.L(), b.L()
    _switch(b, default, [(0, case1)], [case1, default], {case1: (2, "ret"), default: (3, merge)}, merge)


def t_switch_fallthrough_default_break(b, _):
    case1, default, merge = b.L(), b.L(), b.L()
    _switch(b, default, [(0, case1)], [case1, default], {case1: (2, default), default: (3, merge)}, merge)


def t_switch_fallthrough_default_fallthrough_break(b, _):
    case1, default, case2, merge = b.L(), b.L(), b.L(), b.L()
    _switch(b, default, [(0, case1), (1, case1), (2, case2)], [case1, default, case2],
            {case1: (2, default), default: (3, case2), case2: (4, merge)}, merge)


def t_switch_break_default_fallthrough_break(b, _):
    case1, default, case2, merge = b.L(), b.L(), b.L(), b.L()
    _switch(b, default, [(0, case1), (1, case1), (2, case2)], [case1, default, case2],
            {case1: (2, merge), default: (3, case2), case2: (4, merge)}, merge)


def t_switch_break_default_fallthrough_break2(b, _):
    case1, case2, default, merge = b.L(), b.L(), b.L(), b.L()
    _switch(b, default, [(0, case1), (1, case1), (2, case2)], [case1, case2, default],
            {case1: (2, merge), default: (3, case2), case2: (4, merge)}, merge)


def t_while(b, _):
    header, merge = b.L(), b.L()
    b.mark(1)
    b.f.branch(header)
    b.f.label(header)
    b.f.loop_merge(merge, header)
    b.mark(2)
    b.f.branch_cond(b.step(), header, merge)
    b.f.label(merge)
    b.mark(3)
    b.finish()


def t_while_body(b, _):
    header, body, merge = b.L(), b.L(), b.L()
    b.mark(1)
    b.f.branch(header)
    b.f.label(header)
    b.f.loop_merge(merge, header)
    b.f.branch_cond(b.step(), body, merge)
    b.f.label(body)
    b.mark(2)
    b.f.branch(header)
    b.f.label(merge)
    b.mark(3)
    b.finish()


def t_while_body_infinite(b, _):
    header, body, merge = b.L(), b.L(), b.L()
    b.mark(1)
    b.f.branch(header)
    b.f.label(header)
    b.f.loop_merge(merge, header)
    b.f.branch(body)
    b.f.label(body)
    b.mark(2)
    b.f.branch(header)
    b.f.label(merge)
    b.mark(3)
    b.finish()


# --- shapes beyond the templates ---------------------------------------------

def t_for_break_continue(b, _):
    """for (i = 0; step(); ++j) { if (cond && j == sel) break; if (j == 1) continue; mark(3); }
    with the loop's own continue target (mark 4) and ifs whose arms break
    and continue."""
    header, cond_l, body, brk_if, cont_if, tail, cont, merge = (b.L() for _ in range(8))
    f, m = b.f, b.m
    j = f.var(b.i32)
    f.store(j, m.c_int(0))
    b.mark(1)
    f.branch(header)
    f.label(header)
    f.loop_merge(merge, cont)
    f.branch(cond_l)
    f.label(cond_l)
    b.mark(2)
    f.branch_cond(b.step(), body, merge)
    f.label(body)
    jv = f.load(b.i32, j)
    c1 = f.binop(sp.OpLogicalAnd, b.b, b.cond(), f.binop(sp.OpIEqual, b.b, jv, b.sel()))
    f.selection_merge(brk_if)
    f.branch_cond(c1, merge, brk_if)        # break from inside a selection
    f.label(brk_if)
    f.selection_merge(cont_if)
    f.branch_cond(f.binop(sp.OpIEqual, b.b, jv, m.c_int(1)), cont, cont_if)  # continue from inside one
    f.label(cont_if)
    f.branch(tail)
    f.label(tail)
    b.mark(3)
    f.branch(cont)
    f.label(cont)
    b.mark(4)
    f.store(j, f.binop(sp.OpIAdd, b.i32, f.load(b.i32, j), m.c_int(1)))
    f.branch(header)
    f.label(merge)
    b.mark(5)
    b.finish()


def t_nested_loops(b, _):
    """outer: 2 trips; inner: while step() { mark(3) } (the shared counter
    runs out in the first outer trip)."""
    oh, ob, oc, om, ih, ib, im = (b.L() for _ in range(7))
    f, m = b.f, b.m
    k = f.var(b.i32)
    f.store(k, m.c_int(0))
    f.branch(oh)
    f.label(oh)
    f.loop_merge(om, oc)
    kv = f.load(b.i32, k)
    f.branch_cond(f.binop(sp.OpSLessThan, b.b, kv, m.c_int(2)), ob, om)
    f.label(ob)
    b.mark(2)
    f.branch(ih)
    f.label(ih)
    f.loop_merge(im, ih)
    f.branch_cond(b.step(), ib, im)
    f.label(ib)
    b.mark(3)
    f.branch(ih)
    f.label(im)
    f.branch(oc)
    f.label(oc)
    b.mark(4)
    f.store(k, f.binop(sp.OpIAdd, b.i32, f.load(b.i32, k), m.c_int(1)))
    f.branch(oh)
    f.label(om)
    b.mark(5)
    b.finish()


# --- the model ----------------------------------------------------------------

def model(name, cond, sel, n):
    code = [1]
    cnt = [0]

    def mark(k):
        code[0] = wrap32(code[0] * 7 + k)

    def step():
        i = cnt[0]
        cnt[0] += 1
        return i < n

    if name == "return":
        mark(1)
    elif name == "return_value":
        code[0] = wrap32(code[0] * 7 + 5)
    elif name == "simple_discard":
        return None
    elif name == "conditional_none_none":
        mark(1), mark(2)
    elif name == "conditional_merge_none":
        mark(1)
        if cond:
            mark(2)
        mark(3)
    elif name == "conditional_none_merge":
        mark(1)
        if not cond:
            mark(2)
        mark(3)
    elif name == "conditional_return_merge":
        mark(1)
        if cond:
            mark(2)
        else:
            mark(3), mark(4)
    elif name == "conditional_return_return":
        mark(1)
        mark(2 if cond else 3)
    elif name.startswith("switch"):
        mark(1)
        case = {"switch_default_break": {},
                "switch_return_default_break": {0: "case1"},
                "switch_fallthrough_default_break": {0: "case1"},
                "switch_fallthrough_default_fallthrough_break": {0: "case1", 1: "case1", 2: "case2"},
                "switch_break_default_fallthrough_break": {0: "case1", 1: "case1", 2: "case2"},
                "switch_break_default_fallthrough_break2": {0: "case1", 1: "case1", 2: "case2"}}[name].get(
                    sel, "default")
        flow = {"switch_default_break": {"default": [(2, "merge")]},
                "switch_return_default_break": {"case1": [(2, "ret")], "default": [(3, "merge")]},
                "switch_fallthrough_default_break": {"case1": [(2, "default")], "default": [(3, "merge")]},
                "switch_fallthrough_default_fallthrough_break": {
                    "case1": [(2, "default")], "default": [(3, "case2")], "case2": [(4, "merge")]},
                "switch_break_default_fallthrough_break": {
                    "case1": [(2, "merge")], "default": [(3, "case2")], "case2": [(4, "merge")]},
                "switch_break_default_fallthrough_break2": {
                    "case1": [(2, "merge")], "default": [(3, "case2")], "case2": [(4, "merge")]}}[name]
        while case not in ("merge", "ret"):
            k, case = flow[case][0]
            mark(k)
        if case == "merge":
            mark(9)
    elif name == "while":
        mark(1)
        while True:
            mark(2)
            if not step():
                break
        mark(3)
    elif name == "while_body":
        mark(1)
        while step():
            mark(2)
        mark(3)
    elif name == "for_break_continue":
        mark(1)
        j = 0
        while True:
            mark(2)
            if not step():
                break
            if cond and j == sel:
                break
            if j != 1:
                mark(3)
            mark(4)
            j += 1
        mark(5)
    elif name == "nested_loops":
        for _ in range(2):
            mark(2)
            while step():
                mark(3)
            mark(4)
        mark(5)
    else:
        raise KeyError(name)
    return code[0]


TEMPLATES = {
    "return": (t_return, None),
    "return_value": (t_return_value, t_return_value_helpers),
    "simple_discard": (t_simple_discard, None),
    "conditional_none_none": (t_conditional_none_none, None),
    "conditional_merge_none": (t_conditional_merge_none, None),
    "conditional_none_merge": (t_conditional_none_merge, None),
    "conditional_return_merge": (t_conditional_return_merge, None),
    "conditional_return_return": (t_conditional_return_return, None),
    "switch_default_break": (t_switch_default_break, None),
    "switch_return_default_break": (t_switch_return_default_break, None),
    "switch_fallthrough_default_break": (t_switch_fallthrough_default_break, None),
    "switch_fallthrough_default_fallthrough_break": (t_switch_fallthrough_default_fallthrough_break, None),
    "switch_break_default_fallthrough_break": (t_switch_break_default_fallthrough_break, None),
    "switch_break_default_fallthrough_break2": (t_switch_break_default_fallthrough_break2, None),
    "while": (t_while, None),
    "while_body": (t_while_body, None),
    "while_body_infinite": (t_while_body_infinite, None),
    "for_break_continue": (t_for_break_continue, None),
    "nested_loops": (t_nested_loops, None),
}
# never run on a GPU: its loop has no exit
CPU_ONLY = {"while_body_infinite"}


def module_of(name):
    build, helpers = TEMPLATES[name]
    return path_shader(build, helpers)


def _ir_text(driver, words):
    w = np.frombuffer(words, dtype=np.uint32).copy()
    return vk.tool_call(driver, "kazan_translate_to_ir_text", w.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_size_t(w.size), ctypes.c_uint32(4), b"main")


@pytest.mark.parametrize("name", sorted(TEMPLATES))
def test_template_translates(driver, monkeypatch, name):
    rc, text = _ir_text(driver, module_of(name))
    assert rc == 0, text
    if "while" in name or "loop" in name:
        assert "\n            loop " in text or " loop " in text
    if "switch" in name:
        assert "branch " in text and "default ->" in text
    if name == "simple_discard":
        assert "kill;" in text
    monkeypatch.setitem(rd._FS, g.FS_COLOR, lambda: module_of(name))
    rc, src = rd.compile_offline(driver, g.random_triangles(8, 16, 16, seed=1), 1)
    assert rc == 0, src


@pytest.mark.parametrize("name", ["switch_break_default_fallthrough_break2", "for_break_continue", "nested_loops"])
def test_template_compiles(driver, monkeypatch, name):
    monkeypatch.setitem(rd._FS, g.FS_COLOR, lambda: module_of(name))
    rc, txt = rd.compile_offline(driver, g.random_triangles(8, 16, 16, seed=1), 2)
    assert rc == 0, txt


def test_unstructured_flow_is_rejected(driver):
    # two selections sharing one merge block: the second reaches it twice
    def build(b, _):
        t1, m1 = b.L(), b.L()
        b.f.selection_merge(m1)
        b.f.branch_cond(b.cond(), t1, m1)
        b.f.label(t1)
        b.f.branch(m1)
        b.f.label(m1)
        b.f.branch(t1)  # back into a block already emitted, outside any loop
    rc, text = _ir_text(driver, path_shader(build))
    assert rc != 0 and "unstructured" in text


def test_loop_phong_ir(driver):
    rc, text = _ir_text(driver, shaders.fs_phong())
    assert rc == 0, text
    assert "loop " in text and "continue " in text and "ptr_index" in text and "_phi" in text


def test_model_matches_hand_traces():
    # a few hand-traced paths of the model itself
    assert model("conditional_return_merge", True, 0, 0) == wrap32(wrap32(1 * 7 + 1) * 7 + 2)
    c = 1
    for k in (1, 2, 2, 2, 3):  # while with n = 2: header runs n + 1 times
        c = wrap32(c * 7 + k)
    assert model("while", False, 0, 2) == c
    c = 1
    for k in (1, 4, 9):  # sel 2 -> case2 (mark 4) -> merge
        c = wrap32(c * 7 + k)
    assert model("switch_break_default_fallthrough_break2", False, 2, 0) == c
    c = 1
    for k in (1, 3, 4, 9):  # sel 3 -> default (mark 3) falls through to case2 (mark 4)
        c = wrap32(c * 7 + k)
    assert model("switch_break_default_fallthrough_break2", False, 3, 0) == c


# --- on the GPU -----------------------------------------------------------------

R_VALUES = (0.25, 0.75)
Q_VALUES = (0.1, 0.3, 0.6, 0.9)  # int(x * 4) = 0, 1, 2, 3


def flat_scene(seed):
    """Triangles of flat colours from the grid R x Q x Q (all 32 inputs)."""
    r = np.random.default_rng(seed)
    combos = [(a, b_, c) for a in R_VALUES for b_ in Q_VALUES for c in Q_VALUES]
    n = 240
    s = g.random_triangles(n, 96, 80, seed=seed)
    pick = r.integers(0, len(combos), n)
    cols = np.array([combos[k] for k in pick], np.float32)
    v = s.vertices.reshape(n, 3, 8)
    v[:, :, 4:7] = cols[:, None, :]
    v[:, :, 7] = 1.0
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(set(TEMPLATES) - CPU_ONLY))
def test_template_renders_its_paths(ctx, oracle, monkeypatch, name):
    print("control-flow template on the GPU: %s" % name, flush=True)
    s = flat_scene(7)
    oc, od, _, _ = oracle.render(s)  # which triangle (colour) is visible where
    monkeypatch.setitem(rd._FS, g.FS_COLOR, lambda: module_of(name))
    r = rd.SceneRenderer(ctx, s)
    r.render(1)
    c, d, _ = r.read_back()
    covered = od[:, :, 0] < 1.0
    if name == "simple_discard":
        assert (c == 0).all() and (d == np.float32(1.0)).all()
        return
    np.testing.assert_array_equal(d.view(np.uint32), od.view(np.uint32))
    enc = {}
    for a in R_VALUES:
        for b_ in Q_VALUES:
            for cc in Q_VALUES:
                key = tuple(int(np.float32(x) * np.float32(255) + np.float32(0.5)) for x in (a, b_, cc))
                code = model(name, np.float32(a) < np.float32(0.5), int(np.float32(b_) * np.float32(4)),
                             int(np.float32(cc) * np.float32(4)))
                enc[key] = [code & 255, (code >> 8) & 255, (code >> 16) & 255, 255]
    want = np.zeros_like(c[:, :, 0])
    ys, xs = np.nonzero(covered)
    for y, x in zip(ys, xs):
        want[y, x] = enc[tuple(int(t) for t in oc[y, x, 0, :3])]
    np.testing.assert_array_equal(c[:, :, 0][covered], want[covered])
    assert (c[:, :, 0][~covered] == 0).all()
