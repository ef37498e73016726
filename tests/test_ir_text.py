"""The shader IR's text form parsed back (kazan_ir_round_trip; the
reference's IR text parser, shader-compiler-ir/src/text.rs).

- The reference's two IR goldens parse and print back byte for byte:
  test_assert_ir_matches_file.kazan-ir (shader-compiler-ir/test_data, the
  module the IR crate's own assert_ir_matches_file! test builds) and
  trivial_test.kazan-ir (the translator's golden).
- The reference's instruction text tests (instructions_impl.rs:608-622):
  "[arg0: i32, arg1: i32]" + "add [arg0, arg1] -> [result: i32]" and
  "[arg0: data_ptr]" + "load [arg0], align: 0x4 -> [result: i32]" print
  back as the instruction text.
- Every scene shader and every control-flow template translates to text
  that parses back to the same text.
- Malformed text is refused with the line it fails on.
- With KAZAN_IR_ROUND_TRIP=1 the pipeline compiler builds from the parsed
  module: every scene pipeline still compiles (the GPU tests render with it).
- SPIR-V ids are bounded by the module header's `bound` (the reference's
  IdMap, spirv-id-map/src/lib.rs:94-210): a module whose bound is one too
  small is refused as out of bounds.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from kazan_amd import vk
from kazan_amd.scenes import shaders

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _round_trip(driver, text, what=0):
    return vk.tool_call(driver, "kazan_ir_round_trip", text.encode(), ctypes.c_uint32(what))


def _ir_text(driver, module_bytes, stage):
    w = np.frombuffer(module_bytes, dtype=np.uint32).copy()
    return vk.tool_call(driver, "kazan_translate_to_ir_text", w.ctypes.data_as(ctypes.c_void_p),
                        ctypes.c_size_t(w.size), ctypes.c_uint32(stage), b"main")


@pytest.mark.parametrize("name", ["test_assert_ir_matches_file.kazan-ir", "trivial_test.kazan-ir"])
def test_reference_goldens_round_trip(driver, name):
    with open(os.path.join(GOLDEN, name)) as f:
        text = f.read()
    rc, out = _round_trip(driver, text)
    assert rc == 0, out
    assert out == text


@pytest.mark.parametrize("defs,instr", [
    ("[arg0: i32, arg1: i32]\n", "add [arg0, arg1] -> [result: i32]"),
    ("[arg0: data_ptr]\n", "load [arg0], align: 0x4 -> [result: i32]"),
])
def test_reference_instruction_text(driver, defs, instr):
    rc, out = _round_trip(driver, defs + instr, what=1)
    assert rc == 0, out
    assert out == instr


def _scene_shaders():
    return {
        "vs_pos4_color4": (shaders.vs_pos4_color4(), 0),
        "vs_pos3_uv2": (shaders.vs_pos3_uv2(), 0),
        "vs_mvp_pos3_color4": (shaders.vs_mvp_pos3_color4(), 0),
        "vs_mvp_pos3_normal3": (shaders.vs_mvp_pos3_normal3(), 0),
        "fs_color": (shaders.fs_color(), 4),
        "fs_texture": (shaders.fs_texture(), 4),
        "fs_phong": (shaders.fs_phong(), 4),
        "fs_alpha_test": (shaders.fs_alpha_test(), 4),
    }


@pytest.mark.parametrize("name", sorted(_scene_shaders()))
def test_scene_shader_ir_round_trips(driver, name):
    module, stage = _scene_shaders()[name]
    rc, text = _ir_text(driver, module, stage)
    assert rc == 0, text
    rc, out = _round_trip(driver, text)
    assert rc == 0, out
    assert out == text


def test_control_flow_templates_round_trip(driver):
    from test_control_flow import TEMPLATES, module_of
    for name in sorted(TEMPLATES):
        rc, text = _ir_text(driver, module_of(name), 4)
        assert rc == 0, (name, text)
        rc, out = _round_trip(driver, text)
        assert rc == 0, (name, out)
        assert out == text, name


@pytest.mark.parametrize("text,where", [
    ("module {\n    target_properties {\n", "line 3"),
    ("[a: i32]\nadd [a, b] -> [r: i32]", "undefined value b"),
    ("[a: i32]\nadd [a, a] -> [a: i32]", "defined twice"),
    ("[a: i32]\nfrobnicate [a] -> [r: i32]", "unknown instruction frobnicate"),
    ("[a: i32]\nbreak nowhere[]", "does not enclose"),
    ("[a: <5 x f32>]\nadd [a] -> [r: i32]", "5 lanes"),
    ("[a: i32]\nload [a], offset: 0x4 -> [r: i32]", "expected 'align'"),
])
def test_malformed_text_is_refused(driver, text, where):
    rc, out = _round_trip(driver, text, what=0 if text.startswith("module") else 1)
    assert rc != 0 and where in out, out


def test_golden_edit_changes_the_print(driver):
    """The parser keeps what it reads: a changed member offset or alignment
    in the golden prints back changed (nothing is re-derived or dropped)."""
    with open(os.path.join(GOLDEN, "test_assert_ir_matches_file.kazan-ir")) as f:
        text = f.read()
    edited = text.replace("0x4: i32: {};", "0x8: <2 x f32>: {};").replace("align: 0x4 -> global_var1",
                                                                          "align: 0x10 -> global_var1")
    rc, out = _round_trip(driver, edited)
    assert rc == 0, out
    assert out == edited and out != text


def test_pipelines_compile_from_parsed_ir(tmp_path):
    """KAZAN_IR_ROUND_TRIP=1: pipeline creation hands the backend the module
    parsed back from its text; every scene pipeline still compiles (run in a
    child process so the setting reaches the library's first reads)."""
    code = r'''
import sys
sys.path.insert(0, %r)
sys.path.insert(0, %r)
from kazan_amd import vk
from kazan_amd.scenes import generators as g
from kazan_amd.scenes.renderer import compile_offline
d = vk.Driver()
for s in (g.jittered_grid(4, 4, 32, 32, 16), g.zipf_draws(4, 64, 32, 32), g.stress_4k(4, 64, 32, 32, 4),
          g.gaussian_cluster(64, 32, 32)):
    rc, txt = compile_offline(d, s, 2)
    assert rc == 0, txt
print("ok")
''' % (ROOT, os.path.join(ROOT, "tests"))
    env = dict(os.environ, KAZAN_IR_ROUND_TRIP="1", KAZAN_PIPELINE_CACHE=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-3000:]


# --- ids bounded by the module header (spirv-id-map/src/lib.rs:94-210) -----

@pytest.mark.parametrize("name", ["fs_phong", "vs_mvp_pos3_normal3", "fs_texture"])
def test_ids_at_or_past_the_bound_are_refused(driver, name):
    module, stage = _scene_shaders()[name]
    w = np.frombuffer(module, dtype=np.uint32).copy()
    bound = int(w[3])
    rc, text = _ir_text(driver, w.tobytes(), stage)
    assert rc == 0, text
    # every id of the module is below its bound; lowering the bound by one
    # puts the largest id outside it
    assert bound > 1
    low = w.copy()
    low[3] = bound - 1
    rc, text = _ir_text(driver, low.tobytes(), stage)
    assert rc != 0 and "out of bounds" in text and "header bound %d" % (bound - 1) in text, text
    # a generous bound is fine
    high = w.copy()
    high[3] = bound + 1000
    rc, text2 = _ir_text(driver, high.tobytes(), stage)
    assert rc == 0, text2
