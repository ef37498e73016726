"""The shader front end against the reference's own golden outputs.

- kazan_spirv_disassemble vs the expected dumps of spirv-parser's tests
  (spirv-parser/src/lib.rs:45-216), on SPIR-V reassembled from those dumps
  (every instruction offset in the dump is checked against the reassembly).
- kazan_translate_to_ir_text vs trivial_test.kazan-ir
  (shader-compiler-translate-spirv-to-ir/src/lib.rs:232-258): an empty
  GLCompute "main" whose one block ends at "stdin":4:0.
"""
import ctypes
import os

import numpy as np
import pytest

from kazan_amd import vk
from spirv_text import assemble

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def _disassemble(driver, words):
    w = np.asarray(words, np.uint32)
    return vk.tool_call(driver, "kazan_spirv_disassemble", w.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(w.size))


@pytest.mark.parametrize("name", ["parse_test", "parse_test2", "parse_test3"])
def test_spirv_dump_matches_reference(driver, name):
    expected = _read("spirv_%s.dump" % name)
    words, offsets = assemble(expected)
    assert offsets and all(a == b for a, b in offsets), [o for o in offsets if o[0] != o[1]][:3]
    rc, text = _disassemble(driver, words)
    assert rc == 0, text
    for i, (a, b) in enumerate(zip(text.splitlines(), expected.splitlines())):
        assert a == b, "line %d: %r != %r" % (i + 1, a, b)
    assert text == expected


def test_spirv_dump_of_byte_swapped_module(driver):
    # the parser accepts either byte order (the magic tells which)
    words, _ = assemble(_read("spirv_parse_test2.dump"))
    swapped = np.asarray(words, np.uint32).byteswap()
    rc, text = _disassemble(driver, swapped)
    assert rc == 0 and text == _read("spirv_parse_test2.dump")


def test_spirv_parse_errors(driver):
    rc, text = _disassemble(driver, [0x12345678, 0x10000, 0, 1, 0])
    assert rc != 0
    words, _ = assemble(_read("spirv_parse_test2.dump"))
    rc, text = _disassemble(driver, words[:-1] + [(5 << 16) | 56])  # last instruction runs off the end
    assert rc != 0


TRIVIAL = """; SPIR-V
; Version: 1.0
; Generator: 0x80007
; Bound: 7
; Schema: 0
               OpCapability Shader
          %1 = OpExtInstImport "GLSL.std.450"
               OpMemoryModel Logical GLSL450
               OpEntryPoint GLCompute %4 "main"
               OpExecutionMode %4 LocalSize 1 1 1
          %5 = OpString "stdin"
               OpSource GLSL 450
               OpName %4 "main"
          %2 = OpTypeVoid
          %3 = OpTypeFunction %2
          %4 = OpFunction %2 None %3
          %6 = OpLabel
               OpLine %5 4 0
               OpReturn
               OpFunctionEnd
"""


def test_trivial_shader_ir_matches_reference(driver):
    words, _ = assemble(TRIVIAL)
    w = np.asarray(words, np.uint32)
    rc, text = vk.tool_call(driver, "kazan_translate_to_ir_text", w.ctypes.data_as(ctypes.c_void_p),
                            ctypes.c_size_t(w.size), ctypes.c_uint32(5), b"main")
    assert rc == 0, text
    expected = _read("trivial_test.kazan-ir")
    for i, (a, b) in enumerate(zip(text.splitlines(), expected.splitlines())):
        assert a == b, "line %d: %r != %r" % (i + 1, a, b)
    assert text.rstrip("\n") == expected.rstrip("\n")
