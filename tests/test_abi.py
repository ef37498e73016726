"""The C ABI of libkazan_driver.so, without a GPU.

- Every Vulkan struct the Python binding passes has the size the C compiler
  gives include/kazan/vk_abi.h (a layout mismatch would corrupt calls).
- The library exports exactly the reference cdylib's three ICD entry points
  (vulkan-driver/src/lib.rs:37-128) plus the kazan_* tools of
  include/kazan/kazan_tools.h; every vk* entry point, including the KAZAN
  extensions of include/kazan/kazan_icd.h, resolves by name through
  vk_icdGetInstanceProcAddr.
- Interface negotiation clamps to version 5 (lib.rs:51-61).
- The training-data and cost-model structs match their C layouts.
"""
import ctypes
import os
import re
import subprocess

import pytest

from kazan_amd import vk
from kazan_amd.balancer import model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


@pytest.fixture(scope="module")
def c_sizes(tmp_path_factory):
    names = [s.__name__ for s in vk.ABI_STRUCTS] + ["KazanPipelineDesc", "KazanClusterScene"]
    d = tmp_path_factory.mktemp("abi")
    src = d / "sizes.c"
    body = "".join('    printf("%s %%zu\\n", sizeof(%s));\n' % (n, n) for n in names)
    src.write_text('#include <stdio.h>\n#include "kazan/kazan_icd.h"\n#include "kazan/kazan_tools.h"\n'
                   "int main(void) {\n%s    return 0;\n}\n" % body)
    exe = d / "sizes"
    subprocess.run(["gcc", "-std=c11", "-I" + INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_struct_sizes_match_c(c_sizes):
    structs = list(vk.ABI_STRUCTS) + [vk.KazanPipelineDesc, vk.KazanClusterScene]
    bad = [(s.__name__, ctypes.sizeof(s), c_sizes[s.__name__]) for s in structs
           if ctypes.sizeof(s) != c_sizes[s.__name__]]
    assert not bad, bad


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def _header_functions(name, pattern):
    with open(os.path.join(INC, "kazan", name)) as f:
        return set(re.findall(pattern, f.read()))


def test_exported_symbols():
    exports = _exports(vk.DRIVER_PATH)
    icd = {"vk_icdGetInstanceProcAddr", "vk_icdNegotiateLoaderICDInterfaceVersion",
           "vk_icdGetPhysicalDeviceProcAddr"}
    tools = _header_functions("kazan_tools.h", r"\bint (kazan_\w+)\(")
    backend = _header_functions("kazan_backend.h", r"KAZAN_EXPORT [\w\s\*]*?\b(kazan_\w+)\(")
    assert icd <= exports and tools <= exports and backend <= exports, (icd | tools | backend) - exports
    assert len(tools) >= 5 and len(backend) == 8, backend
    # nothing else with a C name: every vk* entry point goes through the loader
    assert {e for e in exports if e.startswith("vk")} <= icd, sorted(e for e in exports if e.startswith("vk"))


def test_every_entry_point_resolves(driver):
    declared = _header_functions("kazan_icd.h", r"PFN_(vk\w+KAZAN)\)")
    assert declared and declared <= set(vk._SIGS), declared - set(vk._SIGS)
    for name in vk._SIGS:
        assert driver._gipa(None, name.encode()), name
    assert not driver._gipa(None, b"vkNoSuchFunction")


def test_negotiation_clamps_to_5():
    lib = ctypes.CDLL(vk.DRIVER_PATH)
    f = lib.vk_icdNegotiateLoaderICDInterfaceVersion
    f.restype = ctypes.c_int32
    for asked, got in ((7, 5), (5, 5), (3, 3)):
        v = ctypes.c_uint32(asked)
        assert f(ctypes.byref(v)) == 0 and v.value == got


def test_cost_model_blob_layout(c_sizes):
    assert M.PARAMS == 9473
    blob = M.pack(M.init_params(0), [0.0] * 16, [1.0] * 16, 0.5)
    assert len(blob) == M.BLOB_BYTES
    m = M.unpack(blob)
    assert m["frag_weight"] == 0.5 and m["w1"].shape == (64, 16)


def test_cost_model_eval_rejects_bad_blob(driver):
    fn = driver.lib.kazan_cost_model_eval
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    blob = bytearray(M.pack(M.init_params(0), [0.0] * 16, [1.0] * 16, 0.5))
    assert fn(bytes(blob[:-4]), len(blob) - 4, None, 0, None) < 0   # wrong size
    blob[4] = 9                                                     # wrong version
    assert fn(bytes(blob), len(blob), None, 0, None) < 0


def test_backend_name():
    """Compiler::name (shader-compiler-backend/src/lib.rs): the backend's
    name, without a device."""
    lib = ctypes.CDLL(vk.DRIVER_PATH)
    lib.kazan_backend_name.restype = ctypes.c_char_p
    assert lib.kazan_backend_name() == b"hip-gfx950"
