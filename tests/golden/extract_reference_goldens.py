"""Writes the shader front end's golden outputs from the reference's own tests.

  python tests/golden/extract_reference_goldens.py [REFERENCE_ROOT]

Sources (expected outputs only; the reference's .spv inputs are not in its
tree, so tests reassemble them from the dumps' own offsets and opcodes):
  spirv-parser/src/lib.rs parse_test, parse_test2, parse_test3 (:45-216):
      the expected text of the parser's dump -> spirv_parse_test{,2,3}.dump
  shader-compiler-translate-spirv-to-ir/test_data/trivial_test.kazan-ir:
      the IR text of a translated empty GLCompute shader -> trivial_test.kazan-ir
  shader-compiler-ir/test_data/test_assert_ir_matches_file.kazan-ir: the IR
      text of the module the IR crate's own text test builds
      -> test_assert_ir_matches_file.kazan-ir
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    src = open(os.path.join(ref, "spirv-parser/src/lib.rs")).read()
    for name in ("parse_test", "parse_test2", "parse_test3"):
        m = re.search(r"fn %s\(\) \{.*?let expected = r#\"(.*?)\"#;" % name, src, re.S)
        if not m:
            raise SystemExit("no expected dump for %s" % name)
        out = os.path.join(HERE, "spirv_%s.dump" % name)
        with open(out, "w") as f:
            f.write(m.group(1))
        print("wrote", out)
    for rel in ("shader-compiler-translate-spirv-to-ir/test_data/trivial_test.kazan-ir",
                "shader-compiler-ir/test_data/test_assert_ir_matches_file.kazan-ir"):
        out = os.path.join(HERE, os.path.basename(rel))
        with open(os.path.join(ref, rel)) as f, open(out, "w") as g:
            g.write(f.read())
        print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
