"""include/gpe_hip.h is a C header: it compiles as strict C99, and a C
program that takes the address of every function it declares links against
libgpehip.so (so each declaration has an exported definition) and runs
without a GPU (argument checks only)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "gpe_hip.h")
LIB = os.path.join(ROOT, "gpe_amd", "libgpehip.so")


def _declared():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpe_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(shutil.which("gcc") is None or not os.path.exists(LIB), reason="needs gcc and the built library")
def test_header_is_c99_and_every_declaration_links(tmp_path):
    names = _declared()
    assert len(names) > 40
    src = ['#include "gpe_hip.h"', "#include <stdio.h>", "typedef void (*fp)(void);",
           "static const fp table[] = {"]
    src += ["  (fp)%s," % n for n in names]
    src += ["};", "int main(void) {",
            "  double v = 0.0;",
            "  if (gpe_sim_norm(NULL, &v) != GPE_EINVAL) return 2;",
            '  printf("%d\\n", (int)(sizeof table / sizeof table[0]));',
            "  return 0;", "}"]
    c = tmp_path / "use_header.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "use_header"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I" + os.path.dirname(HDR), str(c), "-L" + os.path.dirname(LIB), "-lgpehip",
                        "-Wl,-rpath," + os.path.dirname(LIB), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(r.stdout.strip()) == len(names)
