"""The ctypes mirrors in motioncor3_amd/lib.py have the layout of the C
structs in include/mc_abi.h: a small C program built with gcc prints
sizeof and every field's offsetof, and they must equal ctypes' figures."""
import ctypes
import os
import shutil
import subprocess

import pytest

from motioncor3_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = [("mc_params", L.McParams), ("mc_movie", L.McMovie), ("mc_result", L.McResult),
           ("mc_override", L.McOverride), ("mc_raw_stack", L.McRawStack),
           ("mc_eer_stack", L.McEerStack)]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mc_abi.h"', 'int main(void) {']
    for cname, py in STRUCTS:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                   text=True).stdout.splitlines())
    for cname, py in STRUCTS:
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert int(got["%s.%s" % (cname, f[0])]) == getattr(py, f[0]).offset, (cname, f[0])
