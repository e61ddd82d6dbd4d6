"""The ctypes mirror of every C-ABI struct has the header's layout.

A small C program compiled against include/pdds_engine.h prints each struct's sizeof and every
field's offsetof; each must equal the ctypes Structure's, so a field added to one side only, or
in another order, fails here without a GPU.
"""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from pdds_amd import _capi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "pdds_problem": A.Problem, "pdds_run": A.Run, "pdds_counters": A.Counters,
    "pdds_adaptive": A.Adaptive, "pdds_adaptive_report": A.AdaptiveReport,
    "pdds_tune_config": A.TuneConfig, "pdds_tune_report": A.TuneReport,
    "pdds_server_config": A.ServerConfig, "pdds_request": A.Request, "pdds_result": A.Result,
}


def test_struct_layouts_match_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pdds_engine.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.splitlines() if line)
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(cls, f).offset, (cname, f)
