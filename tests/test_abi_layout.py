"""The Python bindings' ctypes structures against include/gpuseed.h: every
structure's size and every field's offset, as a C compiler lays them out
from the header (gcc, no GPU needed)."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import gpuseed_amd as g

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(HERE, "..", "include")

PAIRS = [  # (C typedef, ctypes class)
    ("gs_roofline_result", g.RooflineResult),
    ("gs_synth_ref_params", g.SynthRefParams),
    ("gs_read_class", g.ReadClass),
    ("gs_synth_read_params", g.SynthReadParams),
    ("gs_index_info", g.IndexInfo),
    ("gs_opts", g.Opts),
    ("gs_seed_rec", g.SeedRec),
    ("gs_result", g.Result),
    ("gs_result_buffers", g.ResultBuffers),
    ("gs_batch_stats", g.BatchStats),
    ("gs_tune_result", g.TuneResult),
    ("gs_server_opts", g.ServerOpts),
    ("gs_request_info", g.RequestInfo),
    ("gs_server_stats", g.ServerStats),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_ctypes_structures_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gpuseed.h"', "int main(void) {"]
    for cname, cls in PAIRS:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, "-o", exe, str(src)], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    got = {}
    for line in subprocess.run([exe], check=True, stdout=subprocess.PIPE).stdout.decode().splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, cls in PAIRS:
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
