"""The host C code of libgpuseed (csrc/host/*.c) built with AddressSanitizer
and UBSan (gcc, no GPU code) and driven through the synthetic generators,
the index writer and readers, the bwa 0.7 index converter and reference
reader, and the text formatter (host_sanitize.c)."""
import glob
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from conftest import random_text
from test_bwa07 import write_bwa07_index, write_bwa_ref

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    srcs = sorted(glob.glob(os.path.join(ROOT, "gpuseed_amd", "csrc", "host", "*.c")))
    cmd = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-std=gnu11", "-pthread",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gpuseed_amd", "csrc"),
           os.path.join(HERE, "host_sanitize.c")] + srcs + ["-o", exe]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    rng = np.random.default_rng(5)
    p07 = str(tmp_path / "b07")
    write_bwa07_index(p07, list(rng.integers(0, 4, 3001, dtype=np.uint8)), 7)
    r = random.Random(3)
    pref = str(tmp_path / "ref")
    write_bwa_ref(pref, [("chr1", "first", random_text(r, 400) + "N" * 7 + random_text(r, 93)),
                         ("chrM", "", random_text(r, 61) + "R" + random_text(r, 5)),
                         ("scaffold_3", "x y", "NN" + random_text(r, 200))])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe, str(tmp_path), p07, pref], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       timeout=600, env=env)
    assert p.returncode == 0 and b"host code clean" in p.stdout, p.stderr.decode()[-3000:]
