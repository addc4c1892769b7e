"""The device's code packing (gs::pack_chunk in gs_device.h) compiled for the
host and compared with the byte-by-byte definition of the 4-bit codes for
every read length up to 300, every chunk, both strands and all four
alignments, the read flush against inaccessible pages on either side."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "gpuseed_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_pack_chunk_equals_the_byte_loop(tmp_path):
    exe = str(tmp_path / "pack_chunk_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", CSRC, "-o", exe, os.path.join(HERE, "pack_chunk_check.cpp")],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=300)
    out = p.stdout.decode()
    assert p.returncode == 0, out + p.stderr.decode()
    assert " 0 differ" in out and "1190400 chunks compared" in out
