"""Committed golden fixtures (tests/golden, written by make_golden.py from the
brute-force definitions): index bytes and -o text for a 3,000-base reference
and 48 read lines with the input edge cases seed_gen.cu:2955-2990 meets."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
PREFIX = os.path.join(GOLD, "tiny7")
READS = os.path.join(GOLD, "tiny_reads.fa")
ORACLE_SEED = os.path.join(ROOT, "oracle", "_build", "oracle_seed")
SEED_GEN = os.path.join(ROOT, "gpuseed_amd", "bin", "seed_gen")
MODES = [("smem", 19), ("smem", 14), ("mem", 20), ("mem", 16)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ref_codes():
    text = open(os.path.join(GOLD, "tiny_ref.txt")).read().strip()
    return np.frombuffer(text.encode().translate(bytes.maketrans(b"ACGT", b"\0\1\2\3")), dtype=np.uint8)


def golden_text(mode, k):
    return open(os.path.join(GOLD, "tiny_%s_k%d.txt" % (mode, k)), "rb").read()


def flags(mode, k):
    return (["-s"] if mode == "smem" else []) + ["-k", str(k)]


def test_writer_reproduces_golden_index(built, tmp_path):
    import gpuseed_amd as g
    prefix = os.path.join(str(tmp_path), "w")
    g.build_index_host(ref_codes(), 7, prefix)
    for ext in (".bwt", ".sa"):
        assert open(prefix + ext, "rb").read() == open(PREFIX + ext, "rb").read(), ext


@pytest.mark.parametrize("mode,k", MODES)
def test_oracle_reproduces_golden_text(built, mode, k):
    p = subprocess.run([ORACLE_SEED, "-o", "-t", "4"] + flags(mode, k) + [PREFIX, READS],
                       capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout == golden_text(mode, k)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("mode,k", MODES)
def test_seed_gen_reproduces_golden_text(built, mode, k, kernel):
    say("golden %s k=%d kernel=%d" % (mode, k, kernel))
    p = subprocess.run([SEED_GEN, "-o", "--kernel", str(kernel)] + flags(mode, k) + [PREFIX, READS],
                       capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    assert p.stdout == golden_text(mode, k)


@pytest.mark.gpu
def test_gpu_builder_reproduces_golden_index(built, tmp_path):
    import gpuseed_amd as g
    say("golden GPU build")
    prefix = os.path.join(str(tmp_path), "gb")
    X = g.Index(codes=ref_codes(), sa_intv=7, write_prefix=prefix)
    X.close()
    for ext in (".bwt", ".sa"):
        assert open(prefix + ext, "rb").read() == open(PREFIX + ext, "rb").read(), ext

