"""Indexes in upstream bwa 0.7's layout on the device: gs_index_open and the
seed_gen command line read them and seed exactly as from the fork-layout
files of the same text (the reference followed by its reverse complement,
as `bwa index` builds it); those results equal the oracle's."""
import os
import random
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob
from conftest import build_index, repetitive_text, revcomp, sample_reads
from test_bwa07 import fork_to_bwa07

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEED_GEN = os.path.join(ROOT, "gpuseed_amd", "bin", "seed_gen")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def both(built, tmp_path_factory):
    d = tmp_path_factory.mktemp("bwa07")
    rng = random.Random(70)
    text = repetitive_text(rng, 20000, n_units=6, unit_len=250, copies=5, div=0.02, tandem=3)
    fork = build_index(d, text + revcomp(text), 7, "fork")
    b07 = str(d / "b07")
    fork_to_bwa07(fork, b07)
    reads = sample_reads(rng, text, 500, 60, 200, sub=0.02, indel=0.005, n_frac=0.1)
    return dict(fork=fork, b07=b07, reads=reads, dir=str(d))


def test_open_reads_the_same_index(both):
    import gpuseed_amd as g
    with g.Index(both["fork"]) as A, g.Index(both["b07"]) as B:
        ia, ib = A.info(), B.info()
        assert (ia.seq_len, ia.primary, ia.sa_intv, list(ia.L2)) == (ib.seq_len, ib.primary, ib.sa_intv, list(ib.L2))
        n = int(ia.n_blocks) * 4
        assert np.array_equal(A.read_occ(0, n), B.read_occ(0, n))


@pytest.mark.parametrize("smem,k,kernel,tables", [(True, 19, 1, False), (False, 20, 1, False), (True, 19, 0, False),
                                                  (True, 19, 3, True), (False, 20, 3, True)])
def test_seeds_equal_fork_files_and_oracle(both, smem, k, kernel, tables):
    """tables: the upstream-layout index also gets the 16-mer table with
    presence bits and the direct-extension structures, derived from it."""
    import gpuseed_amd as g
    say("bwa 0.7 index smem=%d k=%d kernel=%d tables=%d" % (smem, k, kernel, tables))
    with g.Index(both["fork"]) as A, g.Index(both["b07"]) as B:
        if tables:
            B.set_kmer_len(16)
            B.set_direct(True)
        a = g.seed(A, both["reads"], k=k, smem=smem, kernel=kernel).as_lists()
        b = g.seed(B, both["reads"], k=k, smem=smem, kernel=kernel).as_lists()
    assert a == b
    ox = ob.Index(both["fork"])
    want, _ = ob.seed(ox, both["reads"], k, smem=smem, prune=1, threads=8)
    ox.close()
    assert a == [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in want]


def test_cli_reads_bwa07_files(both):
    fa = os.path.join(both["dir"], "reads.fa")
    with open(fa, "wb") as f:
        for i, r in enumerate(both["reads"]):
            f.write(b">r%d\n%s\n" % (i, r))
    outs = []
    for prefix in (both["fork"], both["b07"]):
        p = subprocess.run([SEED_GEN, "-o", "-s", "-k", "19", prefix, fa], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, timeout=300)
        assert p.returncode == 0, p.stderr.decode()[-2000:]
        outs.append(p.stdout)
    assert len(outs[0]) > 1000 and outs[0] == outs[1]
