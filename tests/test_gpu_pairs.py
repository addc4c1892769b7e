"""The pair table (gs_index_set_pairs): one byte per 18-base core saying
which bases follow and which precede it in the text.  Its bytes must equal
a direct count over the text; with it the batched kernels' seeds must not
change, and the kernels' table reads must equal the oracle's work model of
the pair probes (pair_probe)."""
import random
import sys

import numpy as np
import pytest

import oracle_bind as ob
from conftest import build_index, repetitive_text, sample_reads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g(built):
    import gpuseed_amd
    return gpuseed_amd


CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def core_index(t):
    """Index of an 18-base core: code of its last base in the low bits."""
    v = 0
    for j, ch in enumerate(reversed(t)):
        v |= CODE[ch] << (2 * j)
    return v


def expected_bytes(text):
    want = {}
    n = len(text)
    for p in range(n - 17):
        i = core_index(text[p:p + 18])
        b = want.get(i, 0)
        if p + 18 < n:
            b |= 1 << CODE[text[p + 18]]
        if p > 0:
            b |= 16 << CODE[text[p - 1]]
        want[i] = b
    return want


@pytest.mark.parametrize("direct", [False, True])
def test_pair_bytes_equal_the_text(g, tmp_path, direct):
    """Every core of the text has exactly the bits of its neighbours; other
    bytes are zero.  Built from the direct structures' text or from a text
    derived by walking the index."""
    rng = random.Random(11)
    text = repetitive_text(rng, 20000, n_units=6, unit_len=150, copies=6, div=0.02, tandem=4)
    prefix = build_index(tmp_path, text, 7)
    want = expected_bytes(text)
    with g.Index(prefix) as X:
        if direct:
            X.set_direct(True)
        X.set_pairs(True)
        assert X.info().pairs == 1
        for i, b in want.items():
            assert X.read_pairs(i, 1)[0] == b, i
        # cores that do not occur: a random sample of bytes, and the bytes
        # around every occurring one
        other = np.random.default_rng(1).integers(0, 4 ** 18 - 8, 2000)
        for i in list(other) + [i + 1 for i in want]:
            i = int(i)
            if i not in want and i < 4 ** 18:
                assert X.read_pairs(i, 1)[0] == 0, i
        X.set_pairs(False)
        assert X.info().pairs == 0


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20), (True, 16)])
def test_seeds_unchanged_and_work_equals_the_model(g, tmp_path, smem, k):
    """Seeds with the pair table equal the oracle's; the kernels' table reads
    (pair bytes included), occ entries and direct reads equal the oracle's
    model with pair probes (reads held whole in the LDS window, no work
    splitting).  With k < 19 the table must not be used."""
    rng = random.Random(29 + k)
    text = repetitive_text(rng, 40000, n_units=8, unit_len=200, copies=6, div=0.02, tandem=6)
    reads = sample_reads(rng, text, 700, 60, 160, sub=0.01, indel=0.002, n_frac=0.1)
    prefix = build_index(tmp_path, text, 7)
    arrays = g.pack_reads(reads)
    ox = ob.Index(prefix)
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *arrays, k, smem, threads=8)
    with g.Index(prefix) as X:
        X.set_kmer_len(16)
        X.set_direct(True)
        X.set_pairs(True)
        for kernel, sat in ((1, 255), (2, 65535), (3, 15)):
            say("pairs smem=%d k=%d kernel=%d" % (smem, k, kernel))
            out = g.seed(X, arrays=arrays, k=k, smem=smem, kernel=kernel, direct_min=3)
            assert np.array_equal(out.read_seed_off, so) and np.array_equal(out.read_pos_off, po)
            assert np.array_equal(out.seeds.view(np.uint32).reshape(-1, 6), seeds)
            assert np.array_equal(out.pos, pos)
            b = g.Batch(X, arrays, k=k, smem=smem, kernel=kernel, split_tasks=g.NO_SPLIT, direct_min=3)
            st = b.run()
            b.close()
            c = ob.seed_arrays(ox, *arrays, k, smem, threads=8, smem_stop=1, ext_len=16, prune_sat=sat, miss_skip=1,
                               direct_min=3, pair_probe=2, tandem=6)[4]
            say("  kernel: table %d occ %d direct %d; model: %d + %d + %d pair, %d, %d" %
                (st.n_table_reads, st.n_occ_reads, st.n_direct_reads, c["w_tasks"], c["w_ext"], c["w_pair"],
                 c["w_blocks96"], c["w_direct"]))
            assert (c["w_pair"] > 0) == (k >= 19)
            assert st.n_table_reads == c["w_tasks"] + c["w_ext"] + c["w_pair"]
            assert st.n_occ_reads == c["w_blocks96"]
            assert st.n_direct_reads == c["w_direct"]
            assert st.n_tandem_reads == c["w_tandem"]
    ox.close()


def test_long_reads_and_splitting_with_pairs(g, tmp_path):
    """Reads past the LDS window (segments) and forced work splitting with the
    pair table: the same seeds as the plain kernel and the oracle."""
    rng = random.Random(71)
    text = repetitive_text(rng, 30000, n_units=8, unit_len=300, copies=6, div=0.02, tandem=5)
    reads = sample_reads(rng, text, 300, 100, 160, sub=0.01, indel=0.002, n_frac=0.05)
    reads += sample_reads(rng, text, 40, 161, 2500, sub=0.03, indel=0.01, n_frac=0.2)
    prefix = build_index(tmp_path, text, 7)
    arrays = g.pack_reads(reads)
    ox = ob.Index(prefix)
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *arrays, 19, True, max_occ=100, threads=8)
    ox.close()
    with g.Index(prefix) as X:
        X.set_kmer_len(16)
        X.set_pairs(True)  # from a derived text: no direct structures
        for kw in (dict(kernel=3), dict(kernel=3, seg_tasks=7), dict(kernel=1, split_tasks=1, split_work=1),
                   dict(kernel=3, split_tasks=2, split_work=1)):
            out = g.seed(X, arrays=arrays, k=19, smem=True, max_occ=100, **kw)
            assert np.array_equal(out.read_seed_off, so) and np.array_equal(out.read_pos_off, po), kw
            assert np.array_equal(out.seeds.view(np.uint32).reshape(-1, 6), seeds), kw
            assert np.array_equal(out.pos, pos), kw
