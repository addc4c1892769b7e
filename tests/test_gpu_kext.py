"""The extended interval table (gs_index_set_kmer_len): its entries are the
SA intervals of all patterns of 14..16 bases with the presence bits of their
three-base left extensions, and seeding with it gives the same output as
without (the 13-mer table plus backward steps)."""
import random
import sys

import numpy as np
import pytest

import brute
import oracle_bind as ob
from conftest import build_index, random_text, repetitive_text, sample_reads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g(built):
    import gpuseed_amd
    return gpuseed_amd


def steps(ox, p, n):
    """Interval after the first n codes of p (backward-search order), or None."""
    l, u = 0, ox.seq_len
    for j in range(n):
        l, u = ox.step(l, u, (p >> (2 * j)) & 3)
        if l > u:
            return None
    return l, u


@pytest.mark.parametrize("length", [14, 15, 16])
def test_entries_are_sa_intervals(g, tmp_path, length):
    rng = random.Random(length)
    text = repetitive_text(rng, 5000)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    bi = brute.BruteIndex(text, 7)
    with g.Index(prefix) as X:
        X.set_kmer_len(length)
        info = X.info()
        assert info.kmer_len == length
        ps = [rng.randrange(4 ** length) for _ in range(2000)]
        for i in range(0, len(text) - length, 5):   # patterns that occur
            ps.append(sum("ACGT".index(text[i + length - 1 - j]) << (2 * j) for j in range(length)))
        for i, p in enumerate(ps):
            l, u, mlo, mhi = (int(v) for v in X.read_kext(p, 1)[0])
            mask = mlo | (mhi << 32)
            want = steps(ox, p, length)
            if want is None:
                assert l > u and mask == 0, p
            else:
                assert (l, u) == want, p
                pat = "".join("ACGT"[(p >> (2 * j)) & 3] for j in reversed(range(length)))
                assert bi.interval(pat) == want
                if i % 4 == 0:  # presence bits: c3.c2.c1.pat occurs, by substring search
                    for bit in range(64):
                        c1, c2, c3 = bit & 3, (bit >> 2) & 3, bit >> 4
                        occurs = bi.interval("ACGT"[c3] + "ACGT"[c2] + "ACGT"[c1] + pat) is not None
                        assert bool(mask >> bit & 1) == occurs, (p, bit)
        X.set_kmer_len(13)
        assert X.info().kmer_len == 13


@pytest.mark.parametrize("length", [14, 15, 16])
def test_seeds_unchanged_by_the_extended_table(g, tmp_path, length):
    rng = random.Random(100 + length)
    text = repetitive_text(rng, 20000, n_units=8, unit_len=150, copies=6, div=0.02)
    reads = sample_reads(rng, text, 300, 60, 150, sub=0.02, indel=0.005, n_frac=0.3)
    reads += sample_reads(rng, text, 20, 161, 1200, sub=0.04, indel=0.02, n_frac=0.3)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    with g.Index(prefix) as X:
        X.set_kmer_len(length)
        for smem, k in ((True, 19), (False, 20), (True, 14), (True, 17)):
            want, _ = ob.seed(ox, reads, k, smem=smem, prune=1, threads=8)
            want = [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in want]
            for kernel in (0, 1):
                say("T=%d smem=%d k=%d kernel=%d" % (length, smem, k, kernel))
                got = g.seed(X, reads, k=k, smem=smem, kernel=kernel).as_lists()
                bad = [i for i in range(len(reads)) if got[i] != want[i]]
                assert not bad, "read %d" % bad[0]


@pytest.mark.parametrize("length", [13, 16])
@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20)])
@pytest.mark.parametrize("kernel,sat", [(1, 255), (2, 65535), (3, 15)])
def test_kernel_work_equals_the_oracle_work_model(g, tmp_path, length, smem, k, kernel, sat):
    """The batched kernel's own memory-operation counters equal the oracle's
    count of the same algorithm (distance-1 pruning with size records that
    saturate like the kernel's 8- or 16-bit ones, the SMEM read-start stop,
    the extended table, tasks whose clean window is absent ending at once):
    8-byte table lookups and 16-byte occ entries, for
    reads held whole in the kernel's LDS window."""
    rng = random.Random(7 + length)
    text = repetitive_text(rng, 30000, n_units=8, unit_len=200, copies=6, div=0.02, tandem=6)
    reads = sample_reads(rng, text, 600, 60, 160, sub=0.01, indel=0.002, n_frac=0.1)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    arrays = g.pack_reads(reads)
    _, _, _, _, c = ob.seed_arrays(ox, *arrays, k, smem, threads=8, smem_stop=1,
                                   ext_len=length if length > 13 else 0, prune_sat=sat, miss_skip=1, tandem=6)
    with g.Index(prefix) as X:
        if length > 13:
            X.set_kmer_len(length)
        b = g.Batch(X, arrays, k=k, smem=smem, kernel=kernel, split_tasks=g.NO_SPLIT)
        st = b.run()
        b.close()
    say("T=%d smem=%d: kernel %d table %d occ, oracle %d table %d occ" %
        (length, smem, st.n_table_reads, st.n_occ_reads, c["w_tasks"] + c["w_ext"], c["w_blocks96"]))
    assert st.n_table_reads == c["w_tasks"] + c["w_ext"]
    assert st.n_occ_reads == c["w_blocks96"]
    assert st.n_tandem_reads == c["w_tandem"]
