"""Direct extension (gs_index_set_direct): the full suffix array, its inverse
and the 2-bit text that the device derives from the index equal the
brute-force ones, and the batched kernels' direct runs leave the seeds
unchanged, with the kernel's own memory-operation counts equal to the
oracle's work model of the same procedure."""
import random
import sys

import numpy as np
import pytest

import brute
import oracle_bind as ob
from conftest import build_index, random_text, repetitive_text, revcomp, sample_reads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g(built):
    import gpuseed_amd
    return gpuseed_amd


@pytest.mark.parametrize("n,sa_intv", [(1, 3), (2, 4), (65, 7), (300, 1), (1000, 4), (5000, 7)])
def test_structures_equal_brute_force(g, tmp_path, n, sa_intv):
    rng = random.Random(n)
    text = repetitive_text(rng, n) if n >= 1000 else random_text(rng, n)
    bi = brute.BruteIndex(text, sa_intv)
    say("n=%d sa_intv=%d" % (n, sa_intv))
    with g.Index(build_index(tmp_path, text, sa_intv)) as X:
        X.set_direct(True)
        assert X.info().direct == 1
        sa, isa, t = X.read_direct(0, n + 1)
        assert sa.tolist() == bi.sa
        assert isa.tolist() == bi.rank
        assert "".join("ACGT"[c] for c in t[:n]) == text
        X.set_direct(False)
        assert X.info().direct == 0
        X.set_direct(True)  # rebuilt after a drop
        assert X.read_direct(0, n + 1)[0].tolist() == bi.sa


def test_structures_of_a_gpu_built_index(g):
    rng = random.Random(5)
    text = repetitive_text(rng, 3000)
    codes = np.frombuffer(text.encode().translate(bytes.maketrans(b"ACGT", b"\0\1\2\3")), dtype=np.uint8)
    bi = brute.BruteIndex(text, 7)
    with g.Index(codes=codes, sa_intv=7) as X:
        X.set_direct(True)
        sa, isa, t = X.read_direct(0, len(text) + 1)
        assert sa.tolist() == bi.sa and isa.tolist() == bi.rank
        assert "".join("ACGT"[c] for c in t[:len(text)]) == text


@pytest.mark.parametrize("text,sa_intv", [("AC", 4), ("ACGTTGCA", 3), (None, 7), (None, 4), (None, 2)])
def test_locate_every_row(g, tmp_path, text, sa_intv):
    """Every row's reported position, by the sampled-suffix-array walk and
    from the full suffix array, equals the oracle's (the reference's walk,
    whose detour through row primary+1 gives some rows another value)."""
    if text is None:
        rng = random.Random(sa_intv)
        text = repetitive_text(rng, 3000 + 500 * sa_intv)
    n = len(text)
    prefix = build_index(tmp_path, text, sa_intv)
    ox = ob.Index(prefix)
    bi = brute.BruteIndex(text, sa_intv)
    rows = np.arange(n + 3, dtype=np.uint32)  # rows past seq_len have no position
    want = [ox.locate(int(r)) for r in rows]
    detour = sum(1 for r in range(1, n + 1) if want[r] != bi.sa[r])
    say("n=%d sa_intv=%d: %d rows take the detour" % (n, sa_intv, detour))
    with g.Index(prefix) as X:
        walk = X.locate(rows)
        X.set_direct(True)
        direct = X.locate(rows)
    assert walk.tolist() == want
    assert direct.tolist() == want
    if text == "AC":
        assert detour == 1  # row 2 reports 0 instead of 1 (seed_gen.cu:455-457)


def edge_reads(text):
    """Reads that match at the very start and end of the text (a run reaching
    text position 0), on both strands, with and without an N inside."""
    out = []
    for s in (text[:150], text[-150:], text[:60], text[-40:]):
        out += [s.encode(), revcomp(s).encode()]
        m = len(s) // 2
        out.append((s[:m] + "N" + s[m + 1:]).encode())
        out.append((s[:m] + "NN" + s[m + 2:]).encode())
    return out


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20), (True, 16), (False, 17)])
def test_seeds_unchanged(g, tmp_path, smem, k):
    rng = random.Random(3 + k + smem)
    text = repetitive_text(rng, 40000, n_units=8, unit_len=200, copies=6, div=0.02, tandem=6)
    reads = sample_reads(rng, text, 700, 40, 160, sub=0.01, indel=0.002, n_frac=0.2)
    reads += sample_reads(rng, text, 25, 161, 1500, sub=0.005, indel=0.001, n_frac=0.3)
    reads += edge_reads(text)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    want, _ = ob.seed(ox, reads, k, smem=smem, prune=1, threads=8)
    want = [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in want]
    with g.Index(prefix) as X:
        X.set_kmer_len(16)
        X.set_direct(True)
        for kernel in (1, 2, 3):
            for dmin in (1, 3, 8):
                say("smem=%d k=%d kernel=%d direct_min=%d" % (smem, k, kernel, dmin))
                got = g.seed(X, reads, k=k, smem=smem, kernel=kernel, direct_min=dmin).as_lists()
                bad = [i for i in range(len(reads)) if got[i] != want[i]]
                assert not bad, "read %d" % bad[0]
        # without the extended table the runs start from 13-mer lookups
        X.set_kmer_len(13)
        say("smem=%d k=%d kernel=1 13-mer table" % (smem, k))
        got = g.seed(X, reads, k=k, smem=smem, kernel=1).as_lists()
        assert got == want


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20)])
@pytest.mark.parametrize("kernel,sat", [(1, 255), (2, 65535), (3, 15)])
def test_kernel_work_equals_the_oracle_work_model(g, tmp_path, smem, k, kernel, sat):
    """Reads held whole in the kernel's LDS window, no work splitting:
    table lookups, occ entries and direct reads (suffix array, text chunks,
    inverse) counted by the kernel equal the oracle's count."""
    rng = random.Random(17 + k)
    text = repetitive_text(rng, 30000, n_units=8, unit_len=200, copies=6, div=0.02, tandem=6)
    reads = sample_reads(rng, text, 600, 60, 160, sub=0.01, indel=0.002, n_frac=0.1) + edge_reads(text)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    arrays = g.pack_reads(reads)
    _, _, _, _, c = ob.seed_arrays(ox, *arrays, k, smem, threads=8, smem_stop=1, ext_len=16, prune_sat=sat,
                                   miss_skip=1, direct_min=3, tandem=6)
    with g.Index(prefix) as X:
        X.set_kmer_len(16)
        X.set_direct(True)
        b = g.Batch(X, arrays, k=k, smem=smem, kernel=kernel, split_tasks=g.NO_SPLIT, direct_min=3)
        st = b.run()
        b.close()
    say("smem=%d kernel %d: table %d occ %d direct %d; oracle %d %d %d" %
        (smem, kernel, st.n_table_reads, st.n_occ_reads, st.n_direct_reads, c["w_tasks"] + c["w_ext"],
         c["w_blocks96"], c["w_direct"]))
    assert c["w_direct"] > 0
    assert st.n_table_reads == c["w_tasks"] + c["w_ext"]
    assert st.n_occ_reads == c["w_blocks96"]
    assert st.n_direct_reads == c["w_direct"]
    assert st.n_tandem_reads == c["w_tandem"]
