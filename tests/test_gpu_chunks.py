"""Launches split into chunks, and the lifetime rules of indexes, batches and
servers.

A dispatch's work-item count is a 32-bit quantity, so the table levels, the
presence bits, the direct-structure chains and the suffix-array walk launch
in chunks of at most 2^30 work-items.  At that size only a 3.1 Gbp index
goes past the first chunk; gs_set_launch_chunk makes the chunks small, so
that a small index runs every chunked loop over many chunks, and every
structure and seed must equal the one built with the default chunk."""
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


@pytest.fixture(scope="module")
def case(g, tmp_path_factory):
    rng = random.Random(501)
    text = repetitive_text(rng, 60000, n_units=10, unit_len=300, copies=8, div=0.02, tandem=6)
    reads = sample_reads(rng, text, 800, 100, 160, sub=0.01, indel=0.002, n_frac=0.05)
    reads += sample_reads(rng, text, 20, 161, 1200, sub=0.01, indel=0.002, n_frac=0.1)
    d = tmp_path_factory.mktemp("chunks")
    return dict(prefix1=build_index(d, text, 1, "s1"), prefix7=build_index(d, text, 7, "s7"),
                arrays=g.pack_reads(reads), n=len(text))


def build_all(g, prefix, table_chunk, chain_chunk):
    """Index with the 14-mer table and the direct structures, built with the
    given launch chunks."""
    g.set_launch_chunk(table_chunk)
    try:
        X = g.Index(prefix)
        X.set_kmer_len(14)
        g.set_launch_chunk(chain_chunk)
        X.set_direct(True)
    finally:
        g.set_launch_chunk(0)
    return X


@pytest.mark.parametrize("sa_intv", [1, 7])
def test_chunked_launches_build_the_same_index(g, case, sa_intv):
    prefix = case["prefix%d" % sa_intv]
    n = case["n"]
    # 65,537 (not a multiple of the 256-lane workgroup): 1,024 launches per
    # 13-mer level and 4,096 per 14-mer pass; 997: the chains of sa_intv 1
    # (n + 1 of them) in 61 launches
    A = build_all(g, prefix, 0, 0)
    B = build_all(g, prefix, 65537, 997)
    try:
        say("sa_intv %d: 13-mer tables" % sa_intv)
        assert np.array_equal(A.read_kmer(0, 4 ** 13), B.read_kmer(0, 4 ** 13))
        say("sa_intv %d: 14-mer table around every chunk boundary" % sa_intv)
        total = 4 ** 14
        for b in range(65537, total, 65537):
            lo, hi = b - 3, min(total, b + 3)
            assert np.array_equal(A.read_kext(lo, hi - lo), B.read_kext(lo, hi - lo)), b
        idx = np.random.default_rng(sa_intv).integers(0, total - 64, 200)
        for i in idx:
            assert np.array_equal(A.read_kext(int(i), 64), B.read_kext(int(i), 64)), i
        say("sa_intv %d: direct structures" % sa_intv)
        sa_a, isa_a, t_a = A.read_direct(0, n + 1)
        sa_b, isa_b, t_b = B.read_direct(0, n + 1)
        assert np.array_equal(sa_a, sa_b) and np.array_equal(isa_a, isa_b) and np.array_equal(t_a[:n], t_b[:n])
    finally:
        A.close()
        B.close()


def test_chunked_locate_and_seeding(g, case):
    """The suffix-array walk in launches of 97 rows (and of 1 row) gives the
    positions of one launch, and the seeds equal the oracle's."""
    prefix = case["prefix7"]
    ox = ob.Index(prefix)
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *case["arrays"], 19, True, threads=8)
    ox.close()
    assert len(pos) > 1000
    with g.Index(prefix) as X:
        rows = np.arange(0, case["n"] + 1, dtype=np.uint32)
        want = X.locate(rows)
        for chunk in (97, 1):
            say("locate and seeding in launches of %d rows" % chunk)
            g.set_launch_chunk(chunk)
            try:
                got = X.locate(rows if chunk > 1 else rows[:3000])
                out = g.seed(X, arrays=case["arrays"], k=19, smem=True, kernel=3)
            finally:
                g.set_launch_chunk(0)
            assert np.array_equal(got, want[:len(got)]), chunk
            assert np.array_equal(out.read_seed_off, so) and np.array_equal(out.read_pos_off, po)
            assert np.array_equal(out.seeds.view(np.uint32).reshape(-1, 6), seeds)
            assert np.array_equal(out.pos, pos)


def test_launch_chunk_limits(g):
    with pytest.raises(g.GsError):
        g.set_launch_chunk((1 << 30) + 1)
    g.set_launch_chunk(0)


def test_batch_outlives_the_index_handle(g, case):
    """A batch keeps its index: closing the caller's handle first leaves the
    batch's runs and results intact."""
    X = g.Index(case["prefix7"])
    want = g.seed(X, arrays=case["arrays"], k=19, smem=True, kernel=3)
    b = g.Batch(X, case["arrays"], k=19, smem=True, kernel=3)
    X.close()
    b.run()
    got = b.fetch()
    b.close()
    assert np.array_equal(got.seeds, want.seeds) and np.array_equal(got.pos, want.pos)


def test_structures_are_not_changed_under_a_live_server(g, case):
    with g.Index(case["prefix7"]) as X:
        with g.Server(X, k=19, smem=True) as S:
            for call in (lambda: X.set_kmer_len(14), lambda: X.set_direct(True)):
                with pytest.raises(g.GsError) as ei:
                    call()
                assert ei.value.rc == g.GS_EINVAL
            t = S.submit(case["arrays"])
            out, _ = S.wait(t)
        X.set_kmer_len(14)  # the server is gone
        X.set_direct(True)
        again = g.seed(X, arrays=case["arrays"], k=19, smem=True, kernel=3)
    assert np.array_equal(out.seeds, again.seeds) and np.array_equal(out.pos, again.pos)
