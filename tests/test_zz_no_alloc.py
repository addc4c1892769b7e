"""gs_seed_noalloc: the seeding pipeline with every device buffer in the
library's fixed scratch array and the results in the caller's arrays.  Its
output must equal gs_seed's bit for bit, on two calls in a row (the second
with a smaller input, so that it runs over what the first left in the
array), and sizes beyond the array or the caller's capacities must be
refused with GS_ERANGE."""
import random
import sys

import numpy as np
import pytest

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
    rng = random.Random(907)
    text = repetitive_text(rng, 50000, n_units=10, unit_len=250, copies=8, div=0.02, tandem=6)
    short = sample_reads(rng, text, 3000, 100, 160, sub=0.01, indel=0.002, n_frac=0.05)
    longer = sample_reads(rng, text, 300, 100, 1500, sub=0.01, indel=0.002, n_frac=0.1)
    prefix = build_index(tmp_path_factory.mktemp("noalloc"), text, 7)
    return dict(prefix=prefix, short=short, longer=longer)


def same(a, b):
    return (np.array_equal(a.seeds, b.seeds) and np.array_equal(a.pos, b.pos) and
            np.array_equal(a.read_seed_off, b.read_seed_off) and np.array_equal(a.read_pos_off, b.read_pos_off) and
            np.array_equal(a.read_nseeds, b.read_nseeds) and np.array_equal(a.read_npos, b.read_npos))


@pytest.mark.parametrize("smem,k,kernel,tables,reads,batch_bases", [
    (True, 19, 3, False, "short", 0),      # persistent kernel, 4-bit records
    (False, 20, 3, False, "short", 3000),  # MEM with small reference batches: ids from the device
    (True, 19, 1, True, "short", 0),       # 16-mer table and direct extension
    (True, 19, 0, False, "short", 0),      # plain kernel
    (False, 20, 3, True, "longer", 0),     # reads past 256 bases: plain kernel with its read scratch
])
def test_noalloc_equals_gs_seed_on_two_calls(g, case, smem, k, kernel, tables, reads, batch_bases):
    pool = case[reads]
    first, second = pool, pool[: len(pool) // 7]
    with g.Index(case["prefix"]) as X:
        if tables:
            X.set_kmer_len(16)
            X.set_direct(True)
        for label, rs in (("first", first), ("second, smaller", second)):
            say("noalloc smem=%d k=%d kernel=%d tables=%d %s call: %d reads" %
                (smem, k, kernel, tables, label, len(rs)))
            arrays = g.pack_reads(rs)
            kw = dict(k=k, smem=smem, kernel=kernel, max_occ=300, batch_bases=batch_bases)
            want = g.seed(X, arrays=arrays, **kw)
            got = g.seed_noalloc(X, arrays, **kw)
            assert len(want.pos) > 0
            assert same(got, want), label


def test_noalloc_refuses_sizes_beyond_its_limits(g, case):
    with g.Index(case["prefix"]) as X:
        # one read of 10 million bases: its candidate slots alone (32 bytes a
        # base) exceed the 256 MiB array; refused before any device work
        big = (np.frombuffer(b"ACGT", np.uint8)[np.arange(10_000_000) % 4].copy(),
               np.zeros(1, np.uint64), np.array([10_000_000], np.uint32))
        with pytest.raises(g.GsError) as ei:
            g.seed_noalloc(X, big, k=19, smem=True, kernel=3)
        assert ei.value.rc == g.GS_ERANGE
        # output capacities too small: refused, with the sizes the call needs
        arrays = g.pack_reads(case["short"][:500])
        want = g.seed(X, arrays=arrays, k=19, smem=True, kernel=3)
        with pytest.raises(g.GsError) as ei:
            g.seed_noalloc(X, arrays, seeds_cap=0, pos_cap=0, k=19, smem=True, kernel=3)
        assert ei.value.rc == g.GS_ERANGE
        assert ei.value.n_seeds == len(want.seeds) and ei.value.n_pos == len(want.pos)
        # exactly the sizes it asked for
        got = g.seed_noalloc(X, arrays, seeds_cap=ei.value.n_seeds, pos_cap=ei.value.n_pos, k=19, smem=True,
                             kernel=3)
        assert same(got, want)
        # an empty call
        empty = (np.zeros(1, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32))
        out = g.seed_noalloc(X, empty, k=19, smem=True, kernel=3)
        assert out.n_reads == 0 and len(out.seeds) == 0 and int(out.read_seed_off[0]) == 0
