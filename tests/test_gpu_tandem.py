"""Tandem jumps of the batched kernels (gs_opts.tandem) on a text full of
tandem arrays: the device's tandem table equals the oracle's rows, the seeds
equal the oracle's, and the kernels' memory operations (table lookups, occ
entries, direct reads and tandem-table entries read) equal the oracle's
work model of the same shortcut (or_opts.tandem), for reads held whole in
the LDS window with no work splitting.  Long reads (segments), work
splitting and graph replays only need equal seeds."""
import random
import sys

import numpy as np
import pytest

import oracle_bind as ob
from conftest import build_index, reads_over, tandem_text

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
    rng = random.Random(2024)
    text, arrays = tandem_text(rng, 80000, n_arrays=160, len_min=30, len_max=500, n_units=24)
    reads = reads_over(rng, text, arrays, 900, 60, 160)
    long_reads = reads_over(rng, text, arrays, 30, 400, 1500)
    d = tmp_path_factory.mktemp("tandem")
    return dict(prefix=build_index(d, text, 7), short=g.pack_reads(reads), mixed=g.pack_reads(reads + long_reads))


def test_tandem_table_equals_the_oracle(g, case):
    ox = ob.Index(case["prefix"])
    with g.Index(case["prefix"]) as X:
        got = X.read_tandem(0, ob.TANDEM_KEYS * ob.TANDEM_ROWS).reshape(ob.TANDEM_KEYS, ob.TANDEM_ROWS, 4)
    long_rows = 0
    for key in range(ob.TANDEM_KEYS):
        want = ob.tandem_row(ox, key)
        assert np.array_equal(got[key], want), key
        long_rows += int(want[-1, 2] >= 100)
    ox.close()
    assert long_rows >= 10  # the text's arrays give rows that occur for 100 bases and more
    say("tandem table: %d keys, %d with repeats of 100 bases or more" % (ob.TANDEM_KEYS, long_rows))


def compare(out, want, what):
    seeds, pos, so, po = want
    assert np.array_equal(out.read_seed_off, so) and np.array_equal(out.read_pos_off, po), what
    assert np.array_equal(out.seeds.view(np.uint32).reshape(-1, 6), seeds), what
    assert np.array_equal(out.pos, pos), what


@pytest.mark.parametrize("index", ["ext16_direct_pairs", "ext16", "kmer13"])
@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20)])
def test_tandem_seeds_and_work_equal_the_model(g, case, index, smem, k):
    arrays = case["short"]
    ox = ob.Index(case["prefix"])
    want = ob.seed_arrays(ox, *arrays, k, smem, threads=8)[:4]
    ext = 0 if index == "kmer13" else 16
    direct = index == "ext16_direct_pairs"
    with g.Index(case["prefix"]) as X:
        if ext:
            X.set_kmer_len(16)
        if direct:
            X.set_direct(True)
            X.set_pairs(True)
        for kernel, sat in ((3, 15), (1, 255), (2, 65535)):
            for tandem in (0, 2, g.NO_TANDEM):
                tp = {0: 6, 2: 2}.get(tandem, 0)
                say("index %s smem=%d k=%d kernel %d tandem %d" % (index, smem, k, kernel, tp))
                b = g.Batch(X, arrays, k=k, smem=smem, kernel=kernel, split_tasks=g.NO_SPLIT, direct_min=3,
                            tandem=tandem)
                st = b.run()
                compare(b.fetch(), want, "kernel %d tandem %d" % (kernel, tp))
                b.close()
                c = ob.seed_arrays(ox, *arrays, k, smem, threads=8, smem_stop=1, ext_len=ext, prune_sat=sat,
                                   miss_skip=1 if ext else 0, direct_min=3 if direct else 0,
                                   pair_probe=2 if direct else 0, tandem=tp, tandem_check=1)[4]
                say("  kernel table %d occ %d direct %d tandem %d; model %d %d %d %d (stretches %d)" %
                    (st.n_table_reads, st.n_occ_reads, st.n_direct_reads, st.n_tandem_reads,
                     c["w_tasks"] + c["w_ext"] + c["w_pair"], c["w_blocks96"], c["w_direct"], c["w_tandem"],
                     c["w_stretches"]))
                assert c["w_tandem_bad"] == 0
                assert (c["w_tandem"] > 0) == (tp > 0)
                assert st.n_table_reads == c["w_tasks"] + c["w_ext"] + c["w_pair"]
                assert st.n_occ_reads == c["w_blocks96"]
                assert st.n_direct_reads == c["w_direct"]
                assert st.n_tandem_reads == c["w_tandem"]
    ox.close()


def test_tandem_with_segments_splitting_and_replays(g, case):
    arrays = case["mixed"]
    ox = ob.Index(case["prefix"])
    want = ob.seed_arrays(ox, *arrays, 19, True, max_occ=200, threads=8)[:4]
    ox.close()
    with g.Index(case["prefix"]) as X:
        X.set_kmer_len(16)
        X.set_direct(True)
        for kw in (dict(kernel=3), dict(kernel=3, seg_tasks=7), dict(kernel=1, split_tasks=1, split_work=1),
                   dict(kernel=3, split_tasks=2, split_work=1), dict(kernel=2, wg=64), dict(kernel=3, wg=256)):
            say("mixed lengths %s" % kw)
            compare(g.seed(X, arrays=arrays, k=19, smem=True, max_occ=200, **kw), want, str(kw))
        say("graph replays")
        b = g.Batch(X, arrays, k=19, smem=True, max_occ=200, kernel=3)
        st = b.run()
        assert st.n_tandem_reads > 0
        compare(b.fetch(), want, "run")
        b.capture()
        for i in range(3):
            b.replay()
            compare(b.fetch(), want, "replay %d" % i)
        b.close()


def test_tandem_option_limits(g, case):
    with g.Index(case["prefix"]) as X:
        with pytest.raises(g.GsError):
            g.seed(X, arrays=case["short"], k=19, smem=True, kernel=3, tandem=7)
