"""Reads longer than the persistent kernel's LDS window: segmented work items.

Reads of more than 160 bases are cut into items of seg_tasks task ends; the
join kernel concatenates the segments and settles the SMEM filter at their
boundaries.  The output must not depend on seg_tasks and must equal the plain
kernel's and the oracle's, bit for bit."""
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
def long_case(tmp_path_factory):
    rng = random.Random(41)
    text = repetitive_text(rng, 8000, n_units=8, unit_len=300, copies=5, div=0.01, tandem=4)
    reads = sample_reads(rng, text, 40, 161, 2500, sub=0.03, indel=0.01, n_frac=0.3)
    # exact copies of long stretches: extensions that run far past a segment window
    for p in (100, 2000, 5000):
        reads.append(text[p:p + 1800].encode())
    reads.append(b"N".join([text[3000:3200].encode()] * 4))
    prefix = build_index(tmp_path_factory.mktemp("long"), text, 7)
    return dict(prefix=prefix, reads=reads)


def oracle_lists(prefix, reads, k, smem, max_occ=0):
    ox = ob.Index(prefix)
    got, _ = ob.seed(ox, reads, k, smem=smem, prune=1, max_occ=max_occ, threads=8)
    ox.close()
    return [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in got]


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20), (True, 14)])
@pytest.mark.parametrize("seg_tasks", [1, 7, 64, 112])
def test_segments_match_oracle(g, long_case, smem, k, seg_tasks):
    say("long reads smem=%d k=%d seg_tasks=%d" % (smem, k, seg_tasks))
    want = oracle_lists(long_case["prefix"], long_case["reads"], k, smem)
    with g.Index(long_case["prefix"]) as X:
        for kernel in (1, 3):  # 8-bit and 4-bit size records
            got = g.seed(X, long_case["reads"], k=k, smem=smem, kernel=kernel, seg_tasks=seg_tasks).as_lists()
            bad = [i for i in range(len(want)) if got[i] != want[i]]
            assert not bad, "kernel %d read %d (len %d)" % (kernel, bad[0], len(long_case["reads"][bad[0]]))


@pytest.mark.parametrize("smem", [True, False])
def test_k_above_the_window(g, long_case, smem):
    """k = 170: reads of 161-169 bases are longer than the LDS window but have
    no task; longer ones are segmented; every read's seeds equal the oracle's."""
    k = 170
    say("long reads k=%d smem=%d" % (k, smem))
    reads = [r for r in long_case["reads"] if len(r) > 150][:300]
    reads += [r[:165] for r in reads[:20]] + [r[:170] for r in reads[20:30]]
    ox = ob.Index(long_case["prefix"])
    want, _ = ob.seed(ox, reads, k, smem=smem, prune=1, threads=8)
    ox.close()
    want = [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in want]
    with g.Index(long_case["prefix"]) as X:
        for kernel in (0, 1, 3):
            got = g.seed(X, reads, k=k, smem=smem, kernel=kernel).as_lists()
            bad = [i for i in range(len(want)) if got[i] != want[i]]
            assert not bad, "kernel %d read %d (len %d)" % (kernel, bad[0], len(reads[bad[0]]))


def test_seg_tasks_out_of_range_is_refused(g, long_case):
    with g.Index(long_case["prefix"]) as X:
        with pytest.raises(g.GsError):
            g.seed(X, long_case["reads"][:2], k=19, smem=True, kernel=1, seg_tasks=113)


@pytest.fixture(scope="module")
def mix(g, tmp_path_factory):
    """The uneven read mix (70 % 100-250 bp, 25 % 1-5 kbp, 5 % 5-20 kbp)
    on a 4 Mbp reference with the repeat structure of the 3.1 Gbp preset."""
    d = tmp_path_factory.mktemp("mix")
    rp, qp = g.preset(4)
    rp.length = 4000000
    qp.n_reads = 1500
    ref = g.synth_reference(rp)
    buf, offs, lens, _, _ = g.synth_reads(ref, qp)
    prefix = str(d / "mix")
    g.build_index_host(ref, 7, prefix)
    return dict(prefix=prefix, arrays=(buf, offs, lens))


@pytest.mark.parametrize("smem", [True, False])
def test_config4_mix_batched_equals_plain_and_oracle(g, mix, smem):
    k = 19
    say("config-4 mix smem=%d" % smem)
    buf, offs, lens = mix["arrays"]
    assert lens.max() > 5000 and lens.min() >= 100
    with g.Index(mix["prefix"]) as X:
        plain = g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=0)
        runs = {
            "kernel 1": g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=1),
            "kernel 2 seg 40": g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=2, seg_tasks=40),
            "kernel 1 no split": g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=1, split_tasks=g.NO_SPLIT),
            "kernel 3": g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=3),
            "kernel 3 seg 32": g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=3, seg_tasks=32),
        }
    for name, out in runs.items():
        assert np.array_equal(out.seeds, plain.seeds), name
        assert np.array_equal(out.pos, plain.pos), name
        assert np.array_equal(out.read_seed_off, plain.read_seed_off), name
    ox = ob.Index(mix["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=8)
    ox.close()
    assert np.array_equal(so, plain.read_seed_off) and np.array_equal(po, plain.read_pos_off)
    assert np.array_equal(seeds, plain.seeds.view(np.uint32).reshape(-1, 6))
    assert np.array_equal(pos, plain.pos)


@pytest.mark.parametrize("smem", [True, False])
def test_config4_mix_with_tables_and_direct_extension(g, mix, smem):
    """The benchmark's configuration-4 set-up: the 16-mer table with presence
    bits and the direct-extension structures, the default kernel (4-bit
    records), with and without forced work splitting; bit-exact to the
    oracle."""
    k = 19
    buf, offs, lens = mix["arrays"]
    ox = ob.Index(mix["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, buf, offs, lens, k, smem, max_occ=500, threads=8)
    ox.close()
    with g.Index(mix["prefix"]) as X:
        X.set_kmer_len(16)
        X.set_direct(True)
        for kernel, kw in ((3, {}), (3, dict(seg_tasks=32, direct_min=1)), (1, dict(direct_min=2)),
                           (3, dict(split_tasks=2, split_work=1, direct_min=1))):
            say("config-4 mix with tables smem=%d kernel=%d %s" % (smem, kernel, kw))
            out = g.seed(X, arrays=mix["arrays"], k=k, smem=smem, kernel=kernel, max_occ=500, **kw)
            assert np.array_equal(so, out.read_seed_off) and np.array_equal(po, out.read_pos_off), (kernel, kw)
            assert np.array_equal(seeds, out.seeds.view(np.uint32).reshape(-1, 6)), (kernel, kw)
            assert np.array_equal(pos, out.pos), (kernel, kw)
