"""Batched pipeline variants: tuning fields, the captured HIP graph and the
autotuner.  None of them may change the output (bit for bit against the
plain kernel)."""
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
    rng = random.Random(77)
    text = repetitive_text(rng, 20000, n_units=10, unit_len=200, copies=8, div=0.02, tandem=5)
    reads = sample_reads(rng, text, 1500, 100, 150, sub=0.01, indel=0.002, n_frac=0.05)
    reads += sample_reads(rng, text, 60, 161, 3000, sub=0.04, indel=0.02, n_frac=0.2)
    rng.shuffle(reads)
    prefix = build_index(tmp_path_factory.mktemp("pipe"), text, 5)
    return dict(prefix=prefix, arrays=g.pack_reads(reads))


def same(a, b):
    return (np.array_equal(a.seeds, b.seeds) and np.array_equal(a.pos, b.pos) and
            np.array_equal(a.read_seed_off, b.read_seed_off) and np.array_equal(a.read_pos_off, b.read_pos_off))


def first_diff(a, b, lens):
    for r in range(len(lens)):
        if a.read(r) != b.read(r):
            return "read %d (len %d): %r vs %r" % (r, lens[r], a.read(r)[:4], b.read(r)[:4])
    return "counts only"


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20)])
def test_tuning_fields_do_not_change_output(g, case, smem, k):
    import oracle_bind as ob
    ox = ob.Index(case["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *case["arrays"], k=k, smem=smem, max_occ=50, threads=8)
    ox.close()
    lens = case["arrays"][2]
    with g.Index(case["prefix"]) as X:
        ref = g.seed(X, arrays=case["arrays"], k=k, smem=smem, kernel=0, max_occ=50)
        ok_ref = (np.array_equal(so, ref.read_seed_off) and np.array_equal(po, ref.read_pos_off) and
                  np.array_equal(seeds, ref.seeds.view(np.uint32).reshape(-1, 6)) and np.array_equal(pos, ref.pos))
        assert ok_ref, "plain kernel differs from the oracle"
        bad = []
        for seg_tasks in (64, 48):
            for kernel in (1, 2):
                for wg in (128, 64, 256):
                    for walk_rows in (1, 2, 4):
                        out = g.seed(X, arrays=case["arrays"], k=k, smem=smem, kernel=kernel, max_occ=50, wg=wg,
                                     walk_rows=walk_rows, seg_tasks=seg_tasks)
                        if not same(out, ref):
                            bad.append((seg_tasks, kernel, wg, walk_rows, first_diff(out, ref, lens)))
                            say("differs:", bad[-1])
        assert not bad, bad[:3]


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 20)])
def test_work_splitting_does_not_change_output(g, case, smem, k):
    """Items hand their remaining task ends to idle lanes after split_work
    memory operations; very small budgets make nearly every stream split."""
    with g.Index(case["prefix"]) as X:
        ref = g.seed(X, arrays=case["arrays"], k=k, smem=smem, kernel=0, max_occ=50)
        for kernel, split_tasks, split_work in ((1, 1, 1), (1, 3, 1), (2, 8, 2), (1, 64, 1), (1, 8, 32),
                                                (2, 5, 4)):
            say("smem=%d kernel=%d split_tasks=%d split_work=%d" % (smem, kernel, split_tasks, split_work))
            out = g.seed(X, arrays=case["arrays"], k=k, smem=smem, kernel=kernel, max_occ=50,
                         split_tasks=split_tasks, split_work=split_work, seg_tasks=48)
            assert same(out, ref), (kernel, split_tasks, split_work,
                                    first_diff(out, ref, case["arrays"][2]))


def test_bad_tuning_fields_are_refused(g, case):
    with g.Index(case["prefix"]) as X:
        for kw in (dict(wg=96), dict(walk_rows=3), dict(kernel=3), dict(split_tasks=65), dict(split_work=0xFFFF)):
            with pytest.raises(g.GsError):
                g.seed(X, arrays=case["arrays"], k=19, smem=True, **kw)


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_graph_replays_equal_eager_runs(g, case, kernel):
    """Every replay, with eager work and synchronisations between replays,
    gives the eager result."""
    with g.Index(case["prefix"]) as X:
        other = g.Batch(X, case["arrays"], k=20, smem=False, kernel=1)
        b = g.Batch(X, case["arrays"], k=19, smem=True, kernel=kernel, max_occ=100)
        b.run()
        eager = b.fetch()
        b.capture()
        for rep in range(4):
            say("replay %d kernel %d" % (rep, kernel))
            s = b.replay()
            assert s.n_seeds == len(eager.seeds) and s.n_pos == len(eager.pos)
            assert same(b.fetch(), eager), rep
            other.run()               # eager work on the same stream in between
            other.fetch()
        b.run()                       # an eager run after replays
        assert same(b.fetch(), eager)
        s = b.replay()                # and a replay after the eager run
        assert same(b.fetch(), eager)
        other.close()
        b.close()


def test_autotune_picks_valid_options(g, case):
    with g.Index(case["prefix"]) as X:
        opts, info = g.autotune(X, case["arrays"], reps=2, k=19, smem=True, max_occ=100)
        say("tuned kernel=%d wg=%d walk_rows=%d seg_tasks=%d: %.3f ms (default %.3f ms, %d tried)" %
            (opts.kernel, opts.wg, opts.walk_rows, opts.seg_tasks, info.seconds * 1e3,
             info.default_seconds * 1e3, info.n_tried))
        assert opts.kernel in (1, 2) and opts.wg in (64, 128, 256) and opts.walk_rows in (1, 2, 4)
        assert opts.seg_tasks in (32, 64, 96, 112)
        assert opts.k == 19 and opts.smem == 1 and opts.max_occ == 100
        # the two times come from separate measurements: allow timing noise
        assert info.n_tried >= 9 and 0 < info.seconds <= info.default_seconds * 1.25
        tuned = g.seed(X, arrays=case["arrays"], opts=opts)
        plain = g.seed(X, arrays=case["arrays"], k=19, smem=True, max_occ=100, kernel=0)
        assert same(tuned, plain)
