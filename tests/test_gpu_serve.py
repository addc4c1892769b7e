"""The seeding server: continuous batching must not change any result.

Each request's result from the server (requests of many sizes submitted
together from several threads, mixed into shared device rounds, large ones
split over rounds) must equal gs_seed on that request alone with the plain
kernel, bit for bit, in SMEM and in MEM mode."""
import random
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g(built):
    import gpuseed_amd
    return gpuseed_amd


@pytest.fixture(scope="module")
def world(g, tmp_path_factory):
    """A 3 Mbp reference with the repeat structure of the 3.1 Gbp preset and a pool of
    uneven reads (100 bp to 20 kbp)."""
    d = tmp_path_factory.mktemp("serve")
    rp, qp = g.preset(4)
    rp.length = 3000000
    qp.n_reads = 6000
    ref = g.synth_reference(rp)
    buf, offs, lens, _, _ = g.synth_reads(ref, qp)
    prefix = str(d / "srv")
    g.build_index_host(ref, 7, prefix)
    return dict(prefix=prefix, buf=buf, offs=offs, lens=lens)


def make_requests(world, n_req, max_reads, seed):
    rng = random.Random(seed)
    buf, offs, lens = world["buf"], world["offs"], world["lens"]
    out = []
    for _ in range(n_req):
        m = int(round(np.exp(rng.uniform(0, np.log(max_reads)))))
        a = rng.randrange(0, len(lens) - m)
        lo, hi = int(offs[a]), int(offs[a + m - 1] + lens[a + m - 1])
        out.append((buf[lo:hi].copy(), (offs[a:a + m] - np.uint64(lo)).copy(), lens[a:a + m].copy()))
    out.append((np.zeros(1, np.uint8), np.zeros(0, np.uint64), np.zeros(0, np.uint32)))  # empty request
    return out


def same(a, b):
    return (np.array_equal(a.seeds, b.seeds) and np.array_equal(a.pos, b.pos) and
            np.array_equal(a.read_seed_off, b.read_seed_off) and np.array_equal(a.read_pos_off, b.read_pos_off))


@pytest.mark.parametrize("smem,k,round_bases,growth,tables", [(True, 19, 2000000, 0, False),
                                                              (False, 20, 1500000, 0, False),
                                                              (True, 19, 0, 0, False), (True, 19, 1000000, 8, False),
                                                              (False, 20, 1000000, 4, False),
                                                              (True, 19, 2000000, 4, True), (False, 20, 0, 0, True)])
def test_each_result_is_the_same_alone_and_in_the_mix(g, world, smem, k, round_bases, growth, tables):
    """tables: the index also holds the 16-mer table with presence bits and
    the direct-extension structures (the served rounds use both; the plain
    kernel the results are compared with uses the table only)."""
    reqs = make_requests(world, 40, 2000, seed=7 + k + round_bases % 7 + growth)
    with g.Index(world["prefix"]) as X:
        if tables:
            X.set_kmer_len(16)
            X.set_direct(True)
        alone = [g.seed(X, arrays=r, k=k, smem=smem, kernel=0, max_occ=200) for r in reqs]
        say("alone done: %d requests, %d reads" % (len(reqs), sum(len(r[2]) for r in reqs)))
        with g.Server(X, round_bases=round_bases, round_growth=growth, k=k, smem=smem, max_occ=200) as S:
            tickets = [None] * len(reqs)

            def submit(idx):
                for i in idx:
                    tickets[i] = S.submit(reqs[i])

            th = [threading.Thread(target=submit, args=(range(j, len(reqs), 4),)) for j in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            results = [None] * len(reqs)

            def wait(idx):
                for i in idx:
                    results[i] = S.wait(tickets[i])

            th = [threading.Thread(target=wait, args=(range(j, len(reqs), 3),)) for j in range(3)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            st = S.stats()
        say("rounds %d for %d requests, %d reads" % (st.n_rounds, st.n_requests, st.n_reads))
        assert st.n_requests == len(reqs)
        assert st.n_rounds < len(reqs) - 1          # requests shared rounds
        multi = sum(1 for _, info in results if info.n_rounds > 1)
        if round_bases:
            assert multi > 0                         # and large ones were split over rounds
        for i, (out, info) in enumerate(results):
            assert info.n_reads == len(reqs[i][2])
            assert same(out, alone[i]), "request %d (%d reads, %d rounds)" % (i, info.n_reads, info.n_rounds)
            assert info.t_submit <= info.t_start <= info.t_done


def test_exclusive_rounds_and_reuse_after_waits(g, world):
    reqs = make_requests(world, 12, 300, seed=3)
    with g.Index(world["prefix"]) as X:
        alone = [g.seed(X, arrays=r, k=19, smem=True, kernel=0) for r in reqs]
        with g.Server(X, exclusive=True, k=19, smem=True, kernel=0) as S:
            for rep in range(2):                     # the server keeps working after waits
                ts = [S.submit(r) for r in reqs]
                outs = [S.wait(t) for t in ts]
                for i, (out, info) in enumerate(outs):
                    assert same(out, alone[i]), (rep, i)
                    assert info.n_rounds == (1 if info.n_reads else 0)
            assert S.stats().n_rounds == 2 * (len(reqs) - 1)


def test_unknown_ticket_is_an_error(g, world):
    with g.Index(world["prefix"]) as X:
        with g.Server(X, k=19, smem=True) as S:
            with pytest.raises(g.GsError):
                S.wait(12345)


def test_failed_round_fails_its_requests_only(g, world):
    """A round that fails (injected before its upload, as a failed upload
    would) fails the request it belongs to; that request is collected while
    its later chunks are still planned, and the requests submitted after it
    are served and equal gs_seed alone (MEM mode, small reference batches)."""
    buf, offs, lens = world["buf"], world["offs"], world["lens"]

    def req(a, m):
        lo, hi = int(offs[a]), int(offs[a + m - 1] + lens[a + m - 1])
        return buf[lo:hi].copy(), (offs[a:a + m] - np.uint64(lo)).copy(), lens[a:a + m].copy()

    kw = dict(k=20, smem=False, batch_bases=20000)
    A, B, C = req(0, 100), req(200, 30), req(400, 5)
    assert int(A[2].sum()) > 4 * 20000  # A spans several reference batches, one per round
    with g.Index(world["prefix"]) as X:
        alone = [g.seed(X, arrays=r, kernel=0, **kw) for r in (B, C)]
        S = g.Server(X, round_bases=1, round_growth=1, workers=1, **kw)
        try:
            S.inject_failure(2, g.GS_EHIP)
            ta = S.submit(A)
            with pytest.raises(g.GsError) as ei:
                S.wait(ta)
            assert ei.value.rc == g.GS_EHIP
            for r, want in zip((B, C), alone):
                out, info = S.wait(S.submit(r))
                assert same(out, want) and info.n_reads == len(r[2])
        finally:
            with pytest.raises(g.GsError) as ei:  # the server reports the failed round
                S.close()
            assert ei.value.rc == g.GS_EHIP


def test_large_request_handed_off_in_parts(g, world):
    """One request of 300,000 150 bp reads in a round of its own: its results
    are copied to the caller in parts (131,072 reads, 131,072 seed records
    and 1,048,576 positions a part, on several threads) and must still equal
    gs_seed on the request alone."""
    rp, _ = g.preset(4)  # the reference of the world's index
    rp.length = 3000000
    _, qp = g.preset(3)  # 150 bp reads
    qp.n_reads = 300000
    ref = g.synth_reference(rp)
    arrays = g.synth_reads(ref, qp)[:3]
    with g.Index(world["prefix"]) as X:
        alone = g.seed(X, arrays=arrays, k=19, smem=True, kernel=3)
        say("alone: %d seeds, %d positions" % (len(alone.seeds), len(alone.pos)))
        assert len(alone.seeds) > 131072 and len(alone.pos) > 1048576  # several parts of each kind
        with g.Server(X, exclusive=True, k=19, smem=True) as S:
            out, info = S.wait(S.submit(arrays))
        assert info.n_reads == 300000 and info.n_rounds == 1
        assert same(out, alone)
