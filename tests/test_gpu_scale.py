"""Parity at the benchmark configurations' own scale.

Configuration 3: the 3.1 Gbp synthetic reference indexed on the GPU.  The
index is first checked against its own text: walking it from every sampled
suffix-array row re-derives every row's position and the text (each walk
must land on the position the next sampled row holds, and the walks must
visit every row once), and the derived text must equal the generated
reference.  A correct text this way means a correct BWT, occurrence counts
and sampled suffix array, so the oracle, which reads the same index files,
checks the seeds against a known-good index.  Then 100,000 reads are seeded
against the oracle on the index the benchmark uses (16-mer table with
presence bits, direct-extension structures, 4-bit-record kernel, positions
of intervals of at most 500 rows) and on the reference's own index shape
(13-mer table, sampled-suffix-array walk at sampling 7), and requests of the
served workload (configuration 5: uneven lengths, 100 bp to 20 kbp) go
through the server on the same index against the oracle.

Configuration 2: the 100 Mbp reference and all 1,000,000 reads, plain and
batched kernels, against the oracle."""
import os
import shutil
import sys
import tempfile
import threading

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu
THREADS = 16  # host threads of the oracle


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g(built):
    import gpuseed_amd
    return gpuseed_amd


def scratch_dir(need):
    where = [d for d in (tempfile.gettempdir(), "/dev/shm")
             if os.path.isdir(d) and os.access(d, os.W_OK) and shutil.disk_usage(d).free > need]
    if not where:
        pytest.skip("no %.1f GB of scratch space for the index files" % (need / 1e9))
    return tempfile.mkdtemp(prefix="gs_scale_", dir=where[0])


def check_text(X, ref, step=1 << 28):
    n = len(ref)
    for first in range(0, n, step):
        cnt = min(step, n - first)
        t = X.read_text(first, cnt)
        if not np.array_equal(t, ref[first:first + cnt]):
            bad = first + int(np.argmax(t != ref[first:first + cnt]))
            raise AssertionError("derived text differs from the reference at position %d" % bad)


def check_L2(X, ref):
    counts = np.bincount(ref, minlength=4).astype(np.int64)
    assert list(X.info().L2) == [0] + list(np.cumsum(counts))


def compare(out, want, what):
    seeds, pos, so, po = want
    assert np.array_equal(out.read_seed_off, so) and np.array_equal(out.read_pos_off, po), what + ": counts"
    got = out.seeds.view(np.uint32).reshape(-1, 6)
    if not np.array_equal(got, seeds):
        i = int(np.argmax(np.any(got != seeds, axis=1)))
        r = int(np.searchsorted(so, i, side="right") - 1)
        raise AssertionError("%s: read %d seed %d: %s vs %s" % (what, r, i, got[i], seeds[i]))
    assert np.array_equal(out.pos, pos), what + ": positions"


@pytest.fixture(scope="module")
def c3(g):
    rp, qp = g.preset(3)
    d = scratch_dir(rp.length * 1.2 + (1 << 30))
    say("config 3: generating the %.2f Gbp reference" % (rp.length / 1e9))
    ref = g.synth_reference(rp)
    prefix = os.path.join(d, "c3")
    say("config 3: building the index on the GPU, files under %s" % d)
    X = g.Index(codes=ref, sa_intv=7, write_prefix=prefix)
    yield dict(ref=ref, prefix=prefix, X=X, qp=qp, dir=d)
    X.close()
    shutil.rmtree(d, ignore_errors=True)


@pytest.fixture(scope="module")
def c3_reads(g, c3):
    qp = c3["qp"]
    qp.n_reads = 100000
    buf, offs, lens, _, _ = g.synth_reads(c3["ref"], qp)
    say("config 3: oracle on %d reads" % len(lens))
    ox = ob.Index(c3["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, buf, offs, lens, 19, True, max_occ=500, threads=THREADS)
    ox.close()
    say("config 3: oracle done, %d seeds, %d positions" % (len(seeds), len(pos)))
    return (buf, offs, lens), (seeds, pos, so, po)


def test_config3_index_is_its_text(g, c3):
    X, ref = c3["X"], c3["ref"]
    info = X.info()
    assert info.seq_len == len(ref) and info.sa_intv == 7
    check_L2(X, ref)
    say("config 3: direct structures (walks checked against the sampled suffix array)")
    X.set_direct(True)  # GS_EFORMAT if a walk lands wrong or a row is missed
    say("config 3: derived text against the generated reference")
    check_text(X, ref)


def test_config3_files_reopen_to_the_same_device_index(g, c3):
    """The .bwt/.sa written by the GPU builder (the oracle reads these) load
    through the reference-format reader into the same occurrence entries
    and 13-mer table."""
    X = c3["X"]
    with g.Index(c3["prefix"]) as Y:
        a, b = X.info(), Y.info()
        assert (a.seq_len, a.primary, list(a.L2), a.n_blocks) == (b.seq_len, b.primary, list(b.L2), b.n_blocks)
        total = (a.n_blocks + 1) * 4
        step = 1 << 24
        for first in range(0, total, step):
            cnt = min(step, total - first)
            assert np.array_equal(X.read_occ(first, cnt), Y.read_occ(first, cnt)), first
        assert np.array_equal(X.read_kmer(0, 4 ** 13), Y.read_kmer(0, 4 ** 13))


def test_config3_sample_on_the_benchmark_index(g, c3, c3_reads):
    X = c3["X"]
    arrays, want = c3_reads
    X.set_kmer_len(16)
    X.set_direct(True)
    say("config 3: 100k reads, 16-mer table + direct extension, kernel 3 (drop-in call)")
    compare(g.seed(X, arrays=arrays, k=19, smem=True, max_occ=500, kernel=3), want, "gs_seed kernel 3")
    say("config 3: the same through a resident batch and its captured graph")
    b = g.Batch(X, arrays, k=19, smem=True, max_occ=500, kernel=3)
    b.run()
    compare(b.fetch(), want, "batch run")
    b.capture()
    b.replay()
    compare(b.fetch(), want, "graph replay")
    b.close()


def test_config3_sample_on_the_reference_shaped_index(g, c3, c3_reads):
    """13-mer table only and no direct structures: positions come from the
    sampled-suffix-array walk at sampling 7, as in the reference."""
    X = c3["X"]
    arrays, want = c3_reads
    X.set_kmer_len(13)
    X.set_direct(False)
    assert X.info().kmer_len == 13 and X.info().direct == 0
    for kernel in (3, 0):
        say("config 3: 100k reads, 13-mer table, suffix-array walk, kernel %d" % kernel)
        compare(g.seed(X, arrays=arrays, k=19, smem=True, max_occ=500, kernel=kernel), want, "kernel %d" % kernel)


def test_config5_requests_served_on_the_config3_index(g, c3):
    """Requests of the served workload (the configuration-4 length and error
    mix, log-uniform sizes) submitted together from two threads; each
    request's result equals the oracle's on that request alone."""
    X = c3["X"]
    X.set_kmer_len(16)
    X.set_direct(True)
    rp, qp = g.preset(5)
    qp.n_reads = 20000
    buf, offs, lens, _, _ = g.synth_reads(c3["ref"], qp)
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(0.0, np.log(6000), 10)).round().astype(np.int64).clip(1, 6000)
    reqs = []
    for m in sizes:
        a = int(rng.integers(0, len(lens) - m + 1))
        lo, hi = int(offs[a]), int(offs[a + m - 1] + lens[a + m - 1])
        reqs.append((buf[lo:hi].copy(), (offs[a:a + m] - np.uint64(lo)).copy(), lens[a:a + m].copy()))
    say("config 5: %d requests, %d reads, %.1f Mbases" % (len(reqs), int(sizes.sum()),
                                                         sum(int(r[2].sum()) for r in reqs) / 1e6))
    ox = ob.Index(c3["prefix"])
    want = [ob.seed_arrays(ox, *r, 19, True, max_occ=500, threads=THREADS)[:4] for r in reqs]
    ox.close()
    results = [None] * len(reqs)
    with g.Server(X, k=19, smem=True, max_occ=500) as S:
        tickets = [None] * len(reqs)

        def submit(idx):
            for i in idx:
                tickets[i] = S.submit(reqs[i])

        th = [threading.Thread(target=submit, args=(range(j, len(reqs), 2),)) for j in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i, t in enumerate(tickets):
            results[i] = S.wait(t)
    X.set_kmer_len(13)  # the 68.7 GB table and the direct structures are not needed further
    X.set_direct(False)
    for i, (out, info) in enumerate(results):
        assert info.n_reads == len(reqs[i][2])
        compare(out, want[i], "request %d (%d reads)" % (i, info.n_reads))


@pytest.fixture(scope="module")
def c2(g):
    rp, qp = g.preset(2)
    d = scratch_dir(rp.length * 1.2 + (1 << 28))
    say("config 2: %.0f Mbp reference and %d reads" % (rp.length / 1e6, qp.n_reads))
    ref = g.synth_reference(rp)
    prefix = os.path.join(d, "c2")
    X = g.Index(codes=ref, sa_intv=7, write_prefix=prefix)
    buf, offs, lens, _, _ = g.synth_reads(ref, qp)
    yield dict(ref=ref, prefix=prefix, X=X, arrays=(buf, offs, lens))
    X.close()
    shutil.rmtree(d, ignore_errors=True)


def test_config2_in_full(g, c2):
    X, ref = c2["X"], c2["ref"]
    assert len(c2["arrays"][2]) == 1000000
    check_L2(X, ref)
    X.set_direct(True)
    check_text(X, ref)
    X.set_direct(False)
    say("config 2: oracle on 1,000,000 reads")
    ox = ob.Index(c2["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *c2["arrays"], 19, True, max_occ=500, threads=THREADS)
    ox.close()
    want = (seeds, pos, so, po)
    say("config 2: plain kernel (one read per lane), 13-mer table, suffix-array walk")
    compare(g.seed(X, arrays=c2["arrays"], k=19, smem=True, max_occ=500, kernel=0), want, "plain")
    X.set_kmer_len(16)
    X.set_direct(True)
    say("config 2: batched kernel 3, 16-mer table, direct extension")
    compare(g.seed(X, arrays=c2["arrays"], k=19, smem=True, max_occ=500, kernel=3), want, "batched")
