"""HIP path (libgpuseed.so through its C ABI) against the oracle and brute force."""
import os
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


def oracle_lists(ox, reads, k, smem, max_occ=0):
    got, _ = ob.seed(ox, reads, k, smem=smem, prune=1, max_occ=max_occ, threads=8)
    return [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in got]


def test_occ_layout_matches_bwt(g, tmp_path):
    rng = random.Random(5)
    for n in (1, 95, 96, 97, 500, 2000):
        text = repetitive_text(rng, n) if n > 100 else random_text(rng, n)
        say("occ layout n=%d" % n)
        prefix = build_index(tmp_path, text, 3, "o%d" % n)
        bi = brute.BruteIndex(text, 3)
        with g.Index(prefix) as X:
            info = X.info()
            assert info.seq_len == n and info.primary == bi.primary and list(info.L2) == bi.L2
            nb = (n + 95) // 96
            assert info.n_blocks == nb
            occ = X.read_occ(0, nb * 4)
            rem = [c for r, c in enumerate(bi.bwt) if r != bi.primary]
            for b in range(nb):
                for c in range(4):
                    cnt, w0, w1, w2 = (int(v) for v in occ[b * 4 + c])
                    assert cnt == sum(1 for ch in rem[:96 * b] if ch == "ACGT"[c])
                    bits = w0 | (w1 << 32) | (w2 << 64)
                    want = sum(1 << q for q in range(96) if 96 * b + q < n and rem[96 * b + q] == "ACGT"[c])
                    assert bits == want, (n, b, c)


def test_kmer_table_matches_oracle_steps(g, tmp_path):
    rng = random.Random(8)
    text = repetitive_text(rng, 4000)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    bi = brute.BruteIndex(text, 7)
    with g.Index(prefix) as X:
        ps = rng.sample(range(4 ** 13), 3000)
        # every 13-mer that occurs, plus random (mostly absent) ones
        for i in range(0, len(text) - 13, 7):
            ps.append(sum("ACGT".index(text[i + 12 - j]) << (2 * j) for j in range(13)))
        for p in ps:
            l, u = X.read_kmer(p, 1)[0]
            ol, ou = 0, len(text)
            for j in range(13):
                ol, ou = ox.step(ol, ou, (p >> (2 * j)) & 3)
                if ol > ou:
                    break
            assert (int(l), int(u)) == (ol, ou), p
            pat = "".join("ACGT"[(p >> (2 * j)) & 3] for j in reversed(range(13)))
            iv = bi.interval(pat)
            if iv is not None:
                assert (int(l), int(u)) == iv


CASES = [
    (1, 400, False, 1, 25, (20, 60), 14),
    (2, 1200, True, 3, 30, (30, 120), 16),
    (3, 3000, True, 7, 30, (60, 150), 19),
    (4, 2500, True, 7, 25, (90, 110), 20),
    (5, 800, True, 5, 30, (15, 50), 14),
    (6, 5000, True, 7, 20, (170, 300), 19),   # reads beyond the 160-base LDS slots
    (7, 6000, True, 7, 8, (300, 900), 20),    # reads kept in device memory
]


@pytest.mark.parametrize("kernel", [0, 1, 2, 3])
@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("smem", [True, False])
def test_tiny_cases_match_oracle_and_brute_force(g, tmp_path, ci, smem, kernel):
    seed, n, rep, sa_intv, nr, (lmin, lmax), k = CASES[ci]
    say("case %d smem=%d kernel=%d" % (ci, smem, kernel))
    rng = random.Random(seed)
    text = repetitive_text(rng, n) if rep else random_text(rng, n)
    reads = sample_reads(rng, text, nr, lmin, lmax)
    for _ in range(3):
        p = rng.randrange(0, n - 40)
        reads.append(text[p:p + 40].encode())
    prefix = build_index(tmp_path, text, sa_intv)
    ox = ob.Index(prefix)
    want = oracle_lists(ox, reads, k, smem)
    if n <= 3000:
        assert want == (brute.smem if smem else brute.mem)(brute.BruteIndex(text, sa_intv), reads, k)
    with g.Index(prefix) as X:
        got = g.seed(X, reads, k=k, smem=smem, kernel=kernel).as_lists()
    for i, (a, b) in enumerate(zip(got, want)):
        assert a == b, "read %d %r" % (i, reads[i])


@pytest.mark.parametrize("kernel", [0, 1, 3])
def test_mem_batch_lookahead_on_device(g, tmp_path, kernel):
    rng = random.Random(99)
    text = random_text(rng, 3000)
    prefix = build_index(tmp_path, text, 1)
    ox = ob.Index(prefix)
    fwd = text[1200:1240].encode()
    rev = revcomp(text[2000:2040]).encode()
    reads = [fwd, fwd, rev, rev, fwd]
    with g.Index(prefix) as X:
        got = g.seed(X, reads, k=20, smem=False, kernel=kernel).as_lists()
    assert got == oracle_lists(ox, reads, 20, False)


@pytest.mark.parametrize("smem", [True, False])
def test_occurrence_cap_on_device(g, tmp_path, smem):
    rng = random.Random(3)
    text = repetitive_text(rng, 3000, n_units=6, copies=12, div=0.0)
    reads = sample_reads(rng, text, 40, 40, 100)
    prefix = build_index(tmp_path, text, 7)
    ox = ob.Index(prefix)
    with g.Index(prefix) as X:
        for cap in (1, 4, 1000):
            got = g.seed(X, reads, k=16, smem=smem, max_occ=cap).as_lists()
            assert got == oracle_lists(ox, reads, 16, smem, max_occ=cap), cap


@pytest.fixture(scope="module")
def config1(g, tmp_path_factory):
    d = tmp_path_factory.mktemp("config1")
    rp, qp = g.preset(1)
    ref = g.synth_reference(rp)
    buf, offs, lens, _, _ = g.synth_reads(ref, qp)
    prefix = os.path.join(str(d), "c1")
    g.build_index_host(ref, 7, prefix)
    reads = [buf[int(o):int(o) + int(l)].tobytes() for o, l in zip(offs, lens)]
    return dict(prefix=prefix, reads=reads, arrays=(buf, offs, lens), dir=str(d))


@pytest.mark.parametrize("kernel", [0, 1, 3])
@pytest.mark.parametrize("smem,k", [(False, 20), (True, 20), (True, 19)])
def test_config1_records_and_text_match_oracle(g, config1, smem, k, kernel):
    """Configuration 1 (1 Mbp reference, 10,000 x 100 bp reads), all records
    and the -o text byte for byte."""
    say("config1 smem=%d k=%d kernel=%d" % (smem, k, kernel))
    ox = ob.Index(config1["prefix"])
    want = oracle_lists(ox, config1["reads"], k, smem)
    txt = os.path.join(config1["dir"], "gpu_%d_%d.txt" % (smem, k))
    with g.Index(config1["prefix"]) as X:
        fd = os.open(txt, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        try:
            out = g.seed(X, arrays=config1["arrays"], k=k, smem=smem, text_fd=fd, kernel=kernel)
        finally:
            os.close(fd)
    got = out.as_lists()
    assert len(got) == len(want)
    bad = [i for i in range(len(got)) if got[i] != want[i]]
    assert not bad, "first mismatching read %d: %r vs %r" % (bad[0], got[bad[0]], want[bad[0]])
    import ctypes as C
    L = ob.lib()
    L.or_format_text.argtypes = [C.POINTER(ob.OrResult), C.c_int, C.c_uint64, C.c_void_p]
    buf, offs, lens = ob.pack_reads(config1["reads"])
    o = ob.OrOpts(k, int(smem), 1, 0, 8, 0)
    res = ob.OrResult()
    assert L.or_seed_reads(ox.h, buf, offs.ctypes.data_as(C.POINTER(C.c_uint64)),
                           lens.ctypes.data_as(C.POINTER(C.c_uint32)), len(config1["reads"]),
                           C.byref(o), C.byref(res)) == 0
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    otxt = txt + ".oracle"
    fp = libc.fopen(otxt.encode(), b"w")
    L.or_format_text(C.byref(res), int(smem), 1, fp)
    libc.fclose(fp)
    L.or_result_free(C.byref(res))
    assert open(txt, "rb").read() == open(otxt, "rb").read()


def test_repeated_calls_are_identical(g, config1):
    with g.Index(config1["prefix"]) as X:
        a = g.seed(X, arrays=config1["arrays"], k=19, smem=True)
        b = g.seed(X, arrays=config1["arrays"], k=19, smem=True)
    assert np.array_equal(a.seeds, b.seeds) and np.array_equal(a.pos, b.pos)


def test_empty_batch_and_short_reads(g, config1):
    with g.Index(config1["prefix"]) as X:
        for kernel in (0, 1):
            out = g.seed(X, [], k=20, smem=True, kernel=kernel)
            assert out.n_reads == 0 and len(out.seeds) == 0
            assert list(out.read_seed_off) == [0] and list(out.read_pos_off) == [0]
        out = g.seed(X, [b"", b"ACGT", b"N" * 30], k=20, smem=False)
        assert out.as_lists() == [[], [], []]
