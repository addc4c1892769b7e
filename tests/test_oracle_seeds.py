"""Oracle seeding (SMEM and MEM) against the brute-force definitions."""
import itertools
import os
import random

import pytest

import brute
import oracle_bind as ob
from conftest import build_index, random_text, repetitive_text, sample_reads

CASES = [
    # (seed, text length, repetitive, sa_intv, n reads, read lengths, k)
    (1, 400, False, 1, 25, (20, 60), 14),
    (2, 1200, True, 3, 30, (30, 120), 16),
    (3, 3000, True, 7, 30, (60, 150), 19),
    (4, 2500, True, 7, 25, (90, 110), 20),
    (5, 800, True, 5, 30, (15, 50), 14),
]


@pytest.fixture(scope="module", params=range(len(CASES)))
def case(request, tmp_path_factory, built):
    seed, n, rep, sa_intv, nr, (lmin, lmax), k = CASES[request.param]
    rng = random.Random(seed)
    text = repetitive_text(rng, n) if rep else random_text(rng, n)
    reads = sample_reads(rng, text, nr, lmin, lmax)
    # a few reads copied verbatim from repeats, to exercise many-hit intervals
    for _ in range(3):
        p = rng.randrange(0, n - 40)
        reads.append(text[p:p + 40].encode())
    d = tmp_path_factory.mktemp("seed%d" % seed)
    prefix = build_index(d, text, sa_intv)
    return dict(bi=brute.BruteIndex(text, sa_intv), ox=ob.Index(prefix), reads=reads, k=k)


def as_tuples(res):
    return [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in res]


@pytest.mark.parametrize("prune", [0, 1])
def test_smem_matches_definition(case, prune):
    got, _ = ob.seed(case["ox"], case["reads"], case["k"], smem=True, prune=prune)
    want = brute.smem(case["bi"], case["reads"], case["k"])
    for i, (g, w) in enumerate(zip(as_tuples(got), want)):
        assert g == w, "read %d %r" % (i, case["reads"][i])


@pytest.mark.parametrize("prune", [0, 1])
def test_mem_matches_definition(case, prune):
    got, _ = ob.seed(case["ox"], case["reads"], case["k"], smem=False, prune=prune)
    want = brute.mem(case["bi"], case["reads"], case["k"])
    for i, (g, w) in enumerate(zip(as_tuples(got), want)):
        assert g == w, "read %d %r" % (i, case["reads"][i])


@pytest.mark.parametrize("smem", [True, False])
def test_occurrence_cap(case, smem):
    cap = 3
    got, _ = ob.seed(case["ox"], case["reads"], case["k"], smem=smem, max_occ=cap)
    want = (brute.smem if smem else brute.mem)(case["bi"], case["reads"], case["k"], max_occ=cap)
    assert as_tuples(got) == want


def test_smem_survivors_are_not_contained(case):
    """Invariant: forward survivors of a read are pairwise non-nested, and so
    are reverse survivors."""
    got, _ = ob.seed(case["ox"], case["reads"], case["k"], smem=True)
    for seeds in got:
        for strand in (0, 1):
            iv = [(x & 0x7FFFFFFF, y & 0x7FFFFFFF) for x, y, l, u, p in seeds if (y >> 31) == strand]
            for a in iv:
                for b in iv:
                    if a != b:
                        assert not (b[0] <= a[0] and a[1] <= b[1])


def test_text_output_matches_reference_format(case, tmp_path):
    """The -o text, checked against a Python restatement of the printf formats
    at seed_gen.cu:3632-3646 over the brute-force seeds."""
    import ctypes as C
    got_dir = tmp_path / "out.txt"
    L = ob.lib()
    L.or_format_text.argtypes = [C.POINTER(ob.OrResult), C.c_int, C.c_uint64, C.c_void_p]
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    for smem in (True, False):
        buf, offs, lens = ob.pack_reads(case["reads"])
        o = ob.OrOpts(case["k"], int(smem), 1, 0, 2, 0)
        res = ob.OrResult()
        assert L.or_seed_reads(case["ox"].h, buf, offs.ctypes.data_as(C.POINTER(C.c_uint64)),
                               lens.ctypes.data_as(C.POINTER(C.c_uint32)), len(case["reads"]),
                               C.byref(o), C.byref(res)) == 0
        fp = libc.fopen(str(got_dir).encode(), b"w")
        L.or_format_text(C.byref(res), int(smem), 1, fp)
        libc.fclose(fp)
        L.or_result_free(C.byref(res))
        want = brute.format_text((brute.smem if smem else brute.mem)(case["bi"], case["reads"],
                                                                     case["k"]), smem)
        assert got_dir.read_text() == want


def test_format_known_answer(built, tmp_path):
    """Byte-level fixture of the -o format for a hand-built result."""
    want = ("/*===================================SMEM seeds in read no. 7 (Read[begin, End] -> "
            "starting position(s) on the reference)===================================*/\n"
            "\nRead[0, 25] -> +10, +400\nRead[3, 30] -> -77\n"
            "/*===================================SMEM seeds in read no. 8 (Read[begin, End] -> "
            "starting position(s) on the reference)===================================*/\n\n")
    res = [[(0, 25, 5, 6, [10, 400]), (3 | 1 << 31, 30 | 1 << 31, 9, 9, [77])], []]
    assert brute.format_text(res, True, first_read_no=7) == want


def test_pruning_does_not_change_config1_output(built, tmp_path):
    """Distance-1 pruning (the baseline algorithm) equals the full definition
    on a config-1-sized workload, for both modes."""
    import numpy as np

    import gpuseed_amd
    rp, qp = gpuseed_amd.preset(1)
    ref = gpuseed_amd.synth_reference(rp)
    qp.n_reads = 3000
    buf, offs, lens, _, _ = gpuseed_amd.synth_reads(ref, qp)
    reads = [buf[int(o):int(o) + int(l)].tobytes() for o, l in zip(offs, lens)]
    prefix = os.path.join(str(tmp_path), "c1")
    gpuseed_amd.build_index_host(ref, 7, prefix)
    ox = ob.Index(prefix)
    for smem in (True, False):
        a, ca = ob.seed(ox, reads, 20, smem, prune=0)
        b, cb = ob.seed(ox, reads, 20, smem, prune=1, use_table=0)
        assert a == b
        assert cb["w_steps"] < ca["w_steps"]
    del np


def test_mem_lookahead_crosses_reads_within_a_batch(built, tmp_path):
    """seeds_to_threads_mem looks for the next same-qbeg entry across the whole
    batch (seed_gen.cu:637) while the count stays inside the read (943).  Two
    identical forward reads from a unique region: each has only forward entries,
    all with qbeg 0 and the same one-row interval [l, l].  Read 1's last entry
    finds read 2's last entry, skips row l and emits row l + 1 instead; read 2
    has nothing after it and reports its true position.  Reverse-strand copies
    mirror this with the look-back (645): there the second read is affected."""
    rng = random.Random(99)
    text = random_text(rng, 3000)
    prefix = build_index(tmp_path, text, 1)
    bi, ox = brute.BruteIndex(text, 1), ob.Index(prefix)
    fwd = text[1200:1240].encode()
    got, _ = ob.seed(ox, [fwd, fwd], 20, smem=False)
    assert as_tuples(got) == brute.mem(bi, [fwd, fwd], 20)
    (x0, y0, l0, u0, p0), = got[0]
    (x1, y1, l1, u1, p1), = got[1]
    assert (x0, y0, l0) == (0, 40, u0) and p1 == [1200]
    assert p0 == [bi.sa[l0 + 1]]
    alone, _ = ob.seed(ox, [fwd], 20, smem=False)
    assert alone[0][0][4] == [1200]
    from conftest import revcomp
    rev = revcomp(text[2000:2040]).encode()
    got, _ = ob.seed(ox, [rev, rev], 20, smem=False)
    assert as_tuples(got) == brute.mem(bi, [rev, rev], 20)
    assert got[0][0][4] == [2000]
    assert got[1][0][4] == [bi.sa[got[1][0][2] + 1]]


def test_work_model_options_do_not_change_output(built, tmp_path):
    """The device kernels' shortcuts (a strand stops once a task reaches the
    read start in SMEM mode; an extended interval table answers the first
    14..16 bases; with k at least the table's length, a task whose clean
    window is absent ends at once; direct reads of the text replace the
    steps of single-row runs) change only the counted work, never the
    seeds."""
    import gpuseed_amd
    rp, qp = gpuseed_amd.preset(1)
    ref = gpuseed_amd.synth_reference(rp)
    qp.n_reads = 2000
    buf, offs, lens, _, _ = gpuseed_amd.synth_reads(ref, qp)
    prefix = os.path.join(str(tmp_path), "c1")
    gpuseed_amd.build_index_host(ref, 7, prefix)
    ox = ob.Index(prefix)
    for smem, k in ((True, 19), (False, 20), (True, 16), (False, 16), (True, 15)):
        base = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4)
        noskip = {}
        for stop, ext, skip, dmin in ((1, 0, 0, 0), (0, 16, 0, 0), (1, 14, 0, 0), (1, 16, 0, 0), (0, 16, 1, 0),
                                      (1, 14, 1, 0), (1, 16, 1, 0), (1, 16, 1, 4), (1, 0, 0, 1)):
            got = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4, smem_stop=stop, ext_len=ext,
                                 miss_skip=skip, direct_min=dmin)
            for a, b in zip(base[:4], got[:4]):
                assert (a == b).all(), (smem, k, stop, ext, skip)
            cb, cg = base[4], got[4]
            if stop and smem:
                assert cg["w_blocks96"] < cb["w_blocks96"]
            if ext:
                assert cg["w_ext"] > 0 and cg["w_steps"] < cb["w_steps"]
            if dmin:  # direct reads replace steps; the output is unchanged (checked above)
                assert cg["w_direct"] > 0 and cg["w_text"] > 0 and cg["w_direct"] > cg["w_text"]
                continue
            if not skip:
                noskip[(stop, ext)] = cg
            elif k >= ext:  # absent windows end their tasks: fewer 13-mer lookups
                assert cg["w_tasks"] < noskip[(stop, ext)]["w_tasks"], (smem, k, stop, ext)
            else:  # k below the window length: the option must not apply
                assert cg == noskip[(stop, ext)], (smem, k, stop, ext)
    # the pair-table model: with the 16-mer table and k >= 19, a lookup of an
    # 18-base core answers two adjacent windows (pair_probe 2: two cores in
    # one operation answer four); same seeds, fewer table reads; with k < 19
    # (or no 16-mer table) it must not apply
    for (smem, k), pp in itertools.product(((True, 19), (False, 20), (True, 16)), (1, 2)):
        off = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4, smem_stop=1, ext_len=16, miss_skip=1)
        for ext in (16, 14):
            on = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4, smem_stop=1, ext_len=ext, miss_skip=1,
                                pair_probe=pp)
            for a, b in zip(off[:4], on[:4]):
                assert (a == b).all(), (smem, k, ext)
            if ext != 16 or k < 19:
                ref = off if ext == 16 else ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4, smem_stop=1,
                                                         ext_len=ext, miss_skip=1)
                assert on[4] == ref[4], (smem, k, ext)
                continue
            co, cn = off[4], on[4]
            assert cn["w_pair"] > 0 and cn["w_ext"] + cn["w_pair"] < co["w_ext"], (smem, k)
            # one operation per byte, or fewer operations than bytes with two per operation
            assert (cn["w_pair_ops"] == cn["w_pair"]) == (pp == 1), (smem, k, pp)
            assert cn["w_ext"] < co["w_ext"]
            # everything after the table lookups is the same work
            assert all(cn[f] == co[f] for f in ("w_tasks", "w_steps", "w_blocks96", "w_direct")), (smem, k)
    # saturating size records only lose pruning: same seeds, at least as
    # much work, and more work the lower the saturation (repeats of the
    # 1 Mbp reference reach interval sizes of 10, so 4 and 8 lose prunes)
    for smem, k in ((True, 19), (False, 20)):
        exact = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4, smem_stop=1)
        prev = exact[4]["w_blocks96"]
        for sat in (255, 8, 4, 2):
            got = ob.seed_arrays(ox, buf, offs, lens, k, smem, threads=4, smem_stop=1, prune_sat=sat)
            for a, b in zip(exact[:4], got[:4]):
                assert (a == b).all(), (smem, sat)
            assert got[4]["w_blocks96"] >= prev, (smem, sat)
            prev = got[4]["w_blocks96"]
        assert prev > exact[4]["w_blocks96"]
    ox.close()
