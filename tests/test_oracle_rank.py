"""Oracle rank machinery against brute force (occ, backward step, LF, SA walk)."""
import itertools
import random

import pytest

import brute
import oracle_bind as ob
from conftest import build_index, random_text, repetitive_text

_r = random.Random(11)
TEXTS = [random_text(_r, n) for n in (1, 5, 16, 63, 64, 65, 300, 777)]
TEXTS += ["A" * 129, "ACGT" * 40, repetitive_text(_r, 900)]


@pytest.fixture(scope="module", params=range(len(TEXTS)))
def pair(request, tmp_path_factory, built):
    text = TEXTS[request.param]
    d = tmp_path_factory.mktemp("rank%d" % request.param)
    sa_intv = (3, 7, 1, 4)[request.param % 4]
    prefix = build_index(d, text, sa_intv)
    return brute.BruteIndex(text, sa_intv), ob.Index(prefix)


def test_header(pair):
    bi, ox = pair
    assert ox.seq_len == bi.n and ox.primary == bi.primary and ox.L2() == bi.L2


def test_occ_every_row(pair):
    bi, ox = pair
    for c in range(4):
        acc = 0
        assert ox.occ(0xFFFFFFFF, c) == 0
        for row in range(bi.n + 1):
            acc += bi.bwt[row] == "ACGT"[c]
            assert ox.occ(row, c) == acc, (row, c)


def test_backward_step_gives_sa_intervals(pair):
    bi, ox = pair
    rng = random.Random(bi.n)
    pats = ["".join(p) for k in (1, 2, 3) for p in itertools.product("ACGT", repeat=k)]
    pats += [bi.T[i:i + rng.randint(1, 20)] for i in rng.sample(range(bi.n), min(40, bi.n))]
    for pat in pats:
        l, u = 0, bi.n
        for ch in reversed(pat):
            l, u = ox.step(l, u, "ACGT".index(ch))
            if l > u:
                break
        want = bi.interval(pat)
        if want is None:
            assert l > u, pat
        else:
            assert (l, u) == want, pat


def test_step_leaves_interval_unchanged_for_codes_above_3(pair):
    """seed_gen.cu:333 returns (l-1-L2[c], u-L2[c]) for c > 3 and 2173-2174
    add L2[c] back (l: +1); in 32-bit arithmetic that is (l, u) whatever
    L2[c] holds (c = 5..15 read past L2_gpu's five entries), l = 0 included.
    The oracle's step must give the same on arbitrary intervals."""
    bi, ox = pair
    M = 0xFFFFFFFF
    rng = random.Random(17)
    for c in (4, 5, 9, 15):
        assert ox.step(1, bi.n, c) == (1, bi.n)
        for _ in range(200):
            l, u = rng.randrange(0, bi.n + 1), rng.randrange(0, bi.n + 1)
            if rng.random() < 0.1:
                l = 0
            L2c = rng.choice([0, bi.n, M, rng.randrange(0, M + 1)])
            x, y = (l - 1 - L2c) & M, (u - L2c) & M  # find_occ_gpu, c > 3
            want = ((L2c + x + 1) & M, (L2c + y) & M)
            assert want == (l, u)
            assert ox.step(l, u, c) == want, (l, u, c)


def test_lf_matches_true_lf_except_row_after_primary(pair):
    bi, ox = pair
    for row in range(bi.n + 1):
        want = 0 if row == bi.primary + 1 else bi.lf_true(row)
        assert ox.lf(row) == want, row


def test_locate_every_row(pair):
    bi, ox = pair
    for row in range(1, bi.n + 1):
        assert ox.locate(row) == bi.locate(row), row


def test_locate_is_sa_when_walk_avoids_row_after_primary(pair):
    """Rows whose walk never enters primary+1 get their true suffix position."""
    bi, ox = pair
    for row in range(1, bi.n + 1):
        r, hit = row, False
        while r % bi.sa_intv:
            hit |= r == bi.primary + 1
            r = 0 if r in (bi.primary, bi.primary + 1) else bi.lf_true(r)
        if not hit:
            assert ox.locate(row) == bi.sa[row]


def test_true_sa_every_row(pair):
    """or_sa_true (the textbook walk, which checks the device's full suffix
    array) is the sorted-suffix order itself, row primary+1 included."""
    bi, ox = pair
    for row in range(bi.n + 1):
        assert ox.sa_true(row) == bi.sa[row], row


def test_known_answer_row_after_primary(built, tmp_path):
    """Hand-worked case of seed_gen.cu:455-457.  T = AC: suffixes $, AC$, C$ give
    SA = [2, 0, 1] and primary = 1.  With sa_intv = 4, row 2 (SA 1) walks once:
    the reference's LF decrements 2 to 1 == primary and returns row 0, whose
    stored value is -1, so it reports -1 + 1 = 0 instead of 1.  Row 1 walks to
    row 0 legitimately and reports 0, its true position."""
    ox = ob.Index(build_index(tmp_path, "AC", 4))
    assert ox.primary == 1
    assert ox.lf(2) == 0 and ox.lf(1) == 0
    assert ox.locate(2) == 0
    assert ox.locate(1) == 0


def test_encoding_of_every_byte(built):
    L = ob.lib()
    meaning = {"A": (0, 3), "C": (1, 2), "G": (2, 1), "T": (3, 0), "N": (4, 4)}
    for ch, (f, r) in meaning.items():
        for b in (ord(ch), ord(ch.lower())):
            assert (L.or_code_fwd(b), L.or_code_rev(b)) == (f, r), ch
    assert (L.or_code_fwd(ord("\r")), L.or_code_rev(ord("\r"))) == (0, 0)
    for b in range(256):
        assert (L.or_code_fwd(b), L.or_code_rev(b)) == (brute.code_fwd(b), brute.code_rev(b)), b
