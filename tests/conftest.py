import os
import random
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def revcomp(s):
    return "".join(COMP[c] for c in reversed(s))


def random_text(rng, n, gc=0.41):
    w = [(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]
    return "".join(rng.choices("ACGT", weights=w, k=n))


def repetitive_text(rng, n, n_units=4, unit_len=40, copies=6, div=0.03, tandem=2):
    t = list(random_text(rng, n))
    for _ in range(n_units):
        unit = random_text(rng, unit_len)
        for _ in range(copies):
            p = rng.randrange(0, max(1, n - unit_len))
            for j, ch in enumerate(unit):
                if p + j < n:
                    t[p + j] = rng.choice("ACGT") if rng.random() < div else ch
    for _ in range(tandem):
        ul = rng.randint(1, 5)
        unit = random_text(rng, ul)
        reps = rng.randint(8, 30)
        p = rng.randrange(0, max(1, n - ul * reps))
        for j in range(min(ul * reps, n - p)):
            t[p + j] = unit[j % ul]
    return "".join(t)


def tandem_text(rng, n, n_arrays=60, pmax=6, len_min=30, len_max=400, n_units=12):
    """Random text with exact tandem arrays of periods 1..pmax; arrays draw
    their unit from a small pool, so a unit recurs in several arrays (large
    intervals, as in a genome's microsatellites).  Returns the text and the
    (start, length) of every array."""
    t = list(random_text(rng, n))
    units = [random_text(rng, rng.randint(1, pmax)) for _ in range(n_units)]
    arrays = []
    for _ in range(n_arrays):
        unit = rng.choice(units)
        ln = rng.randint(len_min, len_max)
        p = rng.randrange(0, n - ln)
        for j in range(ln):
            t[p + j] = unit[j % len(unit)]
        arrays.append((p, ln))
    return "".join(t), arrays


def reads_over(rng, text, arrays, nreads, lmin, lmax, sub=0.01, n_frac=0.05):
    """Reads (both strands) placed over the given (start, length) stretches:
    each read covers part of a stretch and often a flank, with substitutions
    and, in a fraction of reads, an N."""
    comp = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "N"}
    out = []
    n = len(text)
    for _ in range(nreads):
        L = rng.randint(lmin, lmax)
        p0, ln = rng.choice(arrays)
        p = min(max(0, p0 + rng.randint(-L // 2, ln - L // 2)), n - L)
        s = [rng.choice("ACGT".replace(ch, "")) if rng.random() < sub else ch for ch in text[p:p + L]]
        if rng.random() < n_frac:
            s[rng.randrange(L)] = "N"
        if rng.random() < 0.5:
            s = [comp[c] for c in reversed(s)]
        out.append("".join(s).encode())
    return out


def sample_reads(rng, text, nreads, lmin, lmax, sub=0.02, indel=0.005, n_frac=0.1, extra=True):
    """Reads as raw line bytes, sampled from both strands with errors."""
    out = []
    n = len(text)
    for _ in range(nreads):
        L = rng.randint(lmin, lmax)
        if L + 8 >= n:
            L = max(1, n - 8)
        p = rng.randrange(0, n - L - 4)
        s = []
        i = p
        while len(s) < L and i < n:
            x = rng.random()
            if x < indel / 2:
                s.append(rng.choice("ACGT"))
            elif x < indel:
                i += 1
            else:
                ch = text[i]
                if rng.random() < sub:
                    ch = rng.choice("ACGT".replace(ch, ""))
                s.append(ch)
                i += 1
        s = "".join(s)
        if rng.random() < 0.5:
            s = revcomp(s)
        if s and rng.random() < n_frac:
            q = rng.randrange(len(s))
            s = s[:q] + "N" + s[q + 1:]
        out.append(s.encode())
    if extra:
        out += [
            b"",                                   # empty line
            b"ACGTAC",                             # shorter than k
            (text[10:40] + "\r").encode(),         # CRLF line keeps '\r'
            text[5:45].lower().encode(),           # lower case
            ("N" * 5 + text[50:90]).encode(),      # leading Ns
            (text[100:125] + "NN" + text[127:160]).encode() if n > 170 else b"ACGT" * 8,
            b"ACGTRYKMSWBDHVN" * 3,                # IUPAC and others
        ]
    return out


@pytest.fixture(scope="session")
def built():
    """The product library and the oracle must be built beforehand."""
    import gpuseed_amd
    import oracle_bind
    gpuseed_amd.lib()
    oracle_bind.lib()
    return True


def build_index(tmpdir, text, sa_intv, name="idx"):
    import numpy as np

    import gpuseed_amd
    codes = np.frombuffer(text.encode().translate(bytes.maketrans(b"ACGT", b"\0\1\2\3")),
                          dtype=np.uint8)
    prefix = os.path.join(str(tmpdir), name)
    gpuseed_amd.build_index_host(codes, sa_intv, prefix)
    return prefix


@pytest.fixture
def rng():
    return random.Random(12345)
