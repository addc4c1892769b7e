"""The tandem-jump work model (or_opts.tandem), the device kernel's shortcut
through tandem repeats, and the tandem table it reads (or_tandem_row).

Inside a stretch of the read with period p, read[y, e) is its last p bases
repeated, so its interval depends only on that unit and on e - y: a task
reads it from the table instead of a lookup and the steps down to y.  The
table rows must equal the brute-force intervals of the repeated units; the
model must not change the seeds (checked against the brute-force SMEM and
MEM definitions on a text full of tandem arrays), every jumped-to interval
must equal the one recomputed from the read, and a jump that ends its task
must end it where one more base of the read empties the interval
(tandem_check); and the shortcut must actually run and cut the work."""
import random

import numpy as np
import pytest

import brute
import oracle_bind as ob
from conftest import build_index, reads_over, tandem_text

DEVICE = dict(smem_stop=1, ext_len=16, miss_skip=1, direct_min=4, prune_sat=15, pair_probe=2)


def as_tuples(res):
    return [[(x, y, l, u, list(p)) for (x, y, l, u, p) in r] for r in res]


@pytest.fixture(scope="module")
def small(built, tmp_path_factory):
    rng = random.Random(404)
    text, arrays = tandem_text(rng, 6000, n_arrays=30, len_min=30, len_max=300, n_units=8)
    reads = reads_over(rng, text, arrays, 70, 60, 160)
    # reads of one stretch from end to end, with and without an error, and
    # one with two stretches of different periods
    p0, ln = max(arrays, key=lambda a: a[1])
    pure = text[p0:p0 + min(ln, 150)]
    reads.append(pure.encode())
    mid = len(pure) // 2
    reads.append((pure[:mid] + ("A" if pure[mid] != "A" else "C") + pure[mid + 1:]).encode())
    reads.append((text[arrays[0][0]:arrays[0][0] + 60] + text[arrays[1][0]:arrays[1][0] + 60]).encode())
    d = tmp_path_factory.mktemp("tandem_small")
    prefix = build_index(d, text, 5)
    return dict(bi=brute.BruteIndex(text, 5), ox=ob.Index(prefix), reads=reads)


def unit_of(key):
    p = 1
    while p < 6 and key >= (4 ** (p + 1) - 4) // 3:
        p += 1
    code = key - (4 ** p - 4) // 3
    # code j is the base at distance j from the string's end
    return "".join("ACGT"[(code >> (2 * j)) & 3] for j in reversed(range(p)))


def repeated(unit, n):
    """the string of length n that ends with unit repeated"""
    s = unit * (n // len(unit) + 2)
    return s[len(s) - n:]


def test_tandem_row_definition(small):
    """Every length of the rows of units that occur as arrays in the text,
    of units that never occur and of short units, against brute force."""
    bi = small["bi"]
    text = bi.T
    keys = set(range(0, 20)) | {5459, 1364, 340, 84}
    for key in range(5460):
        u = unit_of(key)
        if len(u) >= 3 and u * 4 in text:
            keys.add(key)
    assert len(keys) > 40
    for key in sorted(keys):
        unit = unit_of(key)
        row = ob.tandem_row(small["ox"], key)
        m = 0
        for n in range(1, ob.TANDEM_LEN + 1):
            if bi.interval(repeated(unit, n)) is None:
                break
            m = n
        for n in range(20, ob.TANDEM_LEN + 1):
            want_m = min(m, n)
            iv = bi.interval(repeated(unit, want_m)) if want_m else (0, bi.n)
            assert tuple(row[n - 20]) == (iv[0], iv[1], want_m, 0), (key, unit, n)


@pytest.mark.parametrize("smem,k", [(True, 19), (False, 19), (True, 16), (False, 20)])
@pytest.mark.parametrize("span,length", [(24, 256), (20, 40)])
def test_tandem_model_matches_definition(small, smem, k, span, length):
    got, c = ob.seed(small["ox"], small["reads"], k, smem=smem, prune=1, tandem=6, tandem_check=1, span=span,
                     length=length, **DEVICE)
    want = (brute.smem if smem else brute.mem)(small["bi"], small["reads"], k)
    for i, (g, w) in enumerate(zip(as_tuples(got), want)):
        assert g == w, "read %d %r" % (i, small["reads"][i])
    assert c["w_tandem_bad"] == 0
    assert c["w_stretches"] > 0 and c["w_tandem"] > 0, c


def test_tandem_needs_pruning(small):
    """Without size records (prune 0) the model is off: no stretch opens."""
    _, c = ob.seed(small["ox"], small["reads"], 19, smem=True, prune=0, tandem=6, **DEVICE)
    assert c["w_stretches"] == 0 and c["w_tandem"] == 0


@pytest.fixture(scope="module")
def large(built, tmp_path_factory):
    rng = random.Random(77)
    text, arrays = tandem_text(rng, 200000, n_arrays=400, len_min=30, len_max=600, n_units=40)
    reads = reads_over(rng, text, arrays, 1500, 100, 160)
    d = tmp_path_factory.mktemp("tandem_large")
    prefix = build_index(d, text, 7)
    return dict(ox=ob.Index(prefix), arrays=ob.pack_reads(reads))


@pytest.mark.parametrize("opts", [dict(), dict(ext_len=0, miss_skip=0, pair_probe=0), dict(direct_min=0),
                                  dict(prune_sat=0)])
@pytest.mark.parametrize("tk", [dict(tandem=6), dict(tandem=1), dict(tandem=6, length=24), dict(tandem=6, span=20),
                                dict(tandem=8)])
def test_tandem_cuts_work_and_keeps_seeds(large, opts, tk):
    o = dict(DEVICE, **opts)
    buf, offs, lens = large["arrays"]
    off = ob.seed_arrays(large["ox"], np.frombuffer(buf, np.uint8), offs, lens, 19, True, max_occ=50, threads=8,
                         **o)
    on = ob.seed_arrays(large["ox"], np.frombuffer(buf, np.uint8), offs, lens, 19, True, max_occ=50, threads=8,
                        tandem_check=1, **o, **tk)
    for a, b in zip(off[:4], on[:4]):
        assert np.array_equal(a, b)
    co, cn = off[4], on[4]
    assert cn["w_tandem_bad"] == 0 and cn["w_stretches"] > 0
    work = lambda c: c["w_tasks"] + c["w_ext"] + c["w_pair"] + c["w_blocks96"] + c["w_direct"] + c["w_tandem"]
    if tk.get("tandem") != 1 and tk.get("length", 256) >= 64:
        assert work(cn) < 0.8 * work(co), (work(cn), work(co))
    else:
        assert work(cn) < work(co)
