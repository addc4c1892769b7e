"""The index writer (SA-IS + reference-format .bwt/.sa) against brute force.

The bytes of a reference-format index are fully determined by the text
(seed_gen.cu:52-58 block layout, 2685-2736 readers, build_index.sh:54-75), so
they are checked field by field against a suffix array made by sorting.
"""
import os
import random
import struct

import numpy as np
import pytest

import brute
from conftest import build_index, random_text, repetitive_text


def parse_bwt(prefix):
    raw = open(prefix + ".bwt", "rb").read()
    primary = struct.unpack_from("<I", raw, 0)[0]
    L2 = [0] + list(struct.unpack_from("<4I", raw, 4))
    body = np.frombuffer(raw[20:], dtype="<u4")
    return primary, L2, body


def parse_sa(prefix):
    raw = open(prefix + ".sa", "rb").read()
    h = struct.unpack_from("<7I", raw, 0)
    body = np.frombuffer(raw[28:], dtype="<u4")
    return h, body


def words(n):
    return (n + 15) // 16 + 4 * ((n + 63) // 64 + 1)


TEXTS = []
_r = random.Random(7)
for n in (1, 2, 3, 15, 16, 17, 63, 64, 65, 127, 128, 129, 200, 511, 640, 1000, 1024, 2049):
    TEXTS.append(random_text(_r, n))
TEXTS += ["A" * 300, "ACGT" * 70, "AC" * 151 + "G", repetitive_text(_r, 1500),
          "T" * 64 + "A" * 64, "GATTACA"]


@pytest.mark.parametrize("ti", range(len(TEXTS)))
@pytest.mark.parametrize("sa_intv", [1, 3, 7])
def test_index_files_match_brute_force(built, tmp_path, ti, sa_intv):
    text = TEXTS[ti]
    prefix = build_index(tmp_path, text, sa_intv)
    bi = brute.BruteIndex(text, sa_intv)
    n = len(text)
    primary, L2, body = parse_bwt(prefix)
    assert primary == bi.primary
    assert L2 == bi.L2
    assert len(body) == words(n)
    # $-removed BWT string, MSB-first 16 bases per word, 4 count words per 64 bases
    rem = [c for r, c in enumerate(bi.bwt) if r != bi.primary]
    cnt = [0, 0, 0, 0]
    for i in range(n):
        blk = (i // 64) * 8
        if i % 64 == 0:
            assert list(body[blk:blk + 4]) == cnt, "counts of block %d" % (i // 64)
        w = int(body[blk + 4 + (i % 64) // 16])
        c = (w >> ((15 - i % 16) * 2)) & 3
        assert "ACGT"[c] == rem[i]
        cnt[c] += 1
    assert list(body[-4:]) == cnt  # trailing counts-only block
    h, sa = parse_sa(prefix)
    assert h[0] == bi.primary and list(h[1:5]) == bi.L2[1:] and h[5] == sa_intv and h[6] == n
    n_sa = (n + sa_intv) // sa_intv
    assert len(sa) == n_sa - 1
    assert list(sa) == [bi.sa[i * sa_intv] for i in range(1, n_sa)]


def test_lf_roundtrip_recovers_text(built, tmp_path):
    """Walking LF from the $ row spells the text backwards (independent of SA)."""
    text = repetitive_text(random.Random(3), 3000)
    prefix = build_index(tmp_path, text, 7)
    primary, L2, body = parse_bwt(prefix)
    n = len(text)

    def char(idx):
        return (int(body[(idx // 64) * 8 + 4 + (idx % 64) // 16]) >> ((15 - idx % 16) * 2)) & 3

    rem = [char(i) for i in range(n)]
    occ_before = np.zeros((n + 1, 4), dtype=np.int64)
    for i, c in enumerate(rem):
        occ_before[i + 1] = occ_before[i]
        occ_before[i + 1][c] += 1
    row, out = 0, []
    for _ in range(n):
        idx = row - (1 if row > primary else 0)
        c = rem[idx]
        out.append("ACGT"[c])
        row = L2[c] + occ_before[idx + 1][c]
    assert "".join(reversed(out)) == text


def test_builder_rejects_bad_arguments(built, tmp_path):
    import gpuseed_amd
    with pytest.raises(gpuseed_amd.GsError):
        gpuseed_amd.build_index_host(np.zeros(0, dtype=np.uint8), 7, str(tmp_path / "x"))
    with pytest.raises(gpuseed_amd.GsError):
        gpuseed_amd.build_index_host(np.array([0, 1, 4], dtype=np.uint8), 7, str(tmp_path / "x"))
    with pytest.raises(gpuseed_amd.GsError):
        gpuseed_amd.build_index_host(np.array([0, 1, 2], dtype=np.uint8), 0, str(tmp_path / "x"))


def test_synthetic_generator_is_deterministic(built):
    import gpuseed_amd
    rp, qp = gpuseed_amd.preset(1)
    rp.length = 200000
    a = gpuseed_amd.synth_reference(rp, threads=1)
    b = gpuseed_amd.synth_reference(rp, threads=7)
    assert np.array_equal(a, b)
    assert a.max() <= 3
    gc = np.isin(a, [1, 2]).mean()
    assert abs(gc - 0.41) < 0.01
    qp.n_reads = 500
    r1 = gpuseed_amd.synth_reads(a, qp, threads=1)
    r2 = gpuseed_amd.synth_reads(a, qp, threads=5)
    for x, y in zip(r1, r2):
        assert np.array_equal(x, y)
    buf, offs, lens, origin, strand = r1
    assert (lens == 100).all()
    assert set(np.unique(buf)) <= set(b"ACGTN")
    nN = sum(1 for i in range(500) if b"N" in buf[offs[i]:offs[i] + lens[i]].tobytes())
    assert 0 < nN < 40  # 2 % of reads carry one N


def test_presets_follow_the_configurations(built):
    import gpuseed_amd
    r1, q1 = gpuseed_amd.preset(1)
    assert r1.length == 1000000 and q1.n_reads == 10000 and q1.classes[0].len_min == 100
    r2, q2 = gpuseed_amd.preset(2)
    assert r2.length == 100000000 and q2.n_reads == 1000000 and q2.classes[0].len_max == 150
    r3, q3 = gpuseed_amd.preset(3)
    assert r3.length == 3100000000 and q3.n_reads == 10000000
    r4, q4 = gpuseed_amd.preset(4)
    assert q4.n_classes == 3 and q4.classes[2].len_max == 20000
