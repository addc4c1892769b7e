"""Upstream bwa 0.7 files: the 64-bit .bwt/.sa with 128-base occurrence blocks
and the .pac/.ann/.amb reference files, written here from bwa 0.7's published
layout (bwt_dump_bwt / bwt_dump_sa / bns_dump and the pac packing of
bntseq.c).  No bwa binary or bwa-written file is available, so the layout is
parity unpinned against bwa itself; the reader is pinned against the fork
layout of the same text, which test_index_build pins by brute force."""
import os
import random
import struct

import numpy as np
import pytest

from conftest import random_text, revcomp

CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def sorted_suffixes(codes):
    """Suffix array of codes + $ ($ smallest) by sorting the suffixes."""
    b = bytes(codes)
    return sorted(range(len(b) + 1), key=lambda i: b[i:])


def write_bwa07_index(prefix, codes, sa_intv):
    """bwt_dump_bwt / bwt_dump_sa of bwa 0.7 for the text `codes`."""
    codes = [int(c) for c in codes]
    n = len(codes)
    sa = sorted_suffixes(codes)
    primary = sa.index(0)
    bwt = [codes[sa[row] - 1] for row in range(n + 1) if row != primary]
    L2 = [0] * 5
    for c in codes:
        L2[c + 1] += 1
    for c in range(2, 5):
        L2[c] += L2[c - 1]
    body, cnt = [], [0, 0, 0, 0]
    for i, ch in enumerate(bwt):
        if i % 128 == 0:
            for c in cnt:
                body += [c & 0xFFFFFFFF, c >> 32]
        if i % 16 == 0:
            body.append(0)
        body[-1] |= ch << ((15 - i % 16) * 2)
        cnt[ch] += 1
    for c in cnt:
        body += [c & 0xFFFFFFFF, c >> 32]
    assert len(body) == (n + 15) // 16 + 8 * ((n + 127) // 128 + 1)
    with open(prefix + ".bwt", "wb") as f:
        f.write(struct.pack("<5Q", primary, *L2[1:]))
        f.write(np.array(body, dtype=np.uint32).tobytes())
    n_sa = (n + sa_intv) // sa_intv
    with open(prefix + ".sa", "wb") as f:
        f.write(struct.pack("<7Q", primary, *L2[1:], sa_intv, n))
        f.write(np.array([sa[i * sa_intv] for i in range(1, n_sa)], dtype=np.uint64).tobytes())


def fork_to_bwa07(fork_prefix, prefix):
    """The same index in bwa 0.7's layout, from the fork layout's files."""
    raw = np.fromfile(fork_prefix + ".bwt", dtype=np.uint32)
    primary, L2 = int(raw[0]), [0] + [int(v) for v in raw[1:5]]
    n = L2[4]
    body = raw[5:]
    words = []                     # the base words, 16 bases each, in order
    for i in range(0, n, 16):
        words.append(int(body[(i >> 6) * 8 + 4 + ((i & 63) >> 4)]))
    out, cnt = [], [0, 0, 0, 0]
    for i in range(0, n, 16):
        if i % 128 == 0:
            for c in cnt:
                out += [c & 0xFFFFFFFF, c >> 32]
        w = words[i >> 4]
        out.append(w)
        for j in range(min(16, n - i)):
            cnt[(w >> ((15 - j) * 2)) & 3] += 1
    for c in cnt:
        out += [c & 0xFFFFFFFF, c >> 32]
    with open(prefix + ".bwt", "wb") as f:
        f.write(struct.pack("<5Q", primary, *L2[1:]))
        f.write(np.array(out, dtype=np.uint32).tobytes())
    sa = np.fromfile(fork_prefix + ".sa", dtype=np.uint32)
    with open(prefix + ".sa", "wb") as f:
        f.write(struct.pack("<7Q", primary, *L2[1:], int(sa[5]), int(sa[6])))
        f.write(sa[7:].astype(np.uint64).tobytes())


def write_bwa_ref(prefix, seqs, seed=11):
    """bns_fasta2bntseq + bns_dump of bwa 0.7 for [(name, anno, text)]: runs
    of non-ACGT become random bases in .pac and holes in .amb."""
    rng = random.Random(seed)
    codes, ann, holes = [], [], []
    for name, anno, text in seqs:
        off = len(codes)
        n_ambs, i = 0, 0
        while i < len(text):
            if text[i] in CODE:
                codes.append(CODE[text[i]])
                i += 1
                continue
            j = i
            while j < len(text) and text[j] == text[i]:
                j += 1
            holes.append((off + i, j - i, text[i]))
            n_ambs += 1
            codes += [rng.randrange(4) for _ in range(j - i)]
            i = j
        ann.append((name, anno, off, len(text), n_ambs))
    l_pac = len(codes)
    pac = bytearray((l_pac + 3) // 4)
    for l, c in enumerate(codes):
        pac[l >> 2] |= c << ((~l & 3) << 1)
    if l_pac % 4 == 0:
        pac.append(0)
    pac.append(l_pac % 4)
    with open(prefix + ".pac", "wb") as f:
        f.write(bytes(pac))
    with open(prefix + ".ann", "w") as f:
        f.write("%d %d %u\n" % (l_pac, len(seqs), seed))
        for gi, (name, anno, off, ln, na) in enumerate(ann):
            f.write("%d %s" % (gi, name) + (" %s\n" % anno if anno else "\n"))
            f.write("%d %d %d\n" % (off, ln, na))
    with open(prefix + ".amb", "w") as f:
        f.write("%d %d %u\n" % (l_pac, len(seqs), len(holes)))
        for off, ln, ch in holes:
            f.write("%d %d %c\n" % (off, ln, ch))
    return np.array(codes, dtype=np.uint8), ann, holes


def files(prefix):
    return [open(prefix + ext, "rb").read() for ext in (".bwt", ".sa")]


@pytest.mark.parametrize("n", [1, 15, 16, 17, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 3001])
def test_bwa07_index_converts_to_the_fork_layout(built, tmp_path, n):
    import gpuseed_amd as g
    rng = np.random.default_rng(n)
    codes = rng.integers(0, 4, n, dtype=np.uint8)
    for sa_intv in (1, 4, 7, 32):
        p07, pcv, pfork = (str(tmp_path / ("%s%d" % (s, sa_intv))) for s in ("b07_", "cv_", "fork_"))
        write_bwa07_index(p07, list(codes), sa_intv)
        g.index_convert(p07, pcv)
        g.build_index_host(codes, sa_intv, pfork)
        assert files(pcv) == files(pfork), (n, sa_intv)


def test_bwa07_text_of_both_strands(built, tmp_path):
    """bwa indexes the reference followed by its reverse complement."""
    import gpuseed_amd as g
    rng = random.Random(5)
    text = random_text(rng, 1500)
    both = text + revcomp(text)
    codes = [CODE[c] for c in both]
    write_bwa07_index(str(tmp_path / "b07"), codes, 7)
    g.index_convert(str(tmp_path / "b07"), str(tmp_path / "cv"))
    g.build_index_host(np.array(codes, dtype=np.uint8), 7, str(tmp_path / "fork"))
    assert files(str(tmp_path / "cv")) == files(str(tmp_path / "fork"))


def test_fork_files_convert_to_themselves(built, tmp_path):
    import gpuseed_amd as g
    codes = np.random.default_rng(3).integers(0, 4, 777, dtype=np.uint8)
    g.build_index_host(codes, 5, str(tmp_path / "fork"))
    g.index_convert(str(tmp_path / "fork"), str(tmp_path / "cv"))
    assert files(str(tmp_path / "cv")) == files(str(tmp_path / "fork"))


def test_damaged_bwa07_files_are_refused(built, tmp_path):
    import gpuseed_amd as g
    codes = list(np.random.default_rng(9).integers(0, 4, 600, dtype=np.uint8))
    p = str(tmp_path / "b07")

    def refused(patch_bwt=None, patch_sa=None):
        write_bwa07_index(p, codes, 4)
        for ext, patch in ((".bwt", patch_bwt), (".sa", patch_sa)):
            if patch:
                b = bytearray(open(p + ext, "rb").read())
                b = patch(b)
                open(p + ext, "wb").write(bytes(b))
        with pytest.raises(g.GsError) as e:
            g.index_convert(p, str(tmp_path / "cv"))
        return e.value.rc

    def bump_count(b):  # second block's count of A
        b[40 + 16 * 4] += 1
        return b

    def bump_sa_primary(b):
        b[0] ^= 1
        return b

    def set_sa_value(b):  # a sampled position beyond the text
        b[56:64] = struct.pack("<Q", 600)
        return b

    assert refused(patch_bwt=bump_count) == -5
    assert refused(patch_sa=bump_sa_primary) == -5
    assert refused(patch_sa=lambda b: b[:-8]) == -5
    assert refused(patch_sa=set_sa_value) == -5
    # a .bwt of the wrong length is not recognised as bwa 0.7 and fails the fork checks
    assert refused(patch_bwt=lambda b: b + b"\0\0\0\0") in (-5, -4)


@pytest.mark.parametrize("tail", [0, 1, 2, 3])
def test_bwa_reference_files(built, tmp_path, tail):
    import gpuseed_amd as g
    rng = random.Random(tail)
    seqs = [("chr1", "first one", random_text(rng, 400) + "N" * 7 + random_text(rng, 93)),
            ("chrM", "", random_text(rng, 61) + "R" + random_text(rng, 3 + tail)),
            ("scaffold_3", "x y z", "NN" + random_text(rng, 200))]
    codes, ann, holes = write_bwa_ref(str(tmp_path / "ref"), seqs)
    with g.BwaRef(str(tmp_path / "ref")) as R:
        assert len(R) == len(codes)
        assert np.array_equal(R.codes(), codes)
        assert R.contigs() == [(name, off, ln) for name, _, off, ln, _ in ann]
        assert R.holes() == holes
        for i, (name, _, off, ln, _) in enumerate(ann):
            assert R.locate(off) == (i, 0)
            assert R.locate(off + ln - 1) == (i, ln - 1)
        with pytest.raises(g.GsError) as e:
            R.locate(len(codes))
        assert e.value.rc == -7


def test_bwa_reference_damaged_pac_is_refused(built, tmp_path):
    import gpuseed_amd as g
    write_bwa_ref(str(tmp_path / "ref"), [("a", "", "ACGT" * 50 + "A")])
    p = str(tmp_path / "ref.pac")
    b = open(p, "rb").read()
    open(p, "wb").write(b[:-1] + bytes([(b[-1] + 1) % 4]))  # wrong l_pac % 4 byte
    with pytest.raises(g.GsError):
        g.BwaRef(str(tmp_path / "ref"))
    open(p, "wb").write(b + b"\0")  # wrong size
    with pytest.raises(g.GsError):
        g.BwaRef(str(tmp_path / "ref"))


def test_fork_to_bwa07_helper_round_trips(built, tmp_path):
    """The helper the GPU tests use to make bwa 0.7 files of large texts."""
    import gpuseed_amd as g
    codes = np.random.default_rng(12).integers(0, 4, 2345, dtype=np.uint8)
    g.build_index_host(codes, 7, str(tmp_path / "fork"))
    fork_to_bwa07(str(tmp_path / "fork"), str(tmp_path / "b07"))
    write_bwa07_index(str(tmp_path / "direct"), codes, 7)
    assert files(str(tmp_path / "b07")) == files(str(tmp_path / "direct"))


def test_build_index_tool_reads_bwa_files(built, tmp_path):
    """gs_build_index -b (index of bwa's .pac) and -c (bwa 0.7 .bwt/.sa to the
    fork layout) write the files the library's writer makes for that text."""
    import subprocess

    import gpuseed_amd as g
    tool = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpuseed_amd", "bin",
                        "gs_build_index")
    rng = random.Random(21)
    seqs = [("c1", "", random_text(rng, 900) + "NNNN" + random_text(rng, 300)), ("c2", "two", random_text(rng, 411))]
    codes, _, _ = write_bwa_ref(str(tmp_path / "ref"), seqs)
    g.build_index_host(codes, 5, str(tmp_path / "want"))
    p = subprocess.run([tool, "-r", "5", "-p", str(tmp_path / "got"), "-b", str(tmp_path / "ref")],
                       stderr=subprocess.PIPE, timeout=60)
    assert p.returncode == 0, p.stderr.decode()
    assert b"c2\toffset 1204\tlength 411" in p.stderr
    assert files(str(tmp_path / "got")) == files(str(tmp_path / "want"))
    write_bwa07_index(str(tmp_path / "b07"), codes, 5)
    p = subprocess.run([tool, "-p", str(tmp_path / "cv"), "-c", str(tmp_path / "b07")], stderr=subprocess.PIPE,
                       timeout=60)
    assert p.returncode == 0, p.stderr.decode()
    assert files(str(tmp_path / "cv")) == files(str(tmp_path / "want"))
    p = subprocess.run([tool, "-p", str(tmp_path / "x"), "-c", str(tmp_path / "missing")], stderr=subprocess.PIPE,
                       timeout=60)
    assert p.returncode == 1


def test_bwa07_text_beyond_32_bits_is_out_of_range(built, tmp_path):
    """Both strands of a human genome (6.2 G bases) exceed the 32-bit rows the
    device index uses: GS_ERANGE, decided from the header (the body of this
    file is sparse and never read)."""
    import gpuseed_amd as g
    n = 1 << 32
    p = str(tmp_path / "big")
    with open(p + ".bwt", "wb") as f:
        f.write(struct.pack("<5Q", 12345, n // 4, n // 2, 3 * n // 4, n))
    os.truncate(p + ".bwt", 40 + 4 * ((n + 15) // 16 + 8 * ((n + 127) // 128 + 1)))
    with pytest.raises(g.GsError) as e:
        g.index_convert(p, str(tmp_path / "cv"))
    assert e.value.rc == -7
