"""Writes the golden fixtures of this directory from first principles.

    python tests/golden/make_golden.py

Inputs (deterministic): a 3,000-base reference with interspersed and tandem
repeats, and 48 read lines exercising what seed_gen.cu:2955-2990 reads and
what seeding treats specially: both strands, substitutions, an N inside and
inside the first 13 bases, lowercase, a carriage return, IUPAC letters, a read
shorter than k, an empty line, '>' lines, and a last line without a newline
(it loses its last byte, like the reference's getline loop).

Expected outputs come from tests/brute.py, which computes suffix arrays by
sorting and intervals by substring search; no FM-index code of the product or
of the oracle is involved.  The index files are assembled here byte by byte
from the brute-force BWT in the layout seed_gen reads (seed_gen.cu:52-58,
2685-2736): .bwt = primary, L2[1..4], 64-base blocks of four counts and four
MSB-first 16-base words, a trailing counts-only block; .sa = primary,
L2[1..4], sa_intv, seq_len, SA[i * sa_intv] for i >= 1.
"""
import os
import random
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import brute  # noqa: E402
from conftest import random_text, repetitive_text, revcomp  # noqa: E402

SA_INTV = 7
MODES = [("smem", 19), ("smem", 14), ("mem", 20), ("mem", 16)]


def make_inputs():
    rng = random.Random(20240601)
    text = repetitive_text(rng, 3000, n_units=5, unit_len=60, copies=6, div=0.03, tandem=3)
    lines = [b">golden reads"]
    for i in range(36):
        L = rng.choice([40, 75, 100, 150])
        p = rng.randrange(0, len(text) - L)
        s = list(text[p:p + L])
        for _ in range(rng.randint(0, 3)):
            q = rng.randrange(L)
            s[q] = rng.choice("ACGT")
        s = "".join(s)
        if i % 2:
            s = revcomp(s)
        lines.append(s.encode())
        if i % 9 == 0:
            lines.append(b">r%d" % i)
    base = text[1000:1100]
    lines.append((base[:50] + "N" + base[51:]).encode())        # N after the window
    lines.append((base[:5] + "N" + base[6:]).encode())          # N in the first 13 bases
    lines.append(base.lower().encode())                         # lowercase
    lines.append(base.encode() + b"\r")                         # CR kept as a base
    lines.append((base[:30] + "RYKMSW" + base[36:]).encode())   # IUPAC letters
    lines.append(text[200:215].encode())                        # shorter than k
    lines.append(b"")                                           # empty read
    lines.append(text[2500:2600].encode())
    lines.append(revcomp(text[1500:1620]).encode())
    lines.append(text[0:80].encode())                           # reference start
    lines.append(text[-80:].encode())                           # reference end
    last = text[700:790].encode()
    data = b"\n".join(lines) + b"\n" + last                     # no final newline
    return text, data


def reads_of(data):
    """seed_gen.cu:2955-2990: every line not starting with '>' is a read of its
    bytes minus the last one (the newline, or the last base at EOF)."""
    out = []
    for line in data.splitlines(keepends=True):
        if line.startswith(b">"):
            continue
        out.append(line[:-1])
    return out


def index_bytes(text):
    bi = brute.BruteIndex(text, SA_INTV)
    n = len(text)
    rem = [c for r, c in enumerate(bi.bwt) if r != bi.primary]
    full = []
    cnt = [0, 0, 0, 0]
    for b in range((n + 63) // 64):
        full += cnt
        # a partial last block keeps only the words that hold bases
        for w in range(min(4, (n - 64 * b + 15) // 16)):
            v = 0
            for j in range(16):
                i = 64 * b + 16 * w + j
                c = "ACGT".index(rem[i]) if i < n else 0
                v |= c << (2 * (15 - j))
            full.append(v)
        for i in range(64 * b, min(n, 64 * b + 64)):
            cnt["ACGT".index(rem[i])] += 1
    full += cnt  # trailing counts-only block
    words = (n + 15) // 16 + 4 * ((n + 63) // 64 + 1)
    assert len(full) == words, (len(full), words)
    bwt = struct.pack("<5I", bi.primary, *bi.L2[1:]) + struct.pack("<%dI" % len(full), *full)
    n_sa = (n + SA_INTV) // SA_INTV
    sa = struct.pack("<7I", bi.primary, *bi.L2[1:], SA_INTV, n)
    sa += struct.pack("<%dI" % (n_sa - 1), *[bi.sa[i * SA_INTV] for i in range(1, n_sa)])
    return bi, bwt, sa


def main():
    text, data = make_inputs()
    with open(os.path.join(HERE, "tiny_ref.txt"), "w") as f:
        f.write(text + "\n")
    with open(os.path.join(HERE, "tiny_reads.fa"), "wb") as f:
        f.write(data)
    bi, bwt, sa = index_bytes(text)
    with open(os.path.join(HERE, "tiny7.bwt"), "wb") as f:
        f.write(bwt)
    with open(os.path.join(HERE, "tiny7.sa"), "wb") as f:
        f.write(sa)
    reads = reads_of(data)
    for mode, k in MODES:
        res = (brute.smem if mode == "smem" else brute.mem)(bi, reads, k)
        with open(os.path.join(HERE, "tiny_%s_k%d.txt" % (mode, k)), "w") as f:
            f.write(brute.format_text(res, mode == "smem"))
    print("%d reads, %d bases of reference" % (len(reads), len(text)))


if __name__ == "__main__":
    main()
