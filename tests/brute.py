"""Brute-force restatement of FM-index seeding for tiny texts.

Everything here is computed from first principles, independently of the
oracle's FM-index machinery:
  * suffix array = Python's sort of all suffixes (the `$` suffix sorts first);
  * an SA interval = the ranks of the occurrence positions found by substring
    search;
  * extension = substring search of the longer pattern;
  * SMEM survivors = for each distinct left end, the largest right end;
  * MEM rows = occurrences of P_i that do not extend to the next same-qbeg
    entry, as sets.
The reference behaviour that no definition fixes (the primary+1 step of the
suffix-array walk, seed_gen.cu:455-457; the batch-wide MEM look-ahead,
seed_gen.cu:637/645) is modelled explicitly and named where it is used.
"""

BASES = "ACGT"

# character meaning -> (forward code, reverse-complement code); the reference
# indexes a lookup table by (byte & 7) (seed_gen.cu:2540, 2562), so every byte
# aliases one of the eight rows below.
_ROW_OF_LOW3 = {
    0: (5, 5),  # NUL @ H P X ` h p x : no base
    1: (0, 3),  # A a Q q ...
    2: (0, 0),  # B R b r ...
    3: (1, 2),  # C K S c k s
    4: (3, 0),  # D L T d l t
    5: (0, 0),  # E M U \r ...
    6: (4, 4),  # F N V n ...
    7: (2, 1),  # G O W g o w
}


def code_fwd(byte):
    return _ROW_OF_LOW3[byte & 7][0]


def code_rev(byte):
    return _ROW_OF_LOW3[byte & 7][1]


class BruteIndex:
    def __init__(self, text, sa_intv=1):
        assert set(text) <= set(BASES)
        self.T = text
        n = len(text)
        self.n = n
        self.sa = sorted(range(n + 1), key=lambda i: text[i:])
        self.rank = [0] * (n + 1)
        for r, i in enumerate(self.sa):
            self.rank[i] = r
        self.primary = self.rank[0]
        self.bwt = [text[i - 1] if i > 0 else "$" for i in self.sa]
        self.sa_intv = sa_intv
        self.L2 = [0] * 5
        for c in text:
            self.L2[BASES.index(c) + 1] += 1
        for c in range(2, 5):
            self.L2[c] += self.L2[c - 1]

    def positions(self, pat):
        out, p = [], self.T.find(pat)
        while p >= 0:
            out.append(p)
            p = self.T.find(pat, p + 1)
        return out

    def rows(self, pat):
        return sorted(self.rank[p] for p in self.positions(pat))

    def interval(self, pat):
        r = self.rows(pat)
        if not r:
            return None
        assert r == list(range(r[0], r[-1] + 1)), "SA interval not contiguous"
        return r[0], r[-1]

    def occ(self, row, c):
        """#c in BWT rows [0..row] ($ excluded); row == -1 -> 0"""
        if row < 0:
            return 0
        ch = BASES[c]
        return sum(1 for r in range(row + 1) if self.bwt[r] == ch)

    def lf_true(self, row):
        i = self.sa[row]
        return self.rank[i - 1] if i > 0 else 0

    def locate(self, row):
        """Position the reference's walk reports for `row` (seed_gen.cu:668-684).
        The walk's LF returns 0 for rows primary and primary+1 (455-457); for
        every other row it is the true LF.  Row 0 holds -1 (sa[0] = -1, 2711)."""
        if row > self.n:  # only from the batch-wide MEM row skip; no position
            return 0xFFFFFFFF
        itr = 0
        while row % self.sa_intv:
            itr += 1
            row = 0 if row in (self.primary, self.primary + 1) else self.lf_true(row)
        base = 0xFFFFFFFF if row == 0 else self.sa[row]
        return (base + itr) & 0xFFFFFFFF


PRECALC = 13


def strand_candidates(bi, vcodes, k):
    """Virtual-strand candidates, decreasing end: list of (b, e, pattern).
    A code > 3 inside the 13-base window drops the task; after the window it
    is skipped (the interval is unchanged) but still consumed."""
    L = len(vcodes)
    out = []
    if L < k:
        return out
    for e in range(L, k - 1, -1):
        win = vcodes[e - PRECALC:e]
        if any(c > 3 for c in win):
            continue
        cur = "".join(BASES[c] for c in win)
        if bi.T.find(cur) < 0:
            continue
        b = e - PRECALC
        for x in range(e - PRECALC - 1, -1, -1):
            c = vcodes[x]
            if c > 3:
                b = x
                continue
            nxt = BASES[c] + cur
            if bi.T.find(nxt) < 0:
                break
            cur = nxt
            b = x
        if e - b >= k:
            out.append((b, e, cur))
    return out


def read_lists(bi, read, k):
    """(F, R): forward entries by increasing qend, reverse by increasing qbeg;
    entries are dicts with x, y (strand bit 31 on reverse), pattern, rows."""
    L = len(read)
    fc = [code_fwd(b) for b in read]
    rc = [code_rev(b) for b in reversed(read)]
    F = []
    for b, e, pat in reversed(strand_candidates(bi, fc, k)):
        F.append(dict(x=b, y=e, qb=b, pat=pat, rows=bi.rows(pat)))
    R = []
    for b, e, pat in strand_candidates(bi, rc, k):
        qb, qe = L - e, L - b
        R.append(dict(x=qb | 0x80000000, y=qe | 0x80000000, qb=qb, pat=pat, rows=bi.rows(pat)))
    return F, R


def _finish(bi, entries, max_occ):
    """entries: (qbeg, is_rev, x, y, rows_all, rows_out) -> merged seed list"""
    entries.sort(key=lambda t: (t[0], t[1]))  # stable merge by qbeg, forward first
    out = []
    for qb, is_rev, x, y, rows_all, rows_out in entries:
        l, u = rows_all[0], rows_all[-1]
        size = u - l + 1
        if max_occ and size > max_occ:
            rows_out = []
        out.append((x, y, l, u, [bi.locate(r) for r in rows_out]))
    return out


def smem(bi, reads, k, max_occ=0):
    res = []
    for read in reads:
        F, R = read_lists(bi, read, k)
        ent = []
        best = {}
        for i, f in enumerate(F):  # largest qend per qbeg
            best[f["qb"]] = i
        for i in best.values():
            f = F[i]
            ent.append((f["qb"], 0, f["x"], f["y"], f["rows"], f["rows"]))
        best = {}
        for i, r in enumerate(R):  # smallest qbeg per qend
            best.setdefault(r["y"], i)
        for i in best.values():
            r = R[i]
            ent.append((r["qb"], 1, r["x"], r["y"], r["rows"], r["rows"]))
        res.append(_finish(bi, ent, max_occ))
    return res


def mem(bi, reads, k, max_occ=0):
    lists = [read_lists(bi, read, k) for read in reads]
    # the batch's entry array (read order: F then R), seed_gen.cu:3400-3401
    arr = []
    for F, R in lists:
        arr.extend(F)
        arr.extend(R)

    def killed(F, R, is_rev, i):  # filter_seed_intervals_gpu_mem, 1103-1139
        if not is_rev:
            return i + 1 < len(F) and F[i + 1]["x"] == F[i]["x"] and len(F[i + 1]["rows"]) == len(F[i]["rows"])
        return i > 0 and R[i - 1]["y"] == R[i]["y"] and len(R[i - 1]["rows"]) == len(R[i]["rows"])

    where = []
    for ri, (F, R) in enumerate(lists):
        where += [(ri, 0, i) for i in range(len(F))] + [(ri, 1, i) for i in range(len(R))]

    res = []
    g = 0
    for ri, (F, R) in enumerate(lists):
        ent = []
        for is_rev, lst in ((0, F), (1, R)):
            for i, e in enumerate(lst):
                gi = g + (len(F) + i if is_rev else i)
                rows = e["rows"]
                if not is_rev and i + 1 < len(F) and F[i + 1]["x"] == e["x"]:
                    out = sorted(set(rows) - set(F[i + 1]["rows"]))
                elif is_rev and i > 0 and R[i - 1]["y"] == e["y"]:
                    out = sorted(set(rows) - set(R[i - 1]["rows"]))
                else:
                    # no in-read neighbour: the reference's row expansion looks
                    # at neighbouring entries of the whole batch (637 / 645)
                    skip = None
                    step = 1 if not is_rev else -1
                    key = "x" if not is_rev else "y"
                    j = gi + step
                    while 0 <= j < len(arr) and arr[j][key] == e[key]:
                        rj, rv, ij = where[j]
                        Fj, Rj = lists[rj]
                        if not killed(Fj, Rj, rv, ij):
                            skip = (arr[j]["rows"][0], arr[j]["rows"][-1])
                            break
                        j += step
                    out = []
                    r = rows[0]
                    while len(out) < len(rows):
                        if skip is None or r < skip[0] or r > skip[1]:
                            out.append(r)
                        r += 1
                if out:
                    ent.append((e["qb"], is_rev, e["x"], e["y"], rows, out))
        g += len(F) + len(R)
        res.append(_finish(bi, ent, max_occ))
    return res


def format_text(results, smem_mode, first_read_no=1):
    """The -o text (seed_gen.cu:3625-3648) of a list of per-read seed lists."""
    head = ("/*===================================%s seeds in read no. %d (Read[begin, End] -> "
            "starting position(s) on the reference)===================================*/\n")
    out = []
    for r, seeds in enumerate(results):
        out.append(head % ("SMEM" if smem_mode else "MEM", first_read_no + r))
        pb = pe = None
        for x, y, l, u, pos in seeds:
            for p in pos:
                sign = "-" if y >> 31 else "+"
                if (x, y) == (pb, pe):
                    out.append(", %s%d" % (sign, p))
                else:
                    out.append("\nRead[%d, %d] -> %s%d" % (x & 0x7FFFFFFF, y & 0x7FFFFFFF, sign, p))
                pb, pe = x, y
        out.append("\n")
    return "".join(out)
