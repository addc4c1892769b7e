"""The seed_gen command line and the -o text writer.

CPU: the library's writer (gs_format_text) against the oracle's fprintf
restatement of seed_gen.cu:3625-3648, byte for byte, and the command line's
error exits (seed_gen.cu:2685-2736).  GPU: seed_gen -o against oracle_seed -o
on the configuration-1 workload (1 Mbp reference, 10,000 x 100 bp reads)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEED_GEN = os.path.join(ROOT, "gpuseed_amd", "bin", "seed_gen")
ORACLE_SEED = os.path.join(ROOT, "oracle", "_build", "oracle_seed")
GS_SYNTH = os.path.join(ROOT, "gpuseed_amd", "bin", "gs_synth")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    """Configuration 1 as files: index prefix and reads FASTA."""
    import gpuseed_amd as g
    d = str(tmp_path_factory.mktemp("c1cli"))
    rp, qp = g.preset(1)
    ref = g.synth_reference(rp)
    buf, offs, lens, _, _ = g.synth_reads(ref, qp)
    prefix = os.path.join(d, "c1")
    g.build_index_host(ref, 7, prefix)
    fa = os.path.join(d, "reads.fa")
    with open(fa, "wb") as f:
        for i, (o, l) in enumerate(zip(offs, lens)):
            f.write(b">r%d\n" % i)
            f.write(buf[int(o):int(o) + int(l)].tobytes() + b"\n")
    return dict(dir=d, prefix=prefix, fa=fa, arrays=(buf, offs, lens))


def oracle_text(prefix, arrays, k, smem, max_occ, first, path):
    L = ob.lib()
    L.or_format_text.argtypes = [C.POINTER(ob.OrResult), C.c_int, C.c_uint64, C.c_void_p]
    ox = ob.Index(prefix)
    buf, offs, lens = (np.ascontiguousarray(a) for a in arrays)
    o = ob.OrOpts(k, int(smem), 1, max_occ, 8, 0)
    res = ob.OrResult()
    assert L.or_seed_reads(ox.h, C.cast(buf.ctypes.data, C.c_char_p),
                           offs.ctypes.data_as(C.POINTER(C.c_uint64)),
                           lens.ctypes.data_as(C.POINTER(C.c_uint32)), len(lens), C.byref(o),
                           C.byref(res)) == 0
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(path.encode(), b"w")
    L.or_format_text(C.byref(res), int(smem), first, fp)
    libc.fclose(fp)
    L.or_result_free(C.byref(res))
    ox.close()


def as_gs_result(seeds, pos, so, po):
    """A gs_result viewing the oracle's flat arrays (kept alive by the caller)."""
    import gpuseed_amd as g
    n = len(so) - 1
    nseeds = np.diff(so).astype(np.uint32)
    npos = np.diff(po).astype(np.uint32)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    keep = [nseeds, npos, so, po, seeds, pos]
    r = g.Result()
    r.n_reads, r.n_seeds, r.n_pos = n, len(seeds), len(pos)
    r.read_nseeds = nseeds.ctypes.data_as(C.POINTER(C.c_uint32))
    r.read_npos = npos.ctypes.data_as(C.POINTER(C.c_uint32))
    r.read_seed_off = so.ctypes.data_as(C.POINTER(C.c_uint64))
    r.read_pos_off = po.ctypes.data_as(C.POINTER(C.c_uint64))
    r.seeds = seeds.ctypes.data_as(C.POINTER(g.SeedRec))
    r.pos = pos.ctypes.data_as(C.POINTER(C.c_uint32))
    return r, keep


@pytest.mark.parametrize("smem,k,max_occ,first,threads", [
    (False, 20, 0, 1, None), (True, 19, 0, 1, None), (True, 19, 4, 1, "3"),
    (False, 20, 2, 2147483000, "1"), (True, 20, 0, 1000001, "16")])
def test_writer_matches_oracle_text(built, c1, tmp_path, smem, k, max_occ, first, threads):
    """max_occ leaves seeds without positions (they print nothing and do not
    break a run of equal pairs); the read number is an int, so a large first
    number wraps like the reference's %d."""
    import gpuseed_amd as g
    ox = ob.Index(c1["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *c1["arrays"], k=k, smem=smem, max_occ=max_occ)
    ox.close()
    want = os.path.join(str(tmp_path), "want.txt")
    oracle_text(c1["prefix"], c1["arrays"], k, smem, max_occ, first, want)
    r, keep = as_gs_result(seeds, pos, so, po)
    got = os.path.join(str(tmp_path), "got.txt")
    old = os.environ.get("GS_FORMAT_THREADS")
    if threads:
        os.environ["GS_FORMAT_THREADS"] = threads
    try:
        fd = os.open(got, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
        try:
            assert g.lib().gs_format_text(C.byref(r), int(smem), first, fd) == 0
        finally:
            os.close(fd)
    finally:
        if threads:
            if old is None:
                del os.environ["GS_FORMAT_THREADS"]
            else:
                os.environ["GS_FORMAT_THREADS"] = old
    del keep
    a, b = open(got, "rb").read(), open(want, "rb").read()
    assert len(a) == len(b) and a == b


def test_writer_splits_large_outputs_in_order(built, c1, tmp_path):
    """More text than one 8 MiB slice: the same reads repeated until the
    writer cuts several slices, formatted on several threads."""
    import gpuseed_amd as g
    ox = ob.Index(c1["prefix"])
    seeds, pos, so, po, _ = ob.seed_arrays(ox, *c1["arrays"], k=20, smem=False)
    ox.close()
    rep = 6
    n = len(so) - 1
    seeds4 = np.concatenate([seeds] * rep)
    pos4 = np.concatenate([pos] * rep)
    so4 = np.concatenate([so[:-1] + i * so[-1] for i in range(rep)] + [[so[-1] * rep]]).astype(np.uint64)
    po4 = np.concatenate([po[:-1] + i * po[-1] for i in range(rep)] + [[po[-1] * rep]]).astype(np.uint64)
    r, keep = as_gs_result(seeds4, pos4, so4, po4)
    got = os.path.join(str(tmp_path), "big.txt")
    fd = os.open(got, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    try:
        assert g.lib().gs_format_text(C.byref(r), 0, 1, fd) == 0
    finally:
        os.close(fd)
    del keep
    one = os.path.join(str(tmp_path), "one.txt")
    parts = []
    for i in range(rep):
        oracle_text(c1["prefix"], c1["arrays"], 20, False, 0, 1 + i * n, one)
        parts.append(open(one, "rb").read())
    want = b"".join(parts)
    data = open(got, "rb").read()
    assert len(want) > 8 << 20
    assert data == want


def run(args, **kw):
    return subprocess.run([SEED_GEN] + args, capture_output=True, timeout=300, **kw)


def test_cli_index_errors_exit_like_the_reference(built, c1, tmp_path):
    """Missing .bwt / .sa and header disagreement exit 1 with the reference's
    messages (seed_gen.cu:2692-2723), before any device work."""
    d = str(tmp_path)
    p = run([os.path.join(d, "nope"), c1["fa"]])
    assert p.returncode == 1 and b"Unable to open .bwt file." in p.stderr
    only_bwt = os.path.join(d, "only")
    open(only_bwt + ".bwt", "wb").write(open(c1["prefix"] + ".bwt", "rb").read())
    p = run([only_bwt, c1["fa"]])
    assert p.returncode == 1 and b"Unable to open .sa file." in p.stderr
    for field, msg in ((0, b"primary is not the same"), (6, b"seq_len is not the same")):
        bad = os.path.join(d, "bad%d" % field)
        open(bad + ".bwt", "wb").write(open(c1["prefix"] + ".bwt", "rb").read())
        sa = bytearray(open(c1["prefix"] + ".sa", "rb").read())
        v = int.from_bytes(sa[4 * field:4 * field + 4], "little") ^ 1
        sa[4 * field:4 * field + 4] = v.to_bytes(4, "little")
        open(bad + ".sa", "wb").write(bytes(sa))
        p = run([bad, c1["fa"]])
        assert p.returncode == 1 and b"SA-BWT inconsistency: " + msg in p.stderr


def test_cli_argument_errors(built, c1):
    assert run([]).returncode == 1
    p = run(["-k", "13", c1["prefix"], c1["fa"]])
    assert p.returncode == 1 and b"at least 14" in p.stderr
    p = run([c1["prefix"], c1["fa"] + ".missing"])
    assert p.returncode == 1 and b"cannot open" in p.stderr
    p = run(["--kernel", "7", c1["prefix"], c1["fa"]])
    assert p.returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [
    ["-k", "20"],                                   # MEM, k=20 (the 1 Mbp preset)
    ["-s", "-k", "19"],
    ["-s", "-k", "19", "--kernel", "0"],
    ["-k", "20", "--device-bases", "300000"],       # one reference batch per device batch
    ["-s", "-k", "20", "--max-occ", "5", "--device-bases", "2500000"],
    ["-s", "-k", "19", "--kmer-len", "15", "--autotune"],   # extended table, tuned launch
    ["-k", "20", "--kmer-len", "14", "--device-bases", "400000"],
    ["-s", "-k", "19", "--kmer-len", "16", "--direct"],     # presence bits and direct runs
    ["-k", "20", "--direct", "--kernel", "2"],
])
def test_cli_text_matches_oracle(built, c1, tmp_path, flags):
    say("seed_gen", " ".join(flags))
    got = os.path.join(str(tmp_path), "got.txt")
    with open(got, "wb") as f:
        p = subprocess.run([SEED_GEN, "-o"] + flags + [c1["prefix"], c1["fa"]], stdout=f,
                           stderr=subprocess.PIPE, timeout=300)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    oflags, i = [], 0
    while i < len(flags):           # the oracle takes -k, -s and -m; device options are dropped
        if flags[i] in ("--device-bases", "--kernel", "--kmer-len"):
            i += 2
        elif flags[i] in ("--autotune", "--direct"):
            i += 1
        elif flags[i] == "--max-occ":
            oflags += ["-m", flags[i + 1]]
            i += 2
        else:
            oflags.append(flags[i])
            i += 1
    want = os.path.join(str(tmp_path), "want.txt")
    with open(want, "wb") as f:
        q = subprocess.run([ORACLE_SEED, "-o", "-t", "8"] + oflags + [c1["prefix"], c1["fa"]],
                           stdout=f, stderr=subprocess.PIPE, timeout=600)
    assert q.returncode == 0, q.stderr.decode()[-2000:]
    a, b = open(got, "rb").read(), open(want, "rb").read()
    assert len(a) > 100000 and a == b


@pytest.mark.gpu
def test_cli_reference_capacity_stops_after_earlier_batches(built, tmp_path):
    """A repeat-only reference makes a batch overflow the reference's fixed
    output buffer (seed_gen.cu:3553-3558): with --reference-capacity the run
    prints the batches before it and exits 1."""
    import gpuseed_amd as g
    d = str(tmp_path)
    rng = np.random.default_rng(4)
    unit = rng.integers(0, 4, 200, dtype=np.uint8)
    ref = np.concatenate([rng.integers(0, 4, 50000, dtype=np.uint8), np.tile(unit, 2000)])
    prefix = os.path.join(d, "rep")
    g.build_index_host(ref, 7, prefix)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    uniq = acgt[ref[:50000]].tobytes()
    rpt = acgt[ref[60000:]].tobytes()
    fa = os.path.join(d, "r.fa")
    with open(fa, "wb") as f:
        for i in range(10000):          # batch 1: unique reads, 1,000,000 bases
            f.write(uniq[(i * 37) % 49000:(i * 37) % 49000 + 100] + b"\n")
        for i in range(200):            # batch 2: repeat reads, 2000 positions per seed
            f.write(rpt[i * 50:i * 50 + 100] + b"\n")
    got = os.path.join(d, "got.txt")
    with open(got, "wb") as f:
        p = subprocess.run([SEED_GEN, "-o", "-s", "-k", "19", "--reference-capacity", prefix, fa],
                           stdout=f, stderr=subprocess.PIPE, timeout=300)
    assert p.returncode == 1 and b"is more than allocated size" in p.stderr
    text = open(got, "rb").read()
    assert text.count(b"seeds in read no.") == 10000
    with open(got, "wb") as f:
        p = subprocess.run([SEED_GEN, "-o", "-s", "-k", "19", prefix, fa], stdout=f,
                           stderr=subprocess.PIPE, timeout=300)
    assert p.returncode == 0
    assert open(got, "rb").read().count(b"seeds in read no.") == 10200
