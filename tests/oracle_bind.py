"""ctypes view of the oracle (oracle/_build/liboracle.so) for the tests.

The oracle is test infrastructure: the checker for the HIP path and the CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use
it.  GS_ORACLE_LIB names another build of the same library, e.g. the one
`make -C oracle sanitize` writes with AddressSanitizer and UBSan.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GS_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class OrOpts(C.Structure):
    _fields_ = [("k", C.c_int), ("smem", C.c_int), ("prune", C.c_int),
                ("max_occ", C.c_uint32), ("threads", C.c_int), ("use_table", C.c_int),
                ("smem_stop", C.c_int), ("ext_len", C.c_int), ("prune_sat", C.c_uint32),
                ("direct_min", C.c_uint32), ("miss_skip", C.c_int), ("pair_probe", C.c_int),
                ("tandem", C.c_uint32), ("tandem_span", C.c_uint32), ("tandem_len", C.c_uint32),
                ("tandem_pad", C.c_uint32), ("tandem_check", C.c_int)]


class OrSeed(C.Structure):
    _fields_ = [("qbeg", C.c_uint32), ("qend", C.c_uint32), ("l", C.c_uint32),
                ("u", C.c_uint32), ("n_pos", C.c_uint32), ("pad", C.c_uint32)]


class OrResult(C.Structure):
    _fields_ = [("n_reads", C.c_uint64),
                ("read_nseeds", C.POINTER(C.c_uint32)), ("read_seed_off", C.POINTER(C.c_uint64)),
                ("read_npos", C.POINTER(C.c_uint32)), ("read_pos_off", C.POINTER(C.c_uint64)),
                ("n_seeds", C.c_uint64), ("seeds", C.POINTER(OrSeed)),
                ("n_pos", C.c_uint64), ("pos", C.POINTER(C.c_uint32)),
                ("w_tasks", C.c_uint64), ("w_steps", C.c_uint64), ("w_blocks64", C.c_uint64),
                ("w_blocks96", C.c_uint64), ("w_lf", C.c_uint64), ("w_sa", C.c_uint64),
                ("w_steps_unpruned", C.c_uint64), ("n_short", C.c_uint64), ("w_ext", C.c_uint64),
                ("w_direct", C.c_uint64), ("w_text", C.c_uint64), ("w_pair", C.c_uint64),
                ("w_pair_ops", C.c_uint64), ("w_tandem", C.c_uint64), ("w_tandem_bad", C.c_uint64), ("w_stretches", C.c_uint64)]


_lib = None

# the device's tandem-jump settings (gs_internal.h): the shortest stretch
# that opens and the tandem table's longest length
TANDEM_SPAN, TANDEM_LEN = 24, 256
TANDEM_KEYS, TANDEM_ROWS = 5460, 237


def tandem_opts(tandem, check, span=TANDEM_SPAN, length=TANDEM_LEN):
    return tandem, span, length, 0, check


def tandem_row(index, key):
    """The oracle's tandem-table row of unit `key`: uint32 [237, 4]."""
    out = np.zeros((TANDEM_ROWS, 4), np.uint32)
    if lib().or_tandem_row(index.h, key, out.ctypes.data_as(C.POINTER(C.c_uint32))) != 0:
        raise RuntimeError("or_tandem_row failed")
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.or_index_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.or_index_free.argtypes = [C.c_void_p]
        for f in ("or_index_seq_len", "or_index_primary", "or_index_sa_intv"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_uint32
        L.or_index_L2.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
        L.or_occ.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.or_occ.restype = C.c_uint32
        L.or_step.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_step.restype = C.c_uint64
        L.or_lf.argtypes = [C.c_void_p, C.c_uint32]
        L.or_lf.restype = C.c_uint32
        L.or_locate.argtypes = [C.c_void_p, C.c_uint32]
        L.or_locate.restype = C.c_uint32
        L.or_sa_true.argtypes = [C.c_void_p, C.c_uint32]
        L.or_sa_true.restype = C.c_uint32
        L.or_code_fwd.argtypes = [C.c_uint8]
        L.or_code_fwd.restype = C.c_uint32
        L.or_code_rev.argtypes = [C.c_uint8]
        L.or_code_rev.restype = C.c_uint32
        L.or_seed_reads.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(OrOpts),
                                    C.POINTER(OrResult)]
        L.or_result_free.argtypes = [C.POINTER(OrResult)]
        L.or_tandem_row.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


class Index:
    def __init__(self, prefix):
        self.h = C.c_void_p()
        rc = lib().or_index_load(prefix.encode(), C.byref(self.h))
        if rc != 0:
            raise OSError(rc, "or_index_load failed for %s" % prefix)

    def close(self):
        if self.h:
            lib().or_index_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def seq_len(self):
        return lib().or_index_seq_len(self.h)

    @property
    def primary(self):
        return lib().or_index_primary(self.h)

    @property
    def sa_intv(self):
        return lib().or_index_sa_intv(self.h)

    def L2(self):
        a = (C.c_uint32 * 5)()
        lib().or_index_L2(self.h, a)
        return list(a)

    def occ(self, row, c):
        return lib().or_occ(self.h, row & 0xFFFFFFFF, c)

    def step(self, l, u, c):
        v = lib().or_step(self.h, l & 0xFFFFFFFF, u & 0xFFFFFFFF, c)
        return v & 0xFFFFFFFF, v >> 32

    def lf(self, row):
        return lib().or_lf(self.h, row)

    def locate(self, row):
        return lib().or_locate(self.h, row)

    def sa_true(self, row):
        return lib().or_sa_true(self.h, row)


def pack_reads(reads):
    """reads: list of bytes -> (buffer, offsets u64, lens u32)"""
    lens = np.array([len(r) for r in reads], dtype=np.uint32)
    offs = np.zeros(len(reads), dtype=np.uint64)
    if len(reads):
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    buf = b"".join(reads)
    return buf, offs, lens


def seed(index, reads, k, smem, prune=0, max_occ=0, threads=4, use_table=0, smem_stop=0, ext_len=0, prune_sat=0,
         direct_min=0, miss_skip=0, pair_probe=0, tandem=0, tandem_check=0, **tkw):
    """Returns (per-read list of (qbeg, qend, l, u, [positions])), counters dict)."""
    buf, offs, lens = pack_reads(reads)
    o = OrOpts(k, 1 if smem else 0, prune, max_occ, threads, use_table, smem_stop, ext_len, prune_sat, direct_min,
               miss_skip, pair_probe, *tandem_opts(tandem, tandem_check, **tkw))
    res = OrResult()
    rc = lib().or_seed_reads(index.h, buf if buf else b"\0",
                             offs.ctypes.data_as(C.POINTER(C.c_uint64)),
                             lens.ctypes.data_as(C.POINTER(C.c_uint32)), len(reads),
                             C.byref(o), C.byref(res))
    if rc != 0:
        raise RuntimeError("or_seed_reads failed: %d" % rc)
    try:
        out = []
        for r in range(len(reads)):
            s0 = res.read_seed_off[r]
            p = res.read_pos_off[r]
            lst = []
            for j in range(res.read_nseeds[r]):
                sd = res.seeds[s0 + j]
                pos = [res.pos[p + q] for q in range(sd.n_pos)]
                p += sd.n_pos
                lst.append((sd.qbeg, sd.qend, sd.l, sd.u, pos))
            out.append(lst)
        counters = {f: getattr(res, f) for f in ("w_tasks", "w_steps", "w_blocks64", "w_blocks96",
                                                 "w_lf", "w_sa", "w_steps_unpruned", "n_short", "w_ext",
                                                 "w_direct", "w_text", "w_pair", "w_pair_ops", "w_tandem",
                                                 "w_tandem_bad", "w_stretches")}
        return out, counters
    finally:
        lib().or_result_free(C.byref(res))


def seed_arrays(index, buf, offs, lens, k, smem, prune=1, max_occ=0, threads=0, use_table=1, smem_stop=0,
                ext_len=0, prune_sat=0, direct_min=0, miss_skip=0, pair_probe=0, tandem=0, tandem_check=0, **tkw):
    """Flat numpy copies of the oracle's result: (seeds[n,6] u32, pos u32,
    read_seed_off u64[n+1], read_pos_off u64[n+1]) plus the work counters."""
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    o = OrOpts(k, 1 if smem else 0, prune, max_occ, threads, use_table, smem_stop, ext_len, prune_sat, direct_min,
               miss_skip, pair_probe, *tandem_opts(tandem, tandem_check, **tkw))
    res = OrResult()
    rc = lib().or_seed_reads(index.h, C.cast(buf.ctypes.data, C.c_char_p),
                             offs.ctypes.data_as(C.POINTER(C.c_uint64)),
                             lens.ctypes.data_as(C.POINTER(C.c_uint32)), len(lens), C.byref(o),
                             C.byref(res))
    if rc != 0:
        raise RuntimeError("or_seed_reads failed: %d" % rc)
    try:
        n = len(lens)
        ns, npos = res.n_seeds, res.n_pos
        seeds = (np.ctypeslib.as_array(C.cast(res.seeds, C.POINTER(C.c_uint32)), (ns * 6,)).copy()
                 .reshape(ns, 6) if ns else np.zeros((0, 6), np.uint32))
        pos = np.ctypeslib.as_array(res.pos, (npos,)).copy() if npos else np.zeros(0, np.uint32)
        so = np.ctypeslib.as_array(res.read_seed_off, (n + 1,)).copy()
        po = np.ctypeslib.as_array(res.read_pos_off, (n + 1,)).copy()
        counters = {f: getattr(res, f) for f in ("w_tasks", "w_steps", "w_blocks64", "w_blocks96",
                                                 "w_lf", "w_sa", "w_steps_unpruned", "n_short", "w_ext",
                                                 "w_direct", "w_text", "w_pair", "w_pair_ops", "w_tandem",
                                                 "w_tandem_bad", "w_stretches")}
        return seeds, pos, so, po, counters
    finally:
        lib().or_result_free(C.byref(res))
