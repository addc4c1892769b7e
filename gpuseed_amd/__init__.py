"""Python view of libgpuseed.so, the MI355X FM-index seeder.

The product is the C ABI declared in include/gpuseed.h (HIP kernels for
gfx950 behind it) and the seed_gen command line built on it.  This module
binds that ABI with ctypes for tests, benchmarks and notebooks.  There is no
fallback: if the shared library is missing, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpuseed.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "gpuseed.h")

GS_OK, GS_EINVAL, GS_ENOMEM, GS_EHIP, GS_EIO, GS_EFORMAT, GS_ENODEV, GS_ERANGE = (
    0, -1, -2, -3, -4, -5, -6, -7)


class GsError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__("%s failed with code %d" % (fn, rc))
        self.rc = rc


class RooflineResult(C.Structure):
    _fields_ = [("seconds", C.c_double), ("lines_per_s", C.c_double), ("bytes_per_s", C.c_double)]


class SynthRefParams(C.Structure):
    _fields_ = [("length", C.c_uint64), ("seed", C.c_uint64), ("gc", C.c_double),
                ("n_families", C.c_uint32), ("family_len_min", C.c_uint32),
                ("family_len_max", C.c_uint32), ("copies_per_family", C.c_uint32),
                ("interspersed_fraction", C.c_double), ("family_div_min", C.c_double),
                ("family_div_max", C.c_double), ("tandem_fraction", C.c_double),
                ("tandem_unit_min", C.c_uint32), ("tandem_unit_max", C.c_uint32),
                ("tandem_copies_min", C.c_uint32), ("tandem_copies_max", C.c_uint32),
                ("segdup_fraction", C.c_double), ("segdup_len_min", C.c_uint32),
                ("segdup_len_max", C.c_uint32), ("segdup_div_min", C.c_double),
                ("segdup_div_max", C.c_double)]


class ReadClass(C.Structure):
    _fields_ = [("weight", C.c_double), ("len_min", C.c_uint32), ("len_max", C.c_uint32),
                ("sub_rate", C.c_double), ("indel_rate", C.c_double)]


class SynthReadParams(C.Structure):
    _fields_ = [("n_reads", C.c_uint64), ("seed", C.c_uint64), ("n_classes", C.c_uint32),
                ("classes", ReadClass * 4), ("n_read_fraction", C.c_double),
                ("both_strands", C.c_int)]


class IndexInfo(C.Structure):
    _fields_ = [("seq_len", C.c_uint32), ("primary", C.c_uint32), ("sa_intv", C.c_uint32),
                ("L2", C.c_uint32 * 5), ("n_blocks", C.c_uint64), ("device_bytes", C.c_uint64),
                ("t_read", C.c_double), ("t_upload", C.c_double), ("t_repack", C.c_double),
                ("t_table", C.c_double), ("t_sa", C.c_double), ("t_write", C.c_double),
                ("sa_rounds", C.c_int), ("kmer_len", C.c_int), ("t_ext", C.c_double),
                ("direct", C.c_int), ("t_direct", C.c_double), ("pairs", C.c_int), ("t_pairs", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("k", C.c_int), ("smem", C.c_int), ("max_occ", C.c_uint32),
                ("batch_bases", C.c_uint32), ("kernel", C.c_int), ("seg_tasks", C.c_uint32),
                ("wg", C.c_uint16), ("walk_rows", C.c_uint16), ("split_tasks", C.c_uint16),
                ("split_work", C.c_uint16), ("direct_min", C.c_uint16), ("tandem", C.c_uint16)]


class ServerOpts(C.Structure):
    _fields_ = [("opts", Opts), ("budget_bytes", C.c_uint64), ("round_bases", C.c_uint64),
                ("exclusive", C.c_int), ("workers", C.c_int), ("round_growth", C.c_int),
                ("reserved", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("n_reads", C.c_uint64), ("n_seeds", C.c_uint64), ("n_pos", C.c_uint64),
                ("n_rounds", C.c_uint32), ("pad", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("n_requests", C.c_uint64), ("n_rounds", C.c_uint64), ("n_reads", C.c_uint64),
                ("n_bases", C.c_uint64), ("busy_seconds", C.c_double), ("device_seconds", C.c_double),
                ("t_load", C.c_double), ("t_run", C.c_double), ("t_fetch", C.c_double), ("t_split", C.c_double),
                ("peak_round_bytes", C.c_uint64), ("elapsed", C.c_double), ("t_plan", C.c_double),
                ("t_idle", C.c_double), ("t_seed", C.c_double), ("t_post", C.c_double), ("t_locate", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("opts", Opts), ("seconds", C.c_double), ("default_seconds", C.c_double),
                ("n_tried", C.c_int), ("pad", C.c_int)]


class SeedRec(C.Structure):
    _fields_ = [("qbeg", C.c_uint32), ("qend", C.c_uint32), ("l", C.c_uint32), ("u", C.c_uint32),
                ("n_pos", C.c_uint32), ("pad", C.c_uint32)]


SEED_DTYPE = np.dtype([("qbeg", "<u4"), ("qend", "<u4"), ("l", "<u4"), ("u", "<u4"),
                       ("n_pos", "<u4"), ("pad", "<u4")])


class ResultBuffers(C.Structure):
    _fields_ = [("read_nseeds", C.POINTER(C.c_uint32)), ("read_npos", C.POINTER(C.c_uint32)),
                ("read_seed_off", C.POINTER(C.c_uint64)), ("read_pos_off", C.POINTER(C.c_uint64)),
                ("seeds", C.POINTER(SeedRec)), ("pos", C.POINTER(C.c_uint32)),
                ("seeds_cap", C.c_uint64), ("pos_cap", C.c_uint64),
                ("n_seeds", C.c_uint64), ("n_pos", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("n_reads", C.c_uint64), ("n_seeds", C.c_uint64), ("n_pos", C.c_uint64),
                ("read_nseeds", C.POINTER(C.c_uint32)), ("read_npos", C.POINTER(C.c_uint32)),
                ("read_seed_off", C.POINTER(C.c_uint64)), ("read_pos_off", C.POINTER(C.c_uint64)),
                ("seeds", C.POINTER(SeedRec)), ("pos", C.POINTER(C.c_uint32)),
                ("t_h2d", C.c_double), ("t_seed", C.c_double), ("t_post", C.c_double),
                ("t_locate", C.c_double), ("t_d2h", C.c_double)]


_lib = None


def lib():
    """The loaded libgpuseed.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgpuseed.so not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        u8p, u32p, u64p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        L.gs_roofline_random.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.POINTER(RooflineResult)]
        L.gs_stream_copy_bytes.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_double)]
        L.gs_synth_reference.argtypes = [C.POINTER(SynthRefParams), u8p, C.c_int]
        L.gs_synth_read_lengths.argtypes = [C.POINTER(SynthReadParams), u32p]
        L.gs_synth_reads.argtypes = [u8p, C.c_uint64, C.POINTER(SynthReadParams), u64p, C.c_char_p,
                                     u32p, u8p, C.c_int]
        L.gs_synth_preset.argtypes = [C.c_int, C.POINTER(SynthRefParams),
                                      C.POINTER(SynthReadParams)]
        L.gs_build_index_host.argtypes = [u8p, C.c_uint64, C.c_uint32, C.c_char_p]
        L.gs_index_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        L.gs_index_close.argtypes = [C.c_void_p]
        L.gs_index_build.argtypes = [u8p, C.c_uint64, C.c_uint32, C.c_int, C.c_char_p,
                                     C.POINTER(C.c_void_p)]
        L.gs_index_get_info.argtypes = [C.c_void_p, C.POINTER(IndexInfo)]
        L.gs_index_read_occ.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, u32p]
        L.gs_index_read_kmer.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, u32p]
        L.gs_index_read_kext.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, u32p]
        L.gs_index_read_tandem.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, u32p]
        L.gs_index_set_kmer_len.argtypes = [C.c_void_p, C.c_int]
        L.gs_index_set_direct.argtypes = [C.c_void_p, C.c_int]
        L.gs_index_set_pairs.argtypes = [C.c_void_p, C.c_int]
        L.gs_index_read_pairs.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint8)]
        L.gs_index_locate.argtypes = [C.c_void_p, u32p, C.c_uint64, u32p]
        L.gs_index_read_direct.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, u32p, u32p,
                                           C.POINTER(C.c_uint8)]
        L.gs_seed.argtypes = [C.c_void_p, C.c_char_p, u64p, u32p, C.c_uint64, C.POINTER(Opts),
                              C.POINTER(Result)]
        L.gs_result_free.argtypes = [C.POINTER(Result)]
        L.gs_result_free.restype = None
        L.gs_format_text.argtypes = [C.POINTER(Result), C.c_int, C.c_uint64, C.c_int]
        L.gs_index_convert.argtypes = [C.c_char_p, C.c_char_p]
        L.gs_bwa_ref_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.gs_bwa_ref_free.argtypes = [C.c_void_p]
        L.gs_bwa_ref_free.restype = None
        L.gs_bwa_ref_length.argtypes = [C.c_void_p]
        L.gs_bwa_ref_length.restype = C.c_uint64
        L.gs_bwa_ref_codes.argtypes = [C.c_void_p]
        L.gs_bwa_ref_codes.restype = u8p
        L.gs_bwa_ref_n_contigs.argtypes = [C.c_void_p]
        L.gs_bwa_ref_n_holes.argtypes = [C.c_void_p]
        L.gs_bwa_ref_contig.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), u64p, u64p]
        L.gs_bwa_ref_hole.argtypes = [C.c_void_p, C.c_int, u64p, u64p, C.c_char_p]
        L.gs_bwa_ref_locate.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), u64p]
        L.gs_set_launch_chunk.argtypes = [C.c_uint64]
        L.gs_seed_noalloc.argtypes = [C.c_void_p, C.c_char_p, u64p, u32p, C.c_uint64, C.POINTER(Opts),
                                      C.POINTER(ResultBuffers)]
        _lib = L
    return _lib


def _check(fn, rc):
    if rc != GS_OK:
        raise GsError(fn, rc)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def preset(config):
    rp, qp = SynthRefParams(), SynthReadParams()
    _check("gs_synth_preset", lib().gs_synth_preset(config, C.byref(rp), C.byref(qp)))
    return rp, qp


def synth_reference(params, threads=0):
    codes = np.empty(params.length, dtype=np.uint8)
    _check("gs_synth_reference", lib().gs_synth_reference(C.byref(params), _ptr(codes, C.c_uint8),
                                                          threads))
    return codes


def synth_reads(ref_codes, params, threads=0):
    """-> (buffer bytes as np.uint8, offsets u64, lens u32, origin u32, strand u8)"""
    n = params.n_reads
    lens = np.empty(n, dtype=np.uint32)
    _check("gs_synth_read_lengths", lib().gs_synth_read_lengths(C.byref(params), _ptr(lens, C.c_uint32)))
    offs = np.zeros(n, dtype=np.uint64)
    if n:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    total = int(lens.sum(dtype=np.uint64))
    buf = np.empty(max(total, 1), dtype=np.uint8)
    origin = np.empty(n, dtype=np.uint32)
    strand = np.empty(n, dtype=np.uint8)
    ref_codes = np.ascontiguousarray(ref_codes, dtype=np.uint8)
    _check("gs_synth_reads", lib().gs_synth_reads(
        _ptr(ref_codes, C.c_uint8), len(ref_codes), C.byref(params), _ptr(offs, C.c_uint64),
        C.cast(buf.ctypes.data, C.c_char_p), _ptr(origin, C.c_uint32), _ptr(strand, C.c_uint8),
        threads))
    return buf[:total], offs, lens, origin, strand


def build_index_host(codes, sa_intv, prefix):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    _check("gs_build_index_host", lib().gs_build_index_host(_ptr(codes, C.c_uint8), len(codes),
                                                            sa_intv, prefix.encode()))


def index_convert(src_prefix, dst_prefix):
    """Writes dst_prefix.bwt/.sa in the fork's layout from either the fork's
    or upstream bwa 0.7's files at src_prefix (gs_index_convert)."""
    _check("gs_index_convert", lib().gs_index_convert(src_prefix.encode(), dst_prefix.encode()))


class BwaRef:
    """bwa's <prefix>.pac/.ann/.amb (gs_bwa_ref_*): codes of the forward
    reference, its sequences and its runs of ambiguous bases."""

    def __init__(self, prefix):
        self.h = C.c_void_p()
        _check("gs_bwa_ref_load", lib().gs_bwa_ref_load(prefix.encode(), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().gs_bwa_ref_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return int(lib().gs_bwa_ref_length(self.h))

    def codes(self):
        n = len(self)
        return np.ctypeslib.as_array(lib().gs_bwa_ref_codes(self.h), shape=(n,)).copy()

    def contigs(self):
        out = []
        for i in range(lib().gs_bwa_ref_n_contigs(self.h)):
            name, off, ln = C.c_char_p(), C.c_uint64(), C.c_uint64()
            _check("gs_bwa_ref_contig", lib().gs_bwa_ref_contig(self.h, i, C.byref(name), C.byref(off),
                                                                C.byref(ln)))
            out.append((name.value.decode(), off.value, ln.value))
        return out

    def holes(self):
        out = []
        for i in range(lib().gs_bwa_ref_n_holes(self.h)):
            off, ln, ch = C.c_uint64(), C.c_uint64(), C.create_string_buffer(1)
            _check("gs_bwa_ref_hole", lib().gs_bwa_ref_hole(self.h, i, C.byref(off), C.byref(ln), ch))
            out.append((off.value, ln.value, ch.raw.decode()))
        return out

    def locate(self, pos):
        c, off = C.c_int(), C.c_uint64()
        _check("gs_bwa_ref_locate", lib().gs_bwa_ref_locate(self.h, pos, C.byref(c), C.byref(off)))
        return c.value, off.value


def set_launch_chunk(items):
    """Work-items per launch of the chunked kernels (test support; 0 = the
    default 2^30)."""
    _check("gs_set_launch_chunk", lib().gs_set_launch_chunk(items))


def roofline_random(buffer_bytes, line_bytes=16, mode=0, inflight=1, grid=8192, iters=256, reps=3):
    r = RooflineResult()
    _check("gs_roofline_random", lib().gs_roofline_random(buffer_bytes, line_bytes, mode, inflight,
                                                          grid, iters, reps, C.byref(r)))
    return r


class Index:
    """A device-resident index (gs_index_open)."""

    def __init__(self, prefix=None, device=0, codes=None, sa_intv=7, write_prefix=None):
        """Opens <prefix>.bwt/.sa, or builds the index of `codes` on the device."""
        self.h = C.c_void_p()
        if codes is None:
            _check("gs_index_open", lib().gs_index_open(prefix.encode(), device, C.byref(self.h)))
        else:
            codes = np.ascontiguousarray(codes, dtype=np.uint8)
            wp = write_prefix.encode() if write_prefix else None
            _check("gs_index_build", lib().gs_index_build(_ptr(codes, C.c_uint8), len(codes), sa_intv,
                                                          device, wp, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().gs_index_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = IndexInfo()
        _check("gs_index_get_info", lib().gs_index_get_info(self.h, C.byref(i)))
        return i

    def read_occ(self, first, count):
        out = np.empty((count, 4), dtype=np.uint32)
        _check("gs_index_read_occ", lib().gs_index_read_occ(self.h, first, count, _ptr(out, C.c_uint32)))
        return out

    def set_kmer_len(self, length):
        """Adds (14..16) or drops (13) the extended interval table."""
        _check("gs_index_set_kmer_len", lib().gs_index_set_kmer_len(self.h, length))

    def read_kext(self, first, count):
        """Extended-table entries: rows of (l, u, presence mask low, high)."""
        out = np.empty((count, 4), dtype=np.uint32)
        _check("gs_index_read_kext", lib().gs_index_read_kext(self.h, first, count, _ptr(out, C.c_uint32)))
        return out

    def read_tandem(self, first, count):
        """Tandem-table entries: rows of (l, u, m, 0), 237 per unit key."""
        out = np.empty((count, 4), dtype=np.uint32)
        _check("gs_index_read_tandem", lib().gs_index_read_tandem(self.h, first, count, _ptr(out, C.c_uint32)))
        return out

    def set_direct(self, on=True):
        """Adds or drops the direct-extension structures (full suffix array,
        its inverse, 2-bit text)."""
        _check("gs_index_set_direct", lib().gs_index_set_direct(self.h, 1 if on else 0))

    def set_pairs(self, on=True):
        """Adds or drops the pair table (one byte per 18-base pattern: which
        bases follow and which precede it in the text)."""
        _check("gs_index_set_pairs", lib().gs_index_set_pairs(self.h, 1 if on else 0))

    def read_pairs(self, first, count):
        out = np.empty(count, dtype=np.uint8)
        _check("gs_index_read_pairs", lib().gs_index_read_pairs(self.h, first, count, _ptr(out, C.c_uint8)))
        return out

    def read_direct(self, first, count):
        """(suffix array rows, inverse positions, text codes) over [first, first+count)."""
        sa = np.empty(count, dtype=np.uint32)
        isa = np.empty(count, dtype=np.uint32)
        text = np.empty(count, dtype=np.uint8)
        _check("gs_index_read_direct", lib().gs_index_read_direct(self.h, first, count, _ptr(sa, C.c_uint32),
                                                                  _ptr(isa, C.c_uint32), _ptr(text, C.c_uint8)))
        return sa, isa, text

    def read_text(self, first, count):
        """Text codes 0..3 over [first, first+count) as the direct-extension
        structures hold them (derived from the index by walking it)."""
        text = np.empty(count, dtype=np.uint8)
        _check("gs_index_read_direct", lib().gs_index_read_direct(self.h, first, count, None, None,
                                                                  _ptr(text, C.c_uint8)))
        return text

    def locate(self, rows):
        """Text positions of suffix-array rows as the reference's walk reports them."""
        rows = np.ascontiguousarray(rows, dtype=np.uint32)
        out = np.empty(len(rows), dtype=np.uint32)
        _check("gs_index_locate", lib().gs_index_locate(self.h, _ptr(rows, C.c_uint32), len(rows),
                                                        _ptr(out, C.c_uint32)))
        return out

    def read_kmer(self, first, count):
        out = np.empty((count, 2), dtype=np.uint32)
        _check("gs_index_read_kmer", lib().gs_index_read_kmer(self.h, first, count, _ptr(out, C.c_uint32)))
        return out


class SeedOutput:
    """Host copy of a gs_result as numpy arrays."""

    @classmethod
    def from_arrays(cls, nseeds, npos, seed_off, pos_off, seeds, pos):
        self = cls.__new__(cls)
        self.n_reads = len(nseeds)
        self.read_nseeds, self.read_npos = nseeds, npos
        self.read_seed_off, self.read_pos_off = seed_off, pos_off
        self.seeds = seeds.view(np.uint32).reshape(-1).view(SEED_DTYPE)
        self.pos = pos
        self.times = {}
        return self

    def __init__(self, res):
        n = res.n_reads
        self.n_reads = n
        self.read_nseeds = np.ctypeslib.as_array(res.read_nseeds, (n,)).copy() if n else np.zeros(0, np.uint32)
        self.read_npos = np.ctypeslib.as_array(res.read_npos, (n,)).copy() if n else np.zeros(0, np.uint32)
        self.read_seed_off = np.ctypeslib.as_array(res.read_seed_off, (n + 1,)).copy()
        self.read_pos_off = np.ctypeslib.as_array(res.read_pos_off, (n + 1,)).copy()
        ns, npos = res.n_seeds, res.n_pos
        if ns:
            raw = np.ctypeslib.as_array(C.cast(res.seeds, C.POINTER(C.c_uint32)), (ns * 6,))
            self.seeds = raw.copy().view(SEED_DTYPE)
        else:
            self.seeds = np.zeros(0, SEED_DTYPE)
        self.pos = np.ctypeslib.as_array(res.pos, (npos,)).copy() if npos else np.zeros(0, np.uint32)
        self.times = dict(h2d=res.t_h2d, seed=res.t_seed, post=res.t_post, locate=res.t_locate,
                          d2h=res.t_d2h)

    def read(self, r):
        """[(qbeg, qend, l, u, [positions])] of read r"""
        out = []
        p = int(self.read_pos_off[r])
        for s in self.seeds[int(self.read_seed_off[r]):int(self.read_seed_off[r + 1])]:
            n = int(s["n_pos"])
            out.append((int(s["qbeg"]), int(s["qend"]), int(s["l"]), int(s["u"]),
                        [int(v) for v in self.pos[p:p + n]]))
            p += n
        return out

    def as_lists(self):
        return [self.read(r) for r in range(self.n_reads)]


def pack_reads(reads):
    lens = np.array([len(r) for r in reads], dtype=np.uint32)
    offs = np.zeros(len(reads), dtype=np.uint64)
    if len(reads):
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    return np.frombuffer(b"".join(reads) or b"\0", dtype=np.uint8), offs, lens


NO_SPLIT = 0xFFFF


NO_DIRECT = 0xFFFF


NO_TANDEM = 0xFFFF


def make_opts(k=20, smem=False, max_occ=0, batch_bases=0, kernel=0, seg_tasks=0, wg=0, walk_rows=0,
              split_tasks=0, split_work=0, direct_min=0, tandem=0):
    return Opts(k, 1 if smem else 0, max_occ, batch_bases, kernel, seg_tasks, wg, walk_rows, split_tasks,
                split_work, direct_min, tandem)


def _arrays(arrays):
    buf, offs, lens = arrays
    return (np.ascontiguousarray(buf, dtype=np.uint8), np.ascontiguousarray(offs, dtype=np.uint64),
            np.ascontiguousarray(lens, dtype=np.uint32))


def seed(index, reads=None, k=20, smem=False, max_occ=0, batch_bases=0, kernel=0, arrays=None,
         text_fd=None, seg_tasks=0, wg=0, walk_rows=0, split_tasks=0, split_work=0, opts=None, direct_min=0,
         tandem=0):
    """Seeds a list of reads (bytes) or arrays=(buf u8, offsets u64, lens u32)."""
    buf, offs, lens = _arrays(arrays if arrays is not None else pack_reads(reads))
    o = opts if opts is not None else make_opts(k, smem, max_occ, batch_bases, kernel, seg_tasks, wg,
                                                walk_rows, split_tasks, split_work, direct_min, tandem)
    smem = bool(o.smem)
    res = Result()
    _check("gs_seed", lib().gs_seed(index.h, C.cast(buf.ctypes.data, C.c_char_p),
                                    _ptr(offs, C.c_uint64), _ptr(lens, C.c_uint32), len(lens),
                                    C.byref(o), C.byref(res)))
    try:
        if text_fd is not None:
            _check("gs_format_text", lib().gs_format_text(C.byref(res), 1 if smem else 0, 1, text_fd))
        return SeedOutput(res)
    finally:
        lib().gs_result_free(C.byref(res))


def seed_noalloc(index, arrays, seeds_cap=None, pos_cap=None, opts=None, **kw):
    """gs_seed_noalloc: the caller's output arrays (sized seeds_cap and
    pos_cap; by default large enough for any result of the call that fits
    the library's fixed device array).  Raises GsError with rc GS_ERANGE and
    attributes n_seeds/n_pos when the result does not fit."""
    buf, offs, lens = _arrays(arrays)
    o = opts if opts is not None else make_opts(**kw)
    n = len(lens)
    if seeds_cap is None:
        seeds_cap = (256 << 20) // 24
    if pos_cap is None:
        pos_cap = (256 << 20) // 4
    nseeds = np.zeros(n, np.uint32)
    npos = np.zeros(n, np.uint32)
    so = np.zeros(n + 1, np.uint64)
    po = np.zeros(n + 1, np.uint64)
    seeds = np.zeros((max(seeds_cap, 1), 6), np.uint32)
    pos = np.zeros(max(pos_cap, 1), np.uint32)
    rb = ResultBuffers(_ptr(nseeds, C.c_uint32), _ptr(npos, C.c_uint32), _ptr(so, C.c_uint64), _ptr(po, C.c_uint64),
                       C.cast(seeds.ctypes.data, C.POINTER(SeedRec)), _ptr(pos, C.c_uint32), seeds_cap, pos_cap, 0, 0)
    rc = lib().gs_seed_noalloc(index.h, C.cast(buf.ctypes.data, C.c_char_p), _ptr(offs, C.c_uint64),
                               _ptr(lens, C.c_uint32), n, C.byref(o), C.byref(rb))
    if rc != GS_OK:
        e = GsError("gs_seed_noalloc", rc)
        e.n_seeds, e.n_pos = rb.n_seeds, rb.n_pos
        raise e
    return SeedOutput.from_arrays(nseeds, npos, so, po, seeds[:rb.n_seeds].copy(), pos[:rb.n_pos].copy())


class BatchStats(C.Structure):
    _fields_ = [("t_h2d", C.c_double), ("t_seed", C.c_double), ("t_count", C.c_double),
                ("t_write", C.c_double), ("t_locate", C.c_double), ("t_total", C.c_double),
                ("n_reads", C.c_uint64), ("n_seeds", C.c_uint64), ("n_pos", C.c_uint64),
                ("n_table_reads", C.c_uint64), ("n_occ_reads", C.c_uint64), ("n_direct_reads", C.c_uint64),
                ("n_tandem_reads", C.c_uint64), ("t_kernel", C.c_double)]


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return
    L.gs_batch_create.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(Opts),
                                  C.POINTER(C.c_void_p)]
    L.gs_batch_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(BatchStats)]
    L.gs_batch_fetch.argtypes = [C.c_void_p, C.POINTER(Result)]
    L.gs_batch_destroy.argtypes = [C.c_void_p]
    L.gs_batch_capture.argtypes = [C.c_void_p, C.c_int]
    L.gs_batch_replay.argtypes = [C.c_void_p, C.POINTER(BatchStats)]
    L.gs_autotune.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                              C.c_uint64, C.POINTER(Opts), C.c_int, C.POINTER(TuneResult)]
    L._batch_bound = True


def autotune(index, arrays, reps=3, opts=None, **kw):
    """Tuned Opts for a sample of reads (gs_autotune) and the TuneResult."""
    L = lib()
    _bind_batch(L)
    buf, offs, lens = _arrays(arrays)
    o = opts if opts is not None else make_opts(**kw)
    out = TuneResult()
    _check("gs_autotune", L.gs_autotune(index.h, C.cast(buf.ctypes.data, C.c_char_p), _ptr(offs, C.c_uint64),
                                        _ptr(lens, C.c_uint32), len(lens), C.byref(o), reps, C.byref(out)))
    return out.opts, out


class Batch:
    """Reads uploaded once (gs_batch_create); run() leaves results in HBM."""

    def __init__(self, index, arrays, k=19, smem=True, max_occ=0, batch_bases=0, kernel=0, seg_tasks=0,
                 wg=0, walk_rows=0, split_tasks=0, split_work=0, opts=None, direct_min=0, tandem=0):
        L = lib()
        _bind_batch(L)
        buf, offs, lens = _arrays(arrays)
        o = opts if opts is not None else make_opts(k, smem, max_occ, batch_bases, kernel, seg_tasks, wg,
                                                    walk_rows, split_tasks, split_work, direct_min, tandem)
        self.smem = bool(o.smem)
        self.h = C.c_void_p()
        _check("gs_batch_create", L.gs_batch_create(index.h, C.cast(buf.ctypes.data, C.c_char_p),
                                                    _ptr(offs, C.c_uint64), _ptr(lens, C.c_uint32),
                                                    len(lens), C.byref(o), C.byref(self.h)))

    def run(self, kernel=-1):
        s = BatchStats()
        _check("gs_batch_run", lib().gs_batch_run(self.h, kernel, C.byref(s)))
        return s

    def capture(self, kernel=-1):
        """Captures the whole pipeline (upload ... download) as a HIP graph;
        needs a run() first (it fixes the output sizes)."""
        _check("gs_batch_capture", lib().gs_batch_capture(self.h, kernel))

    def replay(self):
        s = BatchStats()
        _check("gs_batch_replay", lib().gs_batch_replay(self.h, C.byref(s)))
        return s

    def fetch(self):
        res = Result()
        _check("gs_batch_fetch", lib().gs_batch_fetch(self.h, C.byref(res)))
        try:
            return SeedOutput(res)
        finally:
            lib().gs_result_free(C.byref(res))

    def close(self):
        if self.h:
            lib().gs_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _bind_server(L):
    if getattr(L, "_server_bound", False):
        return
    L.gs_server_create.argtypes = [C.c_void_p, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]
    L.gs_server_submit.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                   C.c_uint64, C.POINTER(C.c_uint64)]
    L.gs_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Result), C.POINTER(RequestInfo)]
    L.gs_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
    L.gs_server_clock.argtypes = [C.c_void_p]
    L.gs_server_clock.restype = C.c_double
    L.gs_server_destroy.argtypes = [C.c_void_p]
    L.gs_server_inject_failure.argtypes = [C.c_void_p, C.c_uint64, C.c_int]
    L._server_bound = True


class Server:
    """Continuous-batching seeding server over a resident index (gs_server_*).
    submit() returns a ticket; wait(ticket) returns (SeedOutput, RequestInfo)."""

    def __init__(self, index, budget_bytes=0, round_bases=0, exclusive=False, workers=0, round_growth=0,
                 opts=None, **kw):
        L = lib()
        _bind_server(L)
        o = ServerOpts()
        o.opts = opts if opts is not None else make_opts(**kw)
        o.budget_bytes = budget_bytes
        o.round_bases = round_bases
        o.exclusive = 1 if exclusive else 0
        o.workers = workers
        o.round_growth = round_growth
        self.h = C.c_void_p()
        self._inputs = {}
        _check("gs_server_create", L.gs_server_create(index.h, C.byref(o), C.byref(self.h)))

    def submit(self, arrays):
        buf, offs, lens = _arrays(arrays)
        t = C.c_uint64()
        _check("gs_server_submit", lib().gs_server_submit(self.h, C.cast(buf.ctypes.data, C.c_char_p),
                                                          _ptr(offs, C.c_uint64), _ptr(lens, C.c_uint32),
                                                          len(lens), C.byref(t)))
        self._inputs[t.value] = (buf, offs, lens)  # alive until wait()
        return t.value

    def wait(self, ticket):
        res, info = Result(), RequestInfo()
        try:
            _check("gs_server_wait", lib().gs_server_wait(self.h, ticket, C.byref(res), C.byref(info)))
            return SeedOutput(res), info
        finally:
            lib().gs_result_free(C.byref(res))
            self._inputs.pop(ticket, None)

    def clock(self):
        return lib().gs_server_clock(self.h)

    def inject_failure(self, round_no, rc):
        """Test support: the round_no-th round taken fails with rc."""
        _check("gs_server_inject_failure", lib().gs_server_inject_failure(self.h, round_no, rc))

    def stats(self):
        st = ServerStats()
        _check("gs_server_get_stats", lib().gs_server_get_stats(self.h, C.byref(st)))
        return st

    def close(self):
        if self.h:
            rc = lib().gs_server_destroy(self.h)
            self.h = C.c_void_p()
            self._inputs.clear()
            _check("gs_server_destroy", rc)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
