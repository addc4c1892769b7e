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


def roofline_random(buffer_bytes, line_bytes=16, mode=0, inflight=1, grid=8192, iters=256, reps=3):
    r = RooflineResult()
    _check("gs_roofline_random", lib().gs_roofline_random(buffer_bytes, line_bytes, mode, inflight,
                                                          grid, iters, reps, C.byref(r)))
    return r
