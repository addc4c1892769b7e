"""The C ABI library loads on a machine without a GPU, exports exactly what
include/gpuseed.h declares, and fails loudly (never falls back) without one."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "gpuseed.h")).read()
    return sorted(set(re.findall(r"GS_API\s+[\w\s\*]+?\b(gs_\w+)\s*\(", src)))


def exported():
    so = os.path.join(ROOT, "gpuseed_amd", "libgpuseed.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    return sorted({l.split()[-1] for l in out.splitlines() if " T " in l})


def test_every_declared_symbol_is_exported(built):
    d, e = declared(), exported()
    assert d, "no declarations parsed"
    missing = [s for s in d if s not in e]
    assert not missing, missing
    lib = C.CDLL(os.path.join(ROOT, "gpuseed_amd", "libgpuseed.so"))
    for s in d:
        getattr(lib, s)


def test_nothing_else_is_exported(built):
    extra = [s for s in exported() if s not in declared()]
    assert not extra, extra


def test_index_open_reports_missing_and_inconsistent_files(built, tmp_path):
    import gpuseed_amd as g
    from conftest import build_index, random_text
    import random
    with pytest.raises(g.GsError) as ei:
        g.Index(str(tmp_path / "nothing"))
    assert ei.value.rc == g.GS_EIO
    prefix = build_index(tmp_path, random_text(random.Random(1), 500), 7)
    os.rename(prefix + ".sa", prefix + ".sa.bak")
    with pytest.raises(g.GsError) as ei:
        g.Index(prefix)
    assert ei.value.rc == g.GS_EIO
    raw = bytearray(open(prefix + ".sa.bak", "rb").read())
    raw[0] ^= 1  # primary disagrees with the .bwt (seed_gen.cu:2697-2700)
    open(prefix + ".sa", "wb").write(bytes(raw))
    with pytest.raises(g.GsError) as ei:
        g.Index(prefix)
    assert ei.value.rc == g.GS_EFORMAT
    raw[0] ^= 1
    raw[24] ^= 1  # seq_len disagrees (2704-2707)
    open(prefix + ".sa", "wb").write(bytes(raw))
    with pytest.raises(g.GsError) as ei:
        g.Index(prefix)
    assert ei.value.rc == g.GS_EFORMAT
    raw[24] ^= 1
    open(prefix + ".sa", "wb").write(bytes(raw))
    bwt = bytearray(open(prefix + ".bwt", "rb").read())
    bwt[-4] ^= 1  # closing counts disagree with L2
    open(prefix + ".bwt", "wb").write(bytes(bwt))
    with pytest.raises(g.GsError) as ei:
        g.Index(prefix)
    assert ei.value.rc == g.GS_EFORMAT


def test_no_device_means_an_error_not_a_fallback(built, tmp_path):
    import gpuseed_amd as g
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from conftest import build_index, random_text
    import random
    prefix = build_index(tmp_path, random_text(random.Random(2), 300), 7)
    with pytest.raises(g.GsError) as ei:
        g.Index(prefix)
    assert ei.value.rc == g.GS_ENODEV


def test_seed_rejects_bad_options(built):
    import gpuseed_amd as g
    o = g.Opts(13, 1, 0, 0, 0)
    res = g.Result()
    buf = np.zeros(1, np.uint8)
    rc = g.lib().gs_seed(None, C.cast(buf.ctypes.data, C.c_char_p), None, None, 0, C.byref(o),
                         C.byref(res))
    assert rc == g.GS_EINVAL


def test_null_handles_are_rejected(built):
    """Every entry point that takes an index, batch or server handle returns
    GS_EINVAL for a null one (checked before any device call)."""
    import gpuseed_amd as g
    L = C.CDLL(os.path.join(ROOT, "gpuseed_amd", "libgpuseed.so"))
    vp, u32, u64, i = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    scratch = (C.c_uint8 * 4096)()
    p = C.cast(scratch, vp)
    calls = {
        "gs_index_close": ([vp], [None]),
        "gs_index_get_info": ([vp, vp], [None, p]),
        "gs_index_read_occ": ([vp, u64, u64, vp], [None, 0, 1, p]),
        "gs_index_read_kmer": ([vp, u64, u64, vp], [None, 0, 1, p]),
        "gs_index_read_kext": ([vp, u64, u64, vp], [None, 0, 1, p]),
        "gs_index_set_kmer_len": ([vp, i], [None, 14]),
        "gs_index_set_direct": ([vp, i], [None, 1]),
        "gs_index_set_pairs": ([vp, i], [None, 1]),
        "gs_index_read_pairs": ([vp, u64, u64, vp], [None, 0, 1, p]),
        "gs_index_locate": ([vp, vp, u64, vp], [None, p, 1, p]),
        "gs_index_read_direct": ([vp, u32, u32, vp, vp, vp], [None, 0, 1, p, p, p]),
        "gs_batch_create": ([vp, vp, vp, vp, u64, vp, vp], [None, p, p, p, 0, p, p]),
        "gs_batch_run": ([vp, i, vp], [None, -1, None]),
        "gs_batch_fetch": ([vp, vp], [None, p]),
        "gs_batch_destroy": ([vp], [None]),
        "gs_batch_capture": ([vp, i], [None, 3]),
        "gs_batch_replay": ([vp, vp], [None, None]),
        "gs_server_create": ([vp, vp, vp], [None, p, p]),
        "gs_server_submit": ([vp, vp, vp, vp, u64, vp], [None, p, p, p, 0, p]),
        "gs_server_wait": ([vp, u64, vp, vp], [None, 0, p, None]),
        "gs_server_get_stats": ([vp, vp], [None, p]),
        "gs_server_destroy": ([vp], [None]),
        "gs_server_inject_failure": ([vp, u64, i], [None, 0, -3]),
    }
    for name, (types, args) in calls.items():
        fn = getattr(L, name)
        fn.argtypes, fn.restype = types, C.c_int
        assert fn(*args) == g.GS_EINVAL, name
