"""The GPU index builder (gs_index_build) against the CPU SA-IS writer: the
reference-format files must be byte-identical, and so must the device
structures and the seeds obtained from either index."""
import os
import random

import numpy as np
import pytest

from conftest import random_text, repetitive_text

pytestmark = pytest.mark.gpu


def codes_of(text):
    return np.frombuffer(text.encode().translate(bytes.maketrans(b"ACGT", b"\0\1\2\3")), dtype=np.uint8)


_r = random.Random(21)
TEXTS = [random_text(_r, n) for n in (1, 2, 13, 21, 22, 64, 95, 96, 97, 128, 1000, 4099)]
TEXTS += ["A" * 500, "ACGT" * 300, "AC" * 777, repetitive_text(_r, 20000), "T" * 21 + "A" * 21]


@pytest.mark.parametrize("ti", range(len(TEXTS)))
def test_gpu_index_files_equal_cpu_files(built, tmp_path, ti):
    import gpuseed_amd as g
    text = TEXTS[ti]
    print("build n=%d" % len(text), flush=True)
    codes = codes_of(text)
    for intv in (1, 7):
        cpu = str(tmp_path / ("cpu%d" % intv))
        gpu = str(tmp_path / ("gpu%d" % intv))
        g.build_index_host(codes, intv, cpu)
        with g.Index(codes=codes, sa_intv=intv, write_prefix=gpu) as Xg:
            with g.Index(cpu) as Xc:
                nb = Xc.info().n_blocks
                assert np.array_equal(Xg.read_occ(0, nb * 4 + 4), Xc.read_occ(0, nb * 4 + 4))
                assert np.array_equal(Xg.read_kmer(0, 1 << 20), Xc.read_kmer(0, 1 << 20))
                ig, ic = Xg.info(), Xc.info()
                assert (ig.primary, list(ig.L2), ig.seq_len) == (ic.primary, list(ic.L2), ic.seq_len)
        for ext in (".bwt", ".sa"):
            assert open(cpu + ext, "rb").read() == open(gpu + ext, "rb").read(), ext


def test_config1_index_and_seeds_from_gpu_build(built, tmp_path):
    import gpuseed_amd as g
    rp, qp = g.preset(1)
    ref = g.synth_reference(rp)
    qp.n_reads = 4000
    arrays = g.synth_reads(ref, qp)[:3]
    cpu = str(tmp_path / "cpu")
    gpu = str(tmp_path / "gpu")
    g.build_index_host(ref, 7, cpu)
    with g.Index(codes=ref, sa_intv=7, write_prefix=gpu) as Xg:
        info = Xg.info()
        print("config1 build: sa %.3fs in %d rounds, occ %.3fs, table %.3fs" %
              (info.t_sa, info.sa_rounds, info.t_repack, info.t_table), flush=True)
        a = g.seed(Xg, arrays=arrays, k=19, smem=True)
        b = g.seed(Xg, arrays=arrays, k=20, smem=False)
    for ext in (".bwt", ".sa"):
        assert open(cpu + ext, "rb").read() == open(gpu + ext, "rb").read(), ext
    with g.Index(cpu) as Xc:
        c = g.seed(Xc, arrays=arrays, k=19, smem=True)
        d = g.seed(Xc, arrays=arrays, k=20, smem=False)
    assert np.array_equal(a.seeds, c.seeds) and np.array_equal(a.pos, c.pos)
    assert np.array_equal(b.seeds, d.seeds) and np.array_equal(b.pos, d.pos)


def test_batch_api_matches_gs_seed(built, tmp_path):
    import gpuseed_amd as g
    rp, qp = g.preset(1)
    rp.length = 300000
    ref = g.synth_reference(rp)
    qp.n_reads = 3000
    arrays = g.synth_reads(ref, qp)[:3]
    with g.Index(codes=ref, sa_intv=7) as X:
        one = g.seed(X, arrays=arrays, k=19, smem=True)
        bt = g.Batch(X, arrays, k=19, smem=True)
        s1 = bt.run()
        s2 = bt.run()
        two = bt.fetch()
        s3 = bt.run(kernel=1)
        three = bt.fetch()
        s4 = bt.run(kernel=2)
        four = bt.fetch()
        bt.close()
    assert s1.n_seeds == s2.n_seeds == s3.n_seeds == s4.n_seeds == len(one.seeds)
    for o in (two, three, four):
        assert np.array_equal(one.seeds, o.seeds) and np.array_equal(one.pos, o.pos)
    assert s2.t_seed > 0 and s2.t_total >= s2.t_seed
    # the extension kernel's own launch lies inside the seed stage's events
    for s in (s1, s2, s3, s4):
        assert 0 < s.t_kernel <= s.t_seed
    assert s3.n_table_reads > 0 and s3.n_occ_reads > 0
