"""bench.py's roofline inputs from the committed rocprofv3 summaries: the
counter file parses, applies only to the workload and kernel options it was
measured on, gives the HBM traffic per launch the summary states, and holds
the same counter values as the human-readable summary beside it."""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..")
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_config3_counters_apply_only_to_their_workload():
    kv = bench.profiled_counters(3, 3, 16, 1, 1, 6)
    assert kv is not None and kv["reads_per_launch"] == 10000000
    for other in [(3, 1, 16, 1, 1, 6), (3, 3, 13, 1, 1, 6), (3, 3, 16, 0, 1, 6), (3, 3, 16, 1, 0, 6),
                  (3, 3, 16, 1, 1, 0), (2, 3, 16, 1, 1, 6)]:
        assert bench.profiled_counters(*other) is None, other


def test_config3_traffic_matches_the_summary():
    kv = bench.profiled_counters(3, 3, 16, 1, 1, 6)
    traffic, src = bench.profiled_traffic(kv, 10000000)
    # FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, KiB per dispatch: 328.2 GB per launch
    assert abs(traffic / 1e9 - 328.2) < 0.1 and "seed_persistent_config3.txt" in src
    half, src2 = bench.profiled_traffic(kv, 5000000)
    assert abs(half - traffic / 2) < 1 and "scaled" in src2 and "5000000" in src2
    assert kv["tcc_miss"] / (kv["tcc_hit"] + kv["tcc_miss"]) < 1.0


def test_counter_file_agrees_with_the_markdown_summary():
    kv = bench.profiled_counters(3, 3, 16, 1, 1, 6)
    md = open(os.path.join(ROOT, "profiles", "seed_persistent_config3.md")).read()

    def counter(name):
        m = re.search(r"^%s\s+([\d,.]+)" % name, md, re.M)
        assert m, name
        return float(m.group(1).replace(",", ""))

    assert counter("FETCH_SIZE") == kv["fetch_kib"]
    assert counter("WRITE_SIZE") == kv["write_kib"]
    assert counter("TCC_HIT_sum") == kv["tcc_hit"]
    assert counter("TCC_MISS_sum") == kv["tcc_miss"]


def test_more_than_one_gpu_is_refused_with_a_message():
    import subprocess
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=120)
    assert p.returncode != 0 and p.stdout == b""
    assert b"--gpus 1" in p.stderr
