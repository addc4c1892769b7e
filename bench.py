#!/usr/bin/env python3
"""Throughput benchmark of the MI355X seeder on one GPU.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config 3] ...

Workload (default --config 3): a 3.1 Gbp synthetic reference (41 % GC, 45 %
interspersed repeats, 3 % tandem repeats, 5 % segmental duplications) indexed
on the GPU with suffix-array sampling 7, and 150 bp reads (0.5 %
substitutions, 0.05 % indels, both strands) seeded in SMEM mode with -k 19,
reporting positions for intervals with at most 500 occurrences.  A step is
one pass of the seeding pipeline (extension kernel, SMEM filtering, scans,
suffix-array walk) over one batch of --batch reads that were uploaded to HBM
before the timed region; results stay in HBM.  By default a batch is the
configuration's whole read set (configuration 3: 10,000,000 reads, about
45 GB of device buffers), one launch; at most about 120 GB of batches stay
resident and further steps re-run them in turn.

--config 4 seeds 2,000,000 reads of uneven length on the same index (70 %
100-250 bp, 25 % 1-5 kbp, 5 % 5-20 kbp, the long ones with 8 % errors) as
one batch; every step re-runs the whole pipeline on that resident batch.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

T0 = time.time()


def log(*a):
    print("[bench %7.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpus():
    """(logical CPUs of the machine, CPUs this process may run on)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return os.cpu_count() or avail, avail


def work_model(prefix, arrays, sample, threads, k, max_occ, ext_len, smem=True, direct_min=0, prune_sat=0,
               pair_probe=0, tandem=0):
    """Memory operations per read of the device algorithm, counted by the
    oracle on `sample` reads spread over the batch: 13-mer table lookups (8 bytes),
    extended-table lookups (16 bytes: interval and presence bits), occ
    entries (16 bytes) and direct-extension reads (4-byte suffix-array and
    inverse words, 16-byte text chunks), pair-table bytes and tandem-table
    entries (16 bytes), with distance-1 pruning, the SMEM read-start
    stop and tasks that cannot reach k bases ending at once; size records
    saturate at prune_sat like the kernel's (0: exact)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    sub = strided_sample(arrays, sample)
    n = len(sub[2])
    ox = ob.Index(prefix)
    c = ob.seed_arrays(ox, *sub, k, smem, max_occ=max_occ, threads=threads, smem_stop=1,
                       ext_len=ext_len if ext_len > 13 else 0, miss_skip=1, direct_min=direct_min,
                       prune_sat=prune_sat, pair_probe=pair_probe, tandem=tandem)[4]
    ox.close()
    table = (c["w_tasks"] + c["w_ext"] + c["w_pair"]) / n
    occ = c["w_blocks96"] / n
    direct, text = c["w_direct"] / n, c["w_text"] / n
    tandem_rd = c["w_tandem"] / n
    ext_bytes = 16.0 if ext_len > 13 else 8.0
    return dict(reads=n, table_per_read=table, occ_per_read=occ, direct_per_read=direct, tandem_per_read=tandem_rd,
                requests_per_read=table + occ + direct + tandem_rd,
                pair_per_read=c["w_pair"] / n,
                bytes_per_read=8.0 * c["w_tasks"] / n + ext_bytes * c["w_ext"] / n + 1.0 * c["w_pair"] / n + 16.0 * occ
                + 4.0 * (direct - text) + 16.0 * text + 16.0 * tandem_rd)


def cpu_baseline(prefix, arrays, gpu_out, sample, threads, k, max_occ, smem=True):
    """The oracle (C restatement of the reference path, distance-1 pruning,
    13-mer table built before timing) on the first `sample` reads of a batch;
    its output is also compared bit for bit with the GPU's for those reads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    import oracle_bind as ob
    buf, offs, lens = arrays
    n = min(sample, len(lens))
    end = int(offs[n - 1] + lens[n - 1])
    sub = (buf[:end], offs[:n], lens[:n])
    ox = ob.Index(prefix)
    t = time.time()
    ob.seed_arrays(ox, buf[:int(lens[0])], offs[:1], lens[:1], k, smem, max_occ=max_occ, threads=threads)
    log("oracle: index loaded, 13-mer table built in %.1fs" % (time.time() - t))
    t = time.time()
    seeds, pos, so, po, counters = ob.seed_arrays(ox, *sub, k, smem, max_occ=max_occ, threads=threads)
    dt = time.time() - t
    log("oracle: %d reads in %.2fs on %d threads" % (n, dt, threads))
    ox.close()
    g_seeds = gpu_out.seeds.view(np.uint32).reshape(-1, 6)[:int(gpu_out.read_seed_off[n])]
    g_pos = gpu_out.pos[:int(gpu_out.read_pos_off[n])]
    same_off = (np.array_equal(so, gpu_out.read_seed_off[:n + 1]) and
                np.array_equal(po, gpu_out.read_pos_off[:n + 1]))
    same = same_off and np.array_equal(seeds, g_seeds) and np.array_equal(pos, g_pos)
    bad = 0 if same else n
    if not same and same_off:
        # count the reads whose seeds or positions differ
        sd = np.any(seeds != g_seeds, axis=1)
        pd = pos != g_pos
        rs = set(np.searchsorted(so, np.nonzero(sd)[0], side="right") - 1)
        rs |= set(np.searchsorted(po, np.nonzero(pd)[0], side="right") - 1)
        bad = len(rs)
    return dict(reads=n, seconds=dt, reads_per_s=n / dt, counters=counters, mismatches=bad,
                bit_exact=bool(same))


def make_requests(pool, n, max_reads, seed):
    """n requests of log-uniform size in [1, max_reads] reads, each a run of
    consecutive reads of the pool (views of its bytes, offsets rebased)."""
    import numpy as np
    buf, offs, lens = pool
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0.0, np.log(max_reads), n)).round().astype(np.int64).clip(1, max_reads)
    out = []
    for m in sizes:
        a = int(rng.integers(0, len(lens) - m + 1))
        lo, hi = int(offs[a]), int(offs[a + m - 1] + lens[a + m - 1])
        out.append((buf[lo:hi], offs[a:a + m] - np.uint64(lo), lens[a:a + m]))
    return out


def run_load(g, X, requests, rate, opts, exclusive, seed, label, growth=0, workers=0):
    """Poisson arrivals at `rate` requests/s (0: all at once) into a server;
    returns latencies (server clock: submit to results complete) and stats."""
    import queue
    import threading

    import numpy as np
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(requests)) if rate > 0 else np.zeros(len(requests))
    arrivals = np.cumsum(gaps) - gaps[0]
    lat, infos, errors = [], [], []
    tickets = queue.Queue()
    with g.Server(X, opts=opts, exclusive=exclusive, round_growth=growth, workers=workers) as S:
        def waiter():
            while True:
                t = tickets.get()
                if t is None:
                    return
                try:
                    _, info = S.wait(t)
                    infos.append(info)
                except Exception as e:  # reported below
                    errors.append(e)
        th = [threading.Thread(target=waiter) for _ in range(8)]
        for t in th:
            t.start()
        t0 = S.clock()
        for r, at in zip(requests, arrivals):
            d = t0 + at - S.clock()
            if d > 0:
                time.sleep(d)
            tickets.put(S.submit(r))
        for _ in th:
            tickets.put(None)
        for t in th:
            t.join()
        st = S.stats()
    if errors:
        raise errors[0]
    lat = np.array([i.t_done - i.t_submit for i in infos])
    sizes = np.array([int(i.n_reads) for i in infos])
    # latency by request size: tenths of the requests ordered by read count
    deciles = []
    order = np.argsort(sizes, kind="stable")
    for d in range(10):
        sel = order[d * len(order) // 10:(d + 1) * len(order) // 10]
        if len(sel):
            deciles.append({"reads_min": int(sizes[sel].min()), "reads_max": int(sizes[sel].max()),
                            "p50_ms": float(np.percentile(lat[sel], 50) * 1e3),
                            "p99_ms": float(np.percentile(lat[sel], 99) * 1e3)})
    reads = sum(int(i.n_reads) for i in infos)
    span = max(i.t_done for i in infos) - min(i.t_submit for i in infos)
    out = dict(requests=len(infos), reads=reads, seconds=span, reads_per_s=reads / span,
               requests_per_s=len(infos) / span, p50_ms=float(np.percentile(lat, 50) * 1e3),
               p99_ms=float(np.percentile(lat, 99) * 1e3), max_ms=float(lat.max() * 1e3),
               by_size_decile=deciles,
               rounds=int(st.n_rounds), device_busy=st.busy_seconds / max(span, 1e-9),
               round_ms={"plan": st.t_plan / max(st.n_rounds, 1) * 1e3,
                         "load": st.t_load / max(st.n_rounds, 1) * 1e3, "run": st.t_run / max(st.n_rounds, 1) * 1e3,
                         "pipeline": st.device_seconds / max(st.n_rounds, 1) * 1e3,
                         "seed": st.t_seed / max(st.n_rounds, 1) * 1e3,
                         "post": st.t_post / max(st.n_rounds, 1) * 1e3,
                         "locate": st.t_locate / max(st.n_rounds, 1) * 1e3,
                         "fetch": st.t_fetch / max(st.n_rounds, 1) * 1e3,
                         "split": st.t_split / max(st.n_rounds, 1) * 1e3},
               worker_idle=st.t_idle / max(span, 1e-9), bases=int(st.n_bases))
    log("%s: %d requests, %.0f reads/s, p50 %.1f ms, p99 %.1f ms, %d rounds; per round ms %s; workers idle %.2f" %
        (label, out["requests"], out["reads_per_s"], out["p50_ms"], out["p99_ms"], out["rounds"],
         " ".join("%s %.1f" % kv for kv in out["round_ms"].items()), out["worker_idle"]))
    return out


def serve_bench(g, X, ref, k, max_occ, args):
    """Configuration 5: requests of 1-200,000 reads (log-uniform) of the
    configuration-4 length mix against the resident index, Poisson arrivals
    at fractions of the measured peak; the batched server against a server
    that runs each request alone with the plain kernel."""
    rp, qp = g.preset(5)
    qp.n_reads = args.serve_pool
    log("serving: generating a pool of %d configuration-4 reads" % qp.n_reads)
    pool = g.synth_reads(ref, qp)[:3]
    opts = g.make_opts(k=k, smem=True, max_occ=max_occ, kernel=args.kernel, seg_tasks=args.seg_tasks,
                       split_tasks=args.serve_split_tasks, tandem=g.NO_TANDEM if args.no_tandem else 0)
    plain = g.make_opts(k=k, smem=True, max_occ=max_occ, kernel=0)
    # peak: a burst the server drains at full speed
    burst = make_requests(pool, args.serve_burst, 200000, seed=50)
    peak = run_load(g, X, burst, 0, opts, False, 1, "peak (burst of %d)" % len(burst), args.serve_growth,
                    args.serve_workers)
    levels = []
    for i, frac in enumerate(args.serve_levels):
        rate = frac * peak["requests_per_s"]
        n = max(20, int(rate * args.serve_seconds))
        reqs = make_requests(pool, n, 200000, seed=100 + i)
        b = run_load(g, X, reqs, rate, opts, False, 200 + i, "batched at %.0f%% (%.1f req/s)" % (frac * 100, rate),
                     args.serve_growth, args.serve_workers)
        nb = min(n, args.serve_plain_requests)
        p = None if not nb else run_load(g, X, reqs[:nb], rate, plain, True, 200 + i,
                                         "plain per-request at %.0f%%" % (frac * 100))
        levels.append(dict(load=frac, rate_requests_per_s=rate, batched=b, plain=p))
    return dict(peak=peak, levels=levels,
                workload="requests of 1-200,000 reads (log-uniform), 70%% 100-250 bp / 25%% 1-5 kbp / 5%% 5-20 kbp, "
                         "from a pool of %d reads; Poisson arrivals; latency = submit to results complete" % qp.n_reads)


def profiled_counters(config, kernel, kmer_len, direct, pairs, tandem):
    """The committed rocprofv3 counter summary of the extension kernel
    (profiles/seed_persistent_config<N>.txt) if it was measured on this
    workload, index shape and kernel options, else None."""
    path = os.path.join(ROOT, "profiles", "seed_persistent_config%d.txt" % config)
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#"):
            kv[parts[0]] = float(parts[1])
    if (kv.get("kernel") != kernel or kv.get("kmer_len") != kmer_len or kv.get("direct", 0) != direct
            or kv.get("pairs", 0) != pairs or kv.get("tandem", 0) != tandem):
        return None
    kv["path"] = os.path.relpath(path, ROOT)
    return kv


def profiled_traffic(kv, reads):
    """HBM bytes per launch of the extension kernel from the FETCH_SIZE and
    WRITE_SIZE counters, scaled to `reads` per launch."""
    if not kv:
        return None, None
    per_launch = (kv["fetch_kib"] * kv["fetch_factor"] + kv["write_kib"] * kv["write_factor"]) * 1024.0
    src = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE at %d reads per launch (%s), scaled by the stream-copy "
           "calibration%s" % (kv["reads_per_launch"], kv["path"],
                              "" if kv["reads_per_launch"] == reads else " and to %d reads per launch" % reads))
    return per_launch * reads / kv["reads_per_launch"], src


def sample_arrays(arrays, n):
    """The first n reads."""
    buf, offs, lens = arrays
    n = min(n, len(lens))
    end = int(offs[n - 1] + lens[n - 1])
    return buf[:end], offs[:n], lens[:n]


def strided_sample(arrays, n):
    """n reads spread evenly over the whole batch (every len/n-th read),
    packed into arrays of their own."""
    import numpy as np
    buf, offs, lens = arrays
    if n >= len(lens):
        return arrays
    idx = (np.arange(n, dtype=np.int64) * len(lens)) // n
    L = lens[idx].astype(np.int64)
    starts = offs[idx].astype(np.int64)
    new_offs = np.zeros(n, dtype=np.uint64)
    new_offs[1:] = np.cumsum(L[:-1])
    src = np.repeat(starts - new_offs.astype(np.int64), L) + np.arange(int(L.sum()), dtype=np.int64)
    return buf[src], new_offs, lens[idx].copy()


def cli_end_to_end(arrays, prefix, tmpdir, n_reads, k, max_occ, kmer_len, direct, smem=True, pairs=0):
    """The drop-in command line, seed_gen, on a FASTA file of the first
    n_reads reads: index files loaded and structures built by the process
    itself, reads parsed from the file, seeds formatted to /dev/null with -o
    and not formatted without it.  Wall time of the process, and its own
    summary (index ready, then seeding)."""
    import re
    import subprocess
    buf, offs, lens = sample_arrays(arrays, n_reads)
    fa = os.path.join(tmpdir, "reads.fa")
    with open(fa, "wb") as f:
        mv = memoryview(buf)
        for i in range(len(lens)):
            o = int(offs[i])
            f.write(b">r\n")
            f.write(mv[o:o + int(lens[i])])
            f.write(b"\n")
    exe = os.path.join(ROOT, "gpuseed_amd", "bin", "seed_gen")
    out = {"reads": int(len(lens)), "bases": int(lens.sum()),
           "command": "seed_gen %s-k %d --max-occ %d --kmer-len %d%s%s [-o] <prefix> reads.fa"
                      % ("-s " if smem else "", k, max_occ, kmer_len, " --direct" if direct else "",
                         " --pairs" if pairs else "")}
    # each form twice, alternating; the faster run of each is kept (a run
    # right after another process released its device memory can be slower)
    for rep in range(2):
        for label, extra in (("with_o", ["-o"]), ("without_o", [])):
            cmd = [exe] + (["-s"] if smem else []) + ["-k", str(k), "--max-occ", str(max_occ), "--kmer-len",
                                                      str(kmer_len)] + (["--direct"] if direct else []) + \
                (["--pairs"] if pairs else []) + extra
            cmd += [prefix, fa]
            t = time.perf_counter()
            p = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
            wall = time.perf_counter() - t
            if p.returncode != 0:
                raise RuntimeError("seed_gen failed (%d): %s" % (p.returncode, p.stderr[-2000:]))
            err = p.stderr
            m = re.search(r"index ready after ([0-9.]+) s; seeding ([0-9.]+) s", err)
            t_index, t_seed = (float(m.group(1)), float(m.group(2))) if m else (None, None)
            m = re.search(r"time: total [0-9.]+ s, input ([0-9.]+) s, device ([0-9.]+) s \(kernels ([0-9.]+) s\), "
                          r"output ([0-9.]+) s", err)
            parts = dict(zip(("input_s", "device_s", "kernels_s", "output_s"), map(float, m.groups()))) if m else {}
            r = {"wall_s": wall, "reads_per_s_wall": len(lens) / wall, "index_ready_s": t_index,
                 "seeding_s": t_seed, "reads_per_s_after_index": len(lens) / t_seed if t_seed else None}
            r.update(parts)
            log("seed_gen %s: %.2f s wall (index ready after %s s, seeding %s s: %s)" %
                (label, wall, t_index, t_seed, " ".join("%s %.3f" % kv for kv in parts.items())))
            if label not in out or wall < out[label]["wall_s"]:
                out[label] = r
    os.unlink(fa)
    return out


def timed_steps(batches, warmup, kernel):
    for b in range(warmup):
        s = batches[b][1].run(kernel)
        log("kernel %d warmup %d: %.1f ms (seed %.1f ms)" % (kernel, b, s.t_total * 1e3, s.t_seed * 1e3))
    stats = []
    t0 = time.perf_counter()
    for b in range(warmup, len(batches)):
        stats.append(batches[b][1].run(kernel))
    t1 = time.perf_counter()
    for i, s in enumerate(stats):
        log("kernel %d step %d: %.1f ms total, seed %.1f (kernel %.1f), count %.1f, write %.1f, locate %.1f; "
            "%d seeds, %d positions" % (kernel, i, s.t_total * 1e3, s.t_seed * 1e3, s.t_kernel * 1e3,
                                        s.t_count * 1e3, s.t_write * 1e3, s.t_locate * 1e3, s.n_seeds, s.n_pos))
    if kernel and stats and stats[0].n_reads:
        s = stats[0]
        log("kernel %d requests per read: %.1f table, %.1f occ, %.1f direct, %.1f tandem table" %
            (kernel, s.n_table_reads / s.n_reads, s.n_occ_reads / s.n_reads, s.n_direct_reads / s.n_reads,
             s.n_tandem_reads / s.n_reads))
    return t1 - t0, stats


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, choices=[1, 2, 3, 4])
    ap.add_argument("--batch", type=int, default=0,
                    help="reads per step (default: the configuration's read count, e.g. configuration 3's "
                         "10,000,000 reads in one launch)")
    ap.add_argument("--k", type=int, default=0, help="default: 20 for config 1, 19 otherwise")
    ap.add_argument("--max-occ", type=int, default=-1, help="default: 0 for config 1, 500 otherwise")
    ap.add_argument("--ref-length", type=int, default=0, help="override the reference length")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="reads timed on the CPU (default 1,000,000; configuration 4: 20,000)")
    ap.add_argument("--seg-tasks", type=int, default=0, help="task ends per work item of long reads")
    ap.add_argument("--work-sample", type=int, default=200000, help="reads whose work the oracle counts")
    ap.add_argument("--reuse", action="store_true", help="one resident batch re-run every step")
    ap.add_argument("--kmer-len", type=int, default=0,
                    help="interval table length: 13 is the reference's table, 14..16 add a longer one "
                         "(default 16; configuration 1, a 1 Mbp reference: 13)")
    ap.add_argument("--no-direct", action="store_true",
                    help="no direct-extension structures (full suffix array, inverse, text)")
    ap.add_argument("--no-pairs", action="store_true",
                    help="no pair table (one byte per 18-base pattern: the bases next to it in the text)")
    ap.add_argument("--no-tandem", action="store_true",
                    help="no tandem jumps in the batched kernel (gs_opts.tandem 0xFFFF)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", type=int, default=3,
                    help="batched kernel variant (1: 8-bit, 2: 16-bit, 3: 4-bit size records)")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain kernel")
    ap.add_argument("--no-graph", action="store_true", help="skip the host-to-host (HIP graph) measurement")
    ap.add_argument("--no-reference-shape", action="store_true",
                    help="skip the second timing on the reference's index shape (13-mer table, sampled "
                         "suffix-array walk, no direct-extension structures)")
    ap.add_argument("--cli-reads", type=int, default=2000000,
                    help="reads of the seed_gen end-to-end measurement (0: none)")
    ap.add_argument("--profile-steps", action="store_true", help="stop after the timed steps")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement (configuration 5)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving measurement")
    ap.add_argument("--serve-levels", type=float, nargs="*", default=[0.1, 0.5, 0.8, 0.9],
                    help="offered loads as fractions of the measured peak")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="arrival window per load level")
    ap.add_argument("--serve-pool", type=int, default=600000, help="reads in the request pool")
    ap.add_argument("--serve-burst", type=int, default=60, help="requests of the peak burst")
    ap.add_argument("--serve-plain-requests", type=int, default=25,
                    help="requests per load level for the plain per-request server (0: none)")
    ap.add_argument("--serve-workers", type=int, default=0, help="device workers of the server (0: its default, 3)")
    ap.add_argument("--serve-split-tasks", type=int, default=0,
                    help="work splitting in served rounds: split_tasks (0: the library's default)")
    ap.add_argument("--serve-growth", type=int, default=0,
                    help="rounds grow up to this many times 16 Mbases under backlog (0: the server's "
                         "default, 4; 1: fixed rounds)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU. Reads shard across GPUs without exchange (each GPU holds a "
              "replica of the index and seeds its own share of the reads), so N GPUs give N times the "
              "one-GPU rate; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import gpuseed_amd as g
    smem = args.config != 1  # configuration 1 is the reference's MEM mode
    if not args.kmer_len:
        args.kmer_len = 13 if args.config == 1 else 16
    k = args.k or (20 if args.config == 1 else 19)
    max_occ = args.max_occ if args.max_occ >= 0 else (0 if args.config == 1 else 500)
    rp, qp = g.preset(args.config)
    if args.ref_length:
        rp.length = args.ref_length
    reuse = args.config == 4 or args.reuse  # one resident batch, re-run every step
    if not args.batch:
        args.batch = qp.n_reads
    if not args.cpu_sample:
        args.cpu_sample = 20000 if args.config == 4 else 1000000
    n_batches = 1 if reuse else args.warmup + args.steps
    # at most about 120 GB of resident batches (candidate slots, 32 bytes per
    # task, dominate); further steps cycle through them, each re-running the
    # whole pipeline on its batch
    w = [qp.classes[i].weight for i in range(qp.n_classes)]
    mean_len = sum(wi * (qp.classes[i].len_min + qp.classes[i].len_max) / 2 for i, wi in enumerate(w)) / sum(w)
    resident = max(1, min(n_batches, int((120e9 if args.no_pairs else 60e9) // ((36.0 * mean_len + 200) * args.batch))))
    qp.n_reads = resident * args.batch

    log("generating %.2f Gbp reference (config %d)" % (rp.length / 1e9, args.config))
    ref = g.synth_reference(rp)
    need = rp.length * 1.2 + (1 << 30)  # .bwt + .sa files
    where = [d for d in (tempfile.gettempdir(), "/dev/shm", ROOT)
             if os.path.isdir(d) and os.access(d, os.W_OK) and shutil.disk_usage(d).free > need]
    tmpdir = tempfile.mkdtemp(prefix="gsbench_", dir=where[0] if where else None)
    prefix = os.path.join(tmpdir, "ref")
    log("building the index on the GPU (files under %s, %.0f GB free)" %
        (tmpdir, shutil.disk_usage(tmpdir).free / 1e9))
    t = time.time()
    X = g.Index(codes=ref, sa_intv=7, write_prefix=prefix)
    info = X.info()
    t_build = time.time() - t
    log("index: %.1fs total; suffix sort %.1fs in %d rounds, occ %.1fs, files %.1fs, table %.2fs, "
        "%.2f GB on device" % (t_build, info.t_sa, info.sa_rounds, info.t_repack, info.t_write,
                               info.t_table, info.device_bytes / 1e9))
    if args.kmer_len > 13:
        X.set_kmer_len(args.kmer_len)
        info = X.info()
        log("%d-mer interval table in %.2fs, %.2f GB on device" % (info.kmer_len, info.t_ext, info.device_bytes / 1e9))
    direct = 0 if args.no_direct or args.config == 1 else 1
    if direct:
        X.set_direct(True)
        info = X.info()
        log("direct-extension structures in %.2fs, %.2f GB on device" % (info.t_direct, info.device_bytes / 1e9))
    pairs = 0 if args.no_pairs or args.config == 1 or info.kmer_len != 16 or k < 19 else 1
    tandem = g.NO_TANDEM if args.no_tandem else 0
    if pairs:
        X.set_pairs(True)
        info = X.info()
        log("pair table in %.2fs, %.2f GB on device" % (info.t_pairs, info.device_bytes / 1e9))
    if args.serve_only:
        serving = serve_bench(g, X, ref, k, max_occ, args)
        X.close()
        shutil.rmtree(tmpdir, ignore_errors=True)
        print(json.dumps({"metric": "serving (configuration 5 workload)", "serving": serving}), flush=True)
        return
    log("generating %d reads" % qp.n_reads)
    buf, offs, lens, _, _ = g.synth_reads(ref, qp)
    serve_ref = ref if (args.config == 3 and not args.no_serve and not args.profile_steps) else None
    del ref
    batches = []
    for b in range(resident):
        lo, hi = b * args.batch, (b + 1) * args.batch
        base = int(offs[lo])
        end = int(offs[hi - 1] + lens[hi - 1])
        arrays = (buf[base:end], offs[lo:hi] - np.uint64(base), lens[lo:hi])
        batches.append((arrays, g.Batch(X, arrays, k=k, smem=smem, max_occ=max_occ, seg_tasks=args.seg_tasks,
                                        kernel=args.kernel, tandem=tandem)))
    log("%d batches of %d reads (%.2f Gbases) resident in HBM" % (resident, args.batch, len(buf) / 1e9))
    batches = [batches[i % resident] for i in range(args.warmup + args.steps)]

    shape = {"kmer_table": info.kmer_len, "direct_extension": bool(direct), "pair_table": bool(pairs),
             "device_bytes": int(info.device_bytes)}
    elapsed, stats = timed_steps(batches, args.warmup, args.kernel)
    reads = args.steps * args.batch
    value = reads / elapsed
    log("batched kernel %d: %d reads in %.3f s -> %.0f reads/s" % (args.kernel, reads, elapsed, value))
    if args.profile_steps:
        print(json.dumps({"metric": "profile run", "value": value, "reads_per_step": args.batch,
                          "seed_stage_ms_hip_events": [s.t_seed * 1e3 for s in stats],
                          "seed_persistent_ms_hip_events": [s.t_kernel * 1e3 for s in stats]}), flush=True)
        return
    plain = None
    if not args.no_plain:
        p_elapsed, p_stats = timed_steps(batches, min(1, args.warmup), 0)
        plain = dict(reads_per_s=(len(batches) - min(1, args.warmup)) * args.batch / p_elapsed,
                     seed_ms=sum(s.t_kernel for s in p_stats) / len(p_stats) * 1e3)
        log("plain kernel: %.0f reads/s" % plain["reads_per_s"])
        # leave the batched results in place for the parity check below
        batches[args.warmup][1].run(args.kernel)
    e2e = None
    if not args.no_graph:
        # the whole pipeline of one batch, host memory to host memory: the
        # drop-in call (gs_seed: upload, pipeline, download, one call each
        # time) against replays of the batch's captured HIP graph (upload from
        # pinned copies, pipeline, download into pinned memory, one launch);
        # the graph's results are the ones the parity check below reads.  The
        # resident batches are freed before the drop-in calls, whose own
        # device buffers would not fit beside them for configuration 4.
        fa, fb = batches[args.warmup]
        fb.capture(args.kernel)
        fb.replay()
        graph = []
        for _ in range(3):
            t = time.perf_counter()
            fb.replay()
            graph.append(time.perf_counter() - t)

    seed_s = sum(s.t_seed for s in stats) / len(stats)
    # seed_persistent's launches alone, by HIP events on the batch's stream
    # right around each launch (seed_s also holds code packing, the work
    # queue's set-up and the join of split streams)
    kern_s = sum(s.t_kernel for s in stats) / len(stats)
    if kern_s <= 0:  # no kernel events (an empty batch): the stage's time
        kern_s = seed_s
    first_arrays, first_batch = batches[args.warmup]
    gpu_out = first_batch.fetch()

    # the same resident batches on the reference's own index shape: the
    # 13-mer table only and no direct-extension structures, so positions
    # come from the sampled-suffix-array walk (sampling 7) inside the timed
    # region; the output must be the same bit for bit
    ref_shape = None
    if not args.no_reference_shape and (info.kmer_len > 13 or direct):
        X.set_pairs(False)
        X.set_kmer_len(13)
        X.set_direct(False)
        rinfo = X.info()
        r_elapsed, r_stats = timed_steps(batches, min(1, args.warmup), args.kernel)
        r_steps = len(batches) - min(1, args.warmup)
        r_out = first_batch.fetch()
        same = (np.array_equal(r_out.seeds, gpu_out.seeds) and np.array_equal(r_out.pos, gpu_out.pos) and
                np.array_equal(r_out.read_seed_off, gpu_out.read_seed_off))
        ref_shape = {"kmer_table": 13, "direct_extension": False, "device_bytes": int(rinfo.device_bytes),
                     "reads_per_s": r_steps * args.batch / r_elapsed,
                     "kernel_ms": {"seed": sum(x.t_seed for x in r_stats) / len(r_stats) * 1e3,
                                   "count": sum(x.t_count for x in r_stats) / len(r_stats) * 1e3,
                                   "write": sum(x.t_write for x in r_stats) / len(r_stats) * 1e3,
                                   "locate": sum(x.t_locate for x in r_stats) / len(r_stats) * 1e3},
                     "requests_per_read_issued": sum(x.n_table_reads + x.n_occ_reads + x.n_direct_reads + x.n_tandem_reads
                                                     for x in r_stats) / (len(r_stats) * args.batch),
                     "same_output_as_headline": bool(same)}
        log("reference-shaped index (13-mer table, suffix-array walk): %.0f reads/s, seed %.1f ms, locate %.1f ms, "
            "%.2f GB on device, output %s" % (ref_shape["reads_per_s"], ref_shape["kernel_ms"]["seed"],
                                              ref_shape["kernel_ms"]["locate"], rinfo.device_bytes / 1e9,
                                              "identical" if same else "DIFFERENT"))
        del r_out
        if info.kmer_len > 13:
            X.set_kmer_len(info.kmer_len)
        if direct:
            X.set_direct(True)
        if pairs:
            X.set_pairs(True)

    # the extension kernel's own counters on exactly the reads the work
    # model counts (no work splitting, which the model does not describe)
    sample_counts = None
    if not args.no_cpu:
        sa = strided_sample(first_arrays, args.work_sample)
        sb = g.Batch(X, sa, k=k, smem=smem, max_occ=max_occ, seg_tasks=args.seg_tasks, split_tasks=g.NO_SPLIT,
                     kernel=args.kernel, tandem=tandem)
        st_s = sb.run(args.kernel)
        sb.close()
        sample_counts = dict(reads=len(sa[2]), table=st_s.n_table_reads, occ=st_s.n_occ_reads,
                             direct=st_s.n_direct_reads, tandem=st_s.n_tandem_reads)
    for _, bt in batches:
        bt.close()
    del batches
    if not args.no_graph:
        eager = []
        for _ in range(2):
            t = time.perf_counter()
            g.seed(X, arrays=first_arrays, k=k, smem=smem, max_occ=max_occ, kernel=args.kernel,
                   seg_tasks=args.seg_tasks)
            eager.append(time.perf_counter() - t)
        e2e = {"reads": args.batch, "drop_in_call_ms": min(eager) * 1e3, "graph_replay_ms": min(graph) * 1e3,
               "graph_reads_per_s": args.batch / min(graph)}
        log("one batch host to host: gs_seed %.1f ms, graph replay %.1f ms (%.0f reads/s)" %
            (e2e["drop_in_call_ms"], e2e["graph_replay_ms"], e2e["graph_reads_per_s"]))

    # random-access ceiling of this GPU (16-byte dependent requests, 8 GiB)
    rr = g.roofline_random(8 << 30, line_bytes=16, mode=1, inflight=1, grid=2048, iters=512, reps=3)
    r_max = rr.lines_per_s
    log("random-access ceiling: %.1f G requests/s" % (r_max / 1e9))

    cpu = None
    nproc, avail = host_cpus()
    if not args.no_cpu:
        # every CPU this process may use, at most the 16 host cores that go
        # with one GPU of the machine
        threads = args.cpu_threads or min(16, avail)
        cpu = cpu_baseline(prefix, first_arrays, gpu_out, args.cpu_sample, threads, k, max_occ, smem)
        cpu["threads"] = threads
        log("cpu baseline: %.0f reads/s on %d threads; parity %d/%d reads differ"
            % (cpu["reads_per_s"], threads, cpu["mismatches"], cpu["reads"]))
    wm = None
    if not args.no_cpu:
        wm = work_model(prefix, first_arrays, args.work_sample, threads, k, max_occ, info.kmer_len, smem,
                        direct_min=4 if direct else 0, prune_sat={1: 255, 2: 65535, 3: 15}.get(args.kernel, 0),
                        pair_probe=2 if pairs else 0, tandem=0 if args.no_tandem else 6)
        log("work model: %.1f table + %.1f occ + %.1f direct requests per read (%.0f bytes)" %
            (wm["table_per_read"], wm["occ_per_read"], wm["direct_per_read"], wm["bytes_per_read"]))
    serving = None
    if serve_ref is not None:
        serving = serve_bench(g, X, serve_ref, k, max_occ, args)
        del serve_ref
    X.close()
    cli = None
    if args.cli_reads:
        cli = cli_end_to_end(first_arrays, prefix, tmpdir, args.cli_reads, k, max_occ, info.kmer_len, direct, smem,
                             pairs)
    shutil.rmtree(tmpdir, ignore_errors=True)

    roofline = None
    extra = {}
    if cpu:
        extra["parity"] = {"reads_checked": cpu["reads"], "mismatches": cpu["mismatches"],
                           "bit_exact": cpu["bit_exact"]}
    if wm:
        # extension kernel, algorithmic: the device algorithm's table entries,
        # distinct 96-base occ entries per backward step and direct-extension
        # reads, counted by the oracle; W for the reference's procedure
        # (13-mer table, distance-1 pruning) is reported beside it
        req_per_read = wm["requests_per_read"]
        issued = sum(st.n_table_reads + st.n_occ_reads + st.n_direct_reads + st.n_tandem_reads
                     for st in stats) / (len(stats) * args.batch)
        prof = profiled_counters(args.config, args.kernel, info.kmer_len, direct, pairs, 0 if args.no_tandem else 6)
        traffic, traffic_src = profiled_traffic(prof, args.batch)
        achieved = wm["bytes_per_read"] * args.batch / kern_s
        roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": 8000.0, "unit": "GB/s",
                    "frac": achieved / 8e12, "traffic": traffic, "traffic_source": traffic_src,
                    "kernel": "seed_persistent", "kernel_ms": kern_s * 1e3,
                    "bytes_per_read": wm["bytes_per_read"], "work_sample_reads": wm["reads"]}
        # every request that misses L2 moves one 128-byte line from HBM: the
        # kernel's own request count over the timed launches times the L2
        # miss share rocprofv3 measured on this workload, against the
        # measured ceiling of dependent random 16-byte requests (all misses)
        ra = {"requests_per_read_model": req_per_read, "requests_per_read_issued": issued,
              "measured_ceiling_requests_per_s": r_max}
        if sample_counts:
            sc = sample_counts
            ra["requests_per_read_issued_on_model_sample"] = (sc["table"] + sc["occ"] + sc["direct"] + sc["tandem"]) / sc["reads"]
            ra["model_sample_reads"] = sc["reads"]
        if prof and "tcc_hit" in prof:
            miss = prof["tcc_miss"] / (prof["tcc_hit"] + prof["tcc_miss"])
            miss_rate = issued * args.batch * miss / kern_s
            ra.update({"l2_miss_share": miss, "l2_miss_source": prof["path"],
                       "achieved_l2_miss_requests_per_s": miss_rate, "frac": miss_rate / r_max})
            roofline["line_bytes"] = {"bytes_per_read": issued * miss * 128.0,
                                      "achieved": issued * miss * 128.0 * args.batch / kern_s / 1e9,
                                      "frac": issued * miss * 128.0 * args.batch / kern_s / 8e12,
                                      "unit": "GB/s", "note": "one 128-byte HBM line per request that misses L2"}
        roofline["random_access"] = ra
        if cpu:
            c = cpu["counters"]
            roofline["reference_procedure_requests_per_read"] = (c["w_tasks"] + c["w_blocks96"]) / cpu["reads"]
    mean_ms = elapsed / args.steps * 1e3
    out = {
        "metric": "SMEM reads/s (150bp,k=19,3.1Gbp) batched vs plain, bit-exact; p99 ms",
        "value": value,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "config %d: %s -k %d, %s reads, %.2f Gbp synthetic reference, "
                               "positions for intervals <= %d, batched persistent kernel"
                               % (args.config, "SMEM" if smem else "MEM", k, ("%d bp" % int(lens[0])) if lens.min() == lens.max() else
                                  "%d-%d bp (mean %.0f)" % (int(lens.min()), int(lens.max()), float(lens.mean())),
                                  rp.length / 1e9, max_occ),
                   "reference_bp": int(rp.length), "reads_per_step": args.batch, "k": k,
                   "mode": "SMEM" if smem else "MEM", "max_occ": max_occ, "sa_intv": 7, "kmer_table": info.kmer_len,
                   "direct_extension": bool(direct), "pair_table": bool(pairs), "device_bytes": shape["device_bytes"],
                   "kernel": "batched-%d" % args.kernel},
        "roofline": roofline,
        "cpu_baseline": None if not cpu else {
            "value": cpu["reads_per_s"], "unit": "reads/s", "cores": cpu["threads"], "kind": "port",
            "nproc": nproc, "cpus_available": avail, "cpu_model": cpu_model(),
            "sample": "first %d reads of the first timed batch, oracle with distance-1 pruning and the "
                      "13-mer table (table build excluded), %d threads, compared bit for bit with the GPU output"
                      % (cpu["reads"], cpu["threads"])},
        "index_build_s": t_build,
        "plain": None if not plain else {"reads_per_s": plain["reads_per_s"], "seed_kernel_ms": plain["seed_ms"],
                                         "batched_over_plain": value / plain["reads_per_s"]},
        "host_to_host": e2e,
        "reference_shaped_index": ref_shape,
        "cli_end_to_end": cli,
        "kernel_ms": {"seed": seed_s * 1e3, "seed_persistent": kern_s * 1e3,
                      "count": sum(s.t_count for s in stats) / len(stats) * 1e3,
                      "write": sum(s.t_write for s in stats) / len(stats) * 1e3,
                      "locate": sum(s.t_locate for s in stats) / len(stats) * 1e3},
    }
    if serving:
        out["serving"] = serving
        l80 = [lv for lv in serving["levels"] if lv["load"] == 0.8]
        if l80:
            out["p99_ms_at_80pct_load"] = l80[0]["batched"]["p99_ms"]
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
