#!/usr/bin/env python3
"""Reduce rocprofv3 output directories to a short text summary.

    rocprof_summary.py DIR [--kernels REGEX] [--timeline KERNEL]

Finds *_kernel_stats.csv and *_counter_collection.csv under DIR and prints
the kernels matching REGEX (all kernels by default): for kernel statistics
the call count and average/total duration, for counters the per-dispatch
values averaged over the matching dispatches.  With --timeline, reads
*_kernel_trace.csv instead and reports, from the first dispatch of KERNEL
on: the span, the time any kernel runs (union of dispatch intervals), the
time KERNEL runs and the sum of its dispatch durations (their ratio is how
many of its dispatches overlap on average), its duration quantiles, and the
total time of each kernel.
"""
import argparse
import csv
import glob
import os
import re
from collections import defaultdict


def short(name, width=70):
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[:width - 3] + "..."


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--kernels", default=".")
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--timeline", default="")
    a = ap.parse_args()
    if a.timeline:
        timeline(a.dir, a.timeline, a.top)
        return
    pat = re.compile(a.kernels)
    for f in sorted(glob.glob(os.path.join(a.dir, "**", "*_kernel_stats.csv"), recursive=True)):
        rows = list(csv.DictReader(open(f)))
        rows.sort(key=lambda r: -float(r.get("TotalDurationNs", 0)))
        print("kernel statistics (%s)" % os.path.relpath(f, a.dir))
        print("%-72s %7s %14s %14s %7s" % ("kernel", "calls", "avg ns", "total ns", "pct"))
        for r in rows[:a.top]:
            if not pat.search(r["Name"]) and a.kernels != ".":
                continue
            print("%-72s %7s %14.0f %14.0f %7.2f" % (short(r["Name"]), r["Calls"], float(r["AverageNs"]),
                                                    float(r["TotalDurationNs"]), float(r["Percentage"])))
        print()
    for f in sorted(glob.glob(os.path.join(a.dir, "**", "*_counter_collection.csv"), recursive=True)):
        acc = defaultdict(lambda: defaultdict(list))
        meta = {}
        for r in csv.DictReader(open(f)):
            name = r.get("Kernel_Name", "")
            if not pat.search(name):
                continue
            acc[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
            meta[name] = (r.get("Grid_Size", ""), r.get("Workgroup_Size", ""), r.get("LDS_Block_Size", ""),
                          r.get("VGPR_Count", r.get("Arch_VGPR_Count", "")))
        print("counters (%s), mean over dispatches" % os.path.relpath(f, a.dir))
        for name, cs in acc.items():
            g, w, lds, v = meta[name]
            print("  %s  grid=%s wg=%s lds=%s vgpr=%s" % (short(name, 90), g, w, lds, v))
            for c, vals in sorted(cs.items()):
                print("    %-28s %18.1f  (%d dispatches)" % (c, sum(vals) / len(vals), len(vals)))
        print()


def union(intervals):
    total, cur = 0, None
    for s, e in sorted(intervals):
        if cur is None or s > cur[1]:
            if cur:
                total += cur[1] - cur[0]
            cur = [s, e]
        else:
            cur[1] = max(cur[1], e)
    return total + (cur[1] - cur[0] if cur else 0)


def timeline(d, kernel, top):
    rows = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*_kernel_trace.csv"), recursive=True)):
        rows += [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in csv.DictReader(open(f))]
    pat = re.compile(kernel)
    mine = [(s, e) for s, e, n in rows if pat.search(n)]
    if not mine:
        print("no dispatch of %s" % kernel)
        return
    t0 = min(s for s, _ in mine)
    rows = [r for r in rows if r[0] >= t0]
    span = max(e for _, e, _ in rows) - t0
    busy = union([(s, e) for s, e, _ in rows])
    kb = union(mine)
    ksum = sum(e - s for s, e in mine)
    dur = sorted((e - s) / 1e6 for s, e in mine)
    q = lambda p: dur[min(len(dur) - 1, int(p * len(dur)))]
    print("timeline from the first dispatch of /%s/: span %.1f ms" % (kernel, span / 1e6))
    print("  any kernel running    %9.1f ms (%.0f %% of the span)" % (busy / 1e6, 100.0 * busy / span))
    print("  %-20s  %9.1f ms (%.0f %%), %d dispatches, durations summed %.1f ms (overlap %.2f)"
          % (kernel, kb / 1e6, 100.0 * kb / span, len(mine), ksum / 1e6, ksum / max(kb, 1)))
    print("  dispatch duration ms: median %.2f, p90 %.2f, max %.2f" % (q(0.5), q(0.9), dur[-1]))
    tot, cnt = defaultdict(int), defaultdict(int)
    for s, e, n in rows:
        tot[short(n, 60)] += e - s
        cnt[short(n, 60)] += 1
    for n, t in sorted(tot.items(), key=lambda kv: -kv[1])[:top]:
        print("  %-62s %6d calls %10.1f ms" % (n, cnt[n], t / 1e6))


if __name__ == "__main__":
    main()
