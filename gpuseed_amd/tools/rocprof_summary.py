#!/usr/bin/env python3
"""Reduce rocprofv3 output directories to a short text summary.

    rocprof_summary.py DIR [--kernels REGEX]

Finds *_kernel_stats.csv and *_counter_collection.csv under DIR and prints
the kernels matching REGEX (all kernels by default): for kernel statistics
the call count and average/total duration, for counters the per-dispatch
values averaged over the matching dispatches.
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
    a = ap.parse_args()
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


if __name__ == "__main__":
    main()
