"""Summarise a rocprofv3 counter CSV (``*_counter_collection.csv`` from a
``--pmc`` run): per kernel name, grid and counter, the number of dispatches
and the mean, minimum and maximum over the last 40 dispatches (the steady
state after warm-up and autotuning).

    python tt-fort_amd/pmc_summary.py <counter_collection.csv>
"""
import collections
import csv
import sys


def main(path, last=40):
    vals = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            key = (r.get("Kernel_Name", "")[:70], r.get("Grid_Size", r.get("Grid_Size_X", "")), r["Counter_Name"])
            vals[key].append((int(r.get("Dispatch_Id", 0) or 0), float(r["Counter_Value"])))
    for k, v in sorted(vals.items(), key=lambda kv: -len(kv[1])):
        v.sort()
        tail = [x for _, x in v[-last:]]
        print("%-70s grid %-8s %-12s n=%4d mean_last%d=%.1f min=%.1f max=%.1f"
              % (k[0], k[1], k[2], len(v), last, sum(tail) / len(tail), min(tail), max(tail)))


if __name__ == "__main__":
    main(sys.argv[1])
