"""Summarise a rocprofv3 kernel trace (``*_kernel_trace.csv``) into a short
text report: dispatch count, busy fraction, idle gaps between dispatches, and
per-kernel call counts and durations; with a second argument, the same per
(kernel, grid) for kernel names containing it.

    python tt-fort_amd/trace_summary.py <kernel_trace.csv> [name-substring]
"""
import csv, sys, collections
path = sys.argv[1]
rows = []
with open(path) as f:
    for r in csv.DictReader(f):
        rows.append((int(r['Start_Timestamp']), int(r['End_Timestamp']), r['Kernel_Name'][:90],
                     r['Grid_Size_X'] + 'x' + r['Grid_Size_Y'] + 'x' + r['Grid_Size_Z']))
rows.sort()
t0 = rows[0][0]; t1 = max(r[1] for r in rows)
busy = sum(r[1] - r[0] for r in rows)
print("dispatches %d span_ms %.2f busy_ms %.2f (%.1f%%)" % (len(rows), (t1 - t0) / 1e6, busy / 1e6, 100 * busy / (t1 - t0)))
gaps = sorted(rows[i + 1][0] - rows[i][1] for i in range(len(rows) - 1))
n = len(gaps)
print("gap us percentiles 50/75/90/99:", [round(gaps[int(n * q)] / 1e3, 2) for q in (0.5, 0.75, 0.9, 0.99)])
print("gap total ms %.2f; gaps > 20us: %d totalling %.2f ms" % (sum(g for g in gaps if g > 0) / 1e6,
      sum(1 for g in gaps if g > 20000), sum(g for g in gaps if g > 20000) / 1e6))
# idle gaps over 20 us by the pair of kernels around them (a host
# synchronisation shows as a gap before the next launch)
pairs = collections.defaultdict(list)
for i in range(len(rows) - 1):
    g = rows[i + 1][0] - rows[i][1]
    if g > 20000:
        pairs[(rows[i][2][:44], rows[i + 1][2][:44])].append(g)
print("%-44s -> %-44s %7s %9s" % ("gap after", "before", "count", "total_ms"))
for k, v in sorted(pairs.items(), key=lambda kv: -sum(kv[1]))[:12]:
    print("%-44s -> %-44s %7d %9.2f" % (k[0], k[1], len(v), sum(v) / 1e6))
agg = collections.defaultdict(list)
for r in rows:
    agg[r[2]].append(r[1] - r[0])
print("%-90s %8s %10s %9s" % ("kernel", "calls", "total_ms", "avg_us"))
for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:40]:
    print("%-90s %8d %10.2f %9.2f" % (k, len(v), sum(v) / 1e6, sum(v) / len(v) / 1e3))
if len(sys.argv) > 2:
    # per (kernel, grid) detail for names containing the given substring
    sub = sys.argv[2]
    agg2 = collections.defaultdict(list)
    for r in rows:
        if sub in r[2]:
            agg2[(r[2], r[3])].append(r[1] - r[0])
    for k, v in sorted(agg2.items(), key=lambda kv: -sum(kv[1]))[:30]:
        print("%-70s %-16s %7d %9.2f %9.2f" % (k[0][:70], k[1], len(v), sum(v) / 1e6, sum(v) / len(v) / 1e3))
