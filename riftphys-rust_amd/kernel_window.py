#!/usr/bin/env python3
"""Per-kernel summary of one window of a rocprofv3 trace.

The window is the dispatches between two groups of marker dispatches (the library's plain copy
kernel, launched by `bench.py --mark-window` just before and just after the pass whose roofline
the bench reports).  Input: a kernel-trace CSV (durations) or a counter-collection CSV from
`rocprofv3 --pmc` (counter values, summed per kernel over the window's dispatches).  Prints a
markdown table.

    python3 riftphys-rust_amd/kernel_window.py <prefix>_kernel_trace.csv [--ticks N] [--group K1,K2,...]
    python3 riftphys-rust_amd/kernel_window.py <prefix>_counter_collection.csv [--ticks N]

--group names kernels that run concurrently (on side streams) within one stage of a tick: the
trace summary then also gives the stage's wall span, first start to last end of each cluster of
overlapping dispatches of those kernels (one cluster per tick).
"""
import argparse
import csv
import re
import sys


def short(name):
    name = re.sub(r"\(.*$", "", name.strip())
    return re.sub(r"^void ", "", name)


def window(rows, marker):
    """rows sorted in dispatch order: (key, ...name at [-1]); returns those between the first two
    runs of consecutive marker dispatches."""
    runs, prev = [], None
    for i, r in enumerate(rows):
        if r[-1].startswith(marker):
            if prev is not None and i == prev + 1:
                runs[-1][1] = i
            else:
                runs.append([i, i])
            prev = i
    if len(runs) < 2:
        sys.exit("fewer than two groups of %s dispatches" % marker)
    return rows[runs[0][1] + 1: runs[1][0]]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trace")
    ap.add_argument("--marker", default="k_copy_probe")
    ap.add_argument("--ticks", type=int, default=0, help="ticks in the window (adds per-tick columns)")
    ap.add_argument("--group", default="", help="comma-separated kernels of one concurrent stage (adds its wall span)")
    a = ap.parse_args()
    with open(a.trace) as f:
        rd = csv.DictReader(f)
        fields = rd.fieldnames or []
        recs = list(rd)
    tk = a.ticks or 1
    if "Counter_Name" in fields:
        order = "Dispatch_Id" if "Dispatch_Id" in fields else "Correlation_Id"
        disp = {}
        for r in recs:
            d = int(r[order])
            disp.setdefault(d, [short(r["Kernel_Name"]), {}])
            disp[d][1][r["Counter_Name"]] = disp[d][1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        rows = [(d, v[1], v[0]) for d, v in sorted(disp.items())]
        win = window(rows, a.marker)
        names = sorted({c for _, cv, _ in win for c in cv})
        per = {}
        for _, cv, n in win:
            e = per.setdefault(n, [0, {c: 0.0 for c in names}])
            e[0] += 1
            for c, v in cv.items():
                e[1][c] += v
        print("Window: %d dispatches between %s markers; counters summed per kernel%s" % (
            len(win), a.marker, (" and per tick (%d ticks)" % a.ticks) if a.ticks else ""))
        print()
        print("| kernel | calls | " + " | ".join("%s%s" % (c, " per tick" if a.ticks else "") for c in names) + " |")
        print("|---|---:|" + "---:|" * len(names))
        for n, (cnt, cv) in sorted(per.items(), key=lambda kv: -sum(kv[1][1].values())):
            print("| `%s` | %d | " % (n, cnt) + " | ".join("%.1f" % (cv[c] / tk) for c in names) + " |")
        return
    rows = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"])) for r in recs)
    win = window(rows, a.marker)
    span_ms = (win[-1][1] - win[0][0]) * 1e-6
    per = {}
    for s, e, n in win:
        c, t = per.get(n, (0, 0))
        per[n] = (c + 1, t + (e - s))
    tot = sum(t for _, t in per.values())
    print("Window: %d dispatches between %s markers, wall span %.3f ms%s" % (
        len(win), a.marker, span_ms, (", %.3f ms per tick" % (span_ms / a.ticks)) if a.ticks else ""))
    print()
    print("| kernel | calls | total ms | avg ms |" + (" ms per tick |" if a.ticks else "") + " % of kernel time |")
    print("|---|---:|---:|---:|" + ("---:|" if a.ticks else "") + "---:|")
    for n, (c, t) in sorted(per.items(), key=lambda kv: -kv[1][1]):
        line = "| `%s` | %d | %.3f | %.4f |" % (n, c, t * 1e-6, t * 1e-6 / c)
        if a.ticks:
            line += " %.3f |" % (t * 1e-6 / a.ticks)
        print(line + " %.2f |" % (100.0 * t / tot))
    if a.group:
        names = set(a.group.split(","))
        clusters = []
        for st, en, n in win:
            if n not in names:
                continue
            if clusters and st <= clusters[-1][1]:
                clusters[-1][1] = max(clusters[-1][1], en)
                clusters[-1][2] += 1
            else:
                clusters.append([st, en, 1])
        spans = [(e - s_) * 1e-6 for s_, e, _ in clusters]
        if spans:
            print()
            print("Stage of %s: %d clusters of overlapping dispatches, wall span %.3f ms mean (min %.3f, max %.3f), %.3f ms total" % (
                " + ".join("`%s`" % n for n in sorted(names)), len(spans), sum(spans) / len(spans), min(spans), max(spans), sum(spans)))


if __name__ == "__main__":
    main()
