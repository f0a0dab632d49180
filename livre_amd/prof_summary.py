"""Summarise a rocprofv3 rocpd database (…_results.db) as text: per-kernel
dispatch statistics from the kernel trace and per-kernel averages of any
collected PMC counters.

    python3 livre_amd/prof_summary.py RESULTS.db [RESULTS.db ...] [--match raycast]
"""
import argparse
import sqlite3
from collections import defaultdict


def kernel_stats(db, match=None):
    c = sqlite3.connect(db)
    rows = c.execute("select name, duration, vgpr_count, sgpr_count, lds_size, "
                     "scratch_size, grid_x, workgroup_x from kernels").fetchall()
    agg = defaultdict(list)
    meta = {}
    for name, dur, vgpr, sgpr, lds, scratch, grid, wg in rows:
        if match and match not in name:
            continue
        agg[name].append(dur)
        meta[name] = (vgpr, sgpr, lds, scratch, grid, wg)
    out = []
    for name, d in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append((name, len(d), sum(d) / 1e6, sum(d) / len(d) / 1e3, min(d) / 1e3,
                    max(d) / 1e3) + meta[name])
    return out


def pmc_stats(db, match=None):
    c = sqlite3.connect(db)
    try:
        rows = c.execute("select name, counter_name, counter_value, dispatch_id "
                         "from pmc_events").fetchall()
    except sqlite3.Error:
        return []
    agg = defaultdict(lambda: defaultdict(dict))
    for name, counter, value, disp in rows:
        if match and match not in name:
            continue
        d = agg[name][counter]
        d[disp] = d.get(disp, 0.0) + float(value)
    out = []
    for name, counters in agg.items():
        for counter, per in counters.items():
            vals = list(per.values())
            out.append((name, counter, len(vals), sum(vals) / len(vals)))
    return sorted(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dbs", nargs="+")
    ap.add_argument("--match", default=None)
    a = ap.parse_args()
    for db in a.dbs:
        print("# %s" % db)
        ks = kernel_stats(db, a.match)
        if ks:
            print("%-60s %6s %10s %10s %10s %10s %5s %5s %7s %7s" % (
                "kernel", "calls", "total_ms", "avg_us", "min_us", "max_us", "vgpr",
                "sgpr", "lds", "scratch"))
            for k in ks:
                print("%-60s %6d %10.3f %10.2f %10.2f %10.2f %5d %5d %7d %7d" % (
                    k[0][:60], k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8], k[9]))
        ps = pmc_stats(db, a.match)
        if ps:
            print("%-60s %-16s %6s %18s" % ("kernel", "counter", "disp", "avg_per_dispatch"))
            for name, counter, n, avg in ps:
                print("%-60s %-16s %6d %18.1f" % (name[:60], counter, n, avg))
        print()


if __name__ == "__main__":
    main()
