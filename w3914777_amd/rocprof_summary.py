"""Per-kernel summary of a rocprofv3 output directory (its SQLite database).

    python -m w3914777_amd.rocprof_summary <rocprofv3 -d directory> [--kernel SUBSTRING ...] [--last N]

Prints one line per kernel: dispatches, total and average duration (kernel
trace), and the mean of every collected counter per dispatch (--pmc runs).
Kernel names are shortened to their template signature.
"""
import argparse
import collections
import glob
import os
import sqlite3
import sys


def short(name):
    return name.split("(")[0]


def load(directory):
    """Per kernel: durations (ns) and per-counter values, in dispatch order."""
    dbs = sorted(glob.glob(os.path.join(directory, "**", "*.db"), recursive=True))
    if not dbs:
        raise SystemExit(f"no rocprofv3 database under {directory}")
    con = sqlite3.connect(dbs[0])
    kernels, counters = collections.defaultdict(list), collections.defaultdict(list)
    try:
        for n, dur in con.execute("select name, duration from kernels order by dispatch_id"):
            kernels[short(n)].append(dur)
    except sqlite3.Error:
        pass
    try:
        for n, c, v in con.execute("select kernel_name, counter_name, value from counters_collection order by dispatch_id"):
            counters[(short(n), c)].append(v)
    except sqlite3.Error:
        pass
    return kernels, counters


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("directory")
    ap.add_argument("--kernel", action="append", default=[], help="only kernels whose name contains this")
    ap.add_argument("--last", type=int, default=0,
                    help="summarise only each kernel's last N dispatches (e.g. the timed windows of a run)")
    args = ap.parse_args(argv)
    kernels, counters = load(args.directory)
    if args.last > 0:
        kernels = {k: v[-args.last:] for k, v in kernels.items()}
        counters = {k: v[-args.last:] for k, v in counters.items()}
    keep = (lambda k: any(s in k for s in args.kernel)) if args.kernel else (lambda k: True)
    if kernels:
        print(f"{'kernel':<58} {'calls':>8} {'total_ms':>10} {'avg_us':>10} {'min_us':>9} {'max_us':>9}")
        for k, v in sorted(kernels.items(), key=lambda kv: -sum(kv[1])):
            if keep(k):
                print(f"{k[:58]:<58} {len(v):>8} {sum(v) / 1e6:>10.2f} {sum(v) / len(v) / 1e3:>10.2f} "
                      f"{min(v) / 1e3:>9.2f} {max(v) / 1e3:>9.2f}")
    if counters:
        print(f"{'kernel':<58} {'counter':<22} {'dispatches':>10} {'mean':>14} {'max':>14} {'sum':>14}")
        for (k, c), v in sorted(counters.items()):
            if keep(k):
                print(f"{k[:58]:<58} {c:<22} {len(v):>10} {sum(v) / len(v):>14.6g} {max(v):>14.6g} {sum(v):>14.6g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
