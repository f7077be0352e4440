"""Text summaries of rocprofv3 CSV output (``-f csv``).

    python -m sipg.rocprof_summary stats DIR            # --kernel-trace --stats
    python -m sipg.rocprof_summary pmc DIR KERNEL_RE    # --pmc runs

``stats`` prints the per-kernel table of every ``*_kernel_stats.csv``
under DIR (or, without CSV output, of the ``*_results.db`` databases).  ``pmc`` averages each counter over the dispatches of the
kernels whose name matches KERNEL_RE, read from every
``*_counter_collection.csv`` under DIR.
"""
import csv
import glob
import os
import re
import sys
from collections import defaultdict


def _rows(pattern, root):
    for f in sorted(glob.glob(os.path.join(root, "**", pattern), recursive=True)):
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError(f"none of {names} in {sorted(row)}")


def _db_stats(root):
    """Per-kernel rows (name, calls, avg/min/max ns, total ns) from the
    rocpd databases (``*_results.db``, rocprofv3's default output) under
    root, most total time first."""
    import sqlite3
    agg = {}
    for f in sorted(glob.glob(os.path.join(root, "**", "*_results.db"), recursive=True)):
        con = sqlite3.connect(f)
        for name, calls, mn, mx, tot in con.execute(
                "select name, count(*), min(duration), max(duration), sum(duration) from kernels group by name"):
            a = agg.setdefault(name, [0, float("inf"), 0.0, 0.0])
            a[0] += calls
            a[1] = min(a[1], mn)
            a[2] = max(a[2], mx)
            a[3] += tot
        con.close()
    return sorted(((n, c, t / c, mn, mx, t) for n, (c, mn, mx, t) in agg.items()), key=lambda r: -r[5])


def stats(root):
    rows = list(_rows("*kernel_stats.csv", root))
    if not rows:
        db = _db_stats(root)
        if not db:
            sys.exit(f"no kernel_stats.csv or *_results.db under {root}")
        total = sum(r[5] for r in db)
        print(f"{'calls':>7} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'total_ms':>10} {'pct':>6}  kernel")
        for name, calls, avg, mn, mx, tot in db:
            print(f"{calls:7d} {avg / 1e3:10.2f} {mn / 1e3:10.2f} {mx / 1e3:10.2f} {tot / 1e6:10.3f} "
                  f"{100.0 * tot / total:6.2f}  {name[:150]}")
        return
    print(f"{'calls':>7} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'total_ms':>10} {'pct':>6}  kernel")
    for r in rows:
        calls = int(_col(r, "Calls"))
        avg = float(_col(r, "AverageNs")) / 1e3
        mn = float(_col(r, "MinNs")) / 1e3
        mx = float(_col(r, "MaxNs")) / 1e3
        tot = float(_col(r, "TotalDurationNs")) / 1e6
        pct = float(_col(r, "Percentage"))
        print(f"{calls:7d} {avg:10.2f} {mn:10.2f} {mx:10.2f} {tot:10.3f} {pct:6.2f}  {_col(r, 'Name')[:150]}")


def pmc_means(root, kernel_re):
    """{counter: mean value per dispatch} over the dispatches of the kernels
    matching kernel_re, and the set of matching kernel names."""
    rx = re.compile(kernel_re)
    vals = defaultdict(list)
    names = set()
    for r in _rows("*counter_collection.csv", root):
        name = _col(r, "Kernel_Name", "Kernel-Name", "KernelName")
        if not rx.search(name):
            continue
        names.add(name)
        vals[_col(r, "Counter_Name", "Counter-Name")].append(float(_col(r, "Counter_Value", "Counter-Value")))
    return {c: (sum(v) / len(v), len(v)) for c, v in vals.items()}, names


def pmc(root, kernel_re):
    means, names = pmc_means(root, kernel_re)
    if not means:
        sys.exit(f"no dispatch of /{kernel_re}/ under {root}")
    for n in sorted(names):
        print(f"kernel: {n[:150]}")
    for c, (m, k) in sorted(means.items()):
        print(f"{c:24s} mean {m:16.1f}  dispatches {k}")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3])
    else:
        sys.exit(__doc__)
