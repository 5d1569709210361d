"""Summarises rocprofv3 result databases (rocpd SQLite) as short text.

    python3 rocpd_summary.py DB [DB ...]
    python3 rocpd_summary.py --by-waves KERNEL_SUBSTRING --simds N DB

Per kernel: dispatch count and total / average / min / max duration; when the database
holds counter samples (rocprofv3 --pmc), the per-kernel sum and per-dispatch average of
each counter.  With --by-waves: the dispatches of the kernels whose names hold the
substring, grouped by waves per SIMD (the grid's 64-lane waves over N SIMDs: up to 1, 1 to
2, ...), with each group's count and total / average duration.
"""
import sqlite3
import sys


def _cols(c, view):
    try:
        return [r[1] for r in c.execute("pragma table_info(%s)" % view)]
    except sqlite3.Error:
        return []


def summarize(path):
    c = sqlite3.connect(path)
    out = ["# %s" % path, "kernel | dispatches | total_ms | avg_us | min_us | max_us"]
    for name, n, tot, avg, mn, mx in c.execute(
            "select name, count(*), sum(end-start)/1e6, avg(end-start)/1e3, min(end-start)/1e3, "
            "max(end-start)/1e3 from kernels group by name order by 3 desc"):
        out.append("%s | %d | %.3f | %.1f | %.1f | %.1f" % (name, n, tot, avg, mn, mx))
    cols = _cols(c, "counters_collection")
    if cols:
        name_col = "kernel_name" if "kernel_name" in cols else ("name" if "name" in cols else None)
        ctr_col = "counter_name" if "counter_name" in cols else None
        val_col = "value" if "value" in cols else ("counter_value" if "counter_value" in cols else None)
        if name_col and ctr_col and val_col:
            out.append("counter | kernel | dispatches | sum | avg_per_dispatch")
            q = ("select %s, %s, count(*), sum(%s), avg(%s) from counters_collection group by 1, 2 order by 1, 4 desc"
                 % (ctr_col, name_col, val_col, val_col))
            for ctr, name, n, tot, avg in c.execute(q):
                out.append("%s | %s | %d | %.6g | %.6g" % (ctr, name, n, tot, avg))
        else:
            out.append("counters_collection columns: %s" % ", ".join(cols))
    return "\n".join(out)


def by_waves(path, sub, simds):
    c = sqlite3.connect(path)
    cols = _cols(c, "kernels")
    if "grid_x" not in cols:
        return "# %s: no grid sizes (kernels columns: %s)" % (path, ", ".join(cols))
    groups = {}
    for name, dur, gx, gy, gz in c.execute("select name, end-start, grid_x, grid_y, grid_z from kernels"):
        if sub not in name:
            continue
        waves = (gx * gy * gz + 63) // 64
        key = max(1, -(-waves // simds))  # waves per SIMD, rounded up
        g = groups.setdefault(key, [0, 0, 0])
        g[0] += 1
        g[1] += dur
        g[2] += waves
    out = ["# %s: dispatches of kernels named *%s*, by waves per SIMD (%d SIMDs)" % (path, sub, simds),
           "waves per SIMD | dispatches | avg waves | total_ms | avg_us"]
    for key in sorted(groups):
        n, tot, w = groups[key]
        out.append("%d-%d | %d | %.0f | %.3f | %.1f" % (key - 1, key, n, w / n, tot / 1e6, tot / n / 1e3))
    n = sum(g[0] for g in groups.values())
    tot = sum(g[1] for g in groups.values())
    out.append("all | %d | | %.3f | %.1f" % (n, tot / 1e6, tot / max(1, n) / 1e3))
    return "\n".join(out)


if __name__ == "__main__":
    a = sys.argv[1:]
    if a and a[0] == "--by-waves":
        if len(a) != 5 or a[2] != "--simds":
            sys.exit("usage: rocpd_summary.py --by-waves KERNEL_SUBSTRING --simds N DB")
        print(by_waves(a[4], a[1], int(a[3])))
    else:
        print("\n\n".join(summarize(p) for p in a))
