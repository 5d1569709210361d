"""Summarises rocprofv3 result databases (rocpd SQLite) as short text.

    python3 rocpd_summary.py DB [DB ...]

Per kernel: dispatch count and total / average / min / max duration; when the database
holds counter samples (rocprofv3 --pmc), the per-kernel sum and per-dispatch average of
each counter.
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


if __name__ == "__main__":
    print("\n\n".join(summarize(p) for p in sys.argv[1:]))
