"""Summarise a rocprofv3 SQLite result (``-d DIR -o run``) as a per-kernel table.

    python profiles/summarize.py gpurun_out/<dir>/run_results.db > profiles/<name>.md

Reads the ``kernels`` view (kernel-trace) and, when present, the ``counters``
view (``--pmc``), and prints Markdown: calls, average and total duration per
kernel, and per-dispatch averages of every collected counter.
"""
import sqlite3
import sys
from collections import defaultdict


def main(path):
    db = sqlite3.connect(path)
    cur = db.cursor()
    views = {r[0] for r in cur.execute("select name from sqlite_master where type in ('table','view')")}
    if "kernels" in views:
        cols = [r[1] for r in cur.execute("pragma table_info(kernels)")]
        ki = cols.index("name") if "name" in cols else cols.index("kernel_name")
        si, ei = cols.index("start"), cols.index("end")
        d = defaultdict(list)
        for r in cur.execute("select * from kernels"):
            d[r[ki]].append((r[ei] - r[si]) / 1e3)
        print("| kernel | calls | avg us | total us |\n|---|---:|---:|---:|")
        for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
            name = k.replace("(anonymous namespace)::", "").split("(")[0]
            print(f"| `{name}` | {len(v)} | {sum(v) / len(v):.1f} | {sum(v):.1f} |")
    view = next((v for v in ("counters", "pmc_events") if v in views), None)
    if view is not None:
        cols = [r[1] for r in cur.execute(f"pragma table_info({view})")]

        def col(*names):
            return next(cols.index(n) for n in names if n in cols)

        ki, ci, vi = col("kernel_name", "name"), col("counter_name"), col("value", "counter_value")
        di = col("dispatch_id", "id")
        agg = defaultdict(lambda: defaultdict(float))
        nd = defaultdict(set)
        for r in cur.execute(f"select * from {view}"):
            agg[r[ki]][r[ci]] += r[vi]
            nd[r[ki]].add(r[di])
        print()
        for k, dd in agg.items():
            name = k.replace("(anonymous namespace)::", "").split("(")[0]
            n = len(nd[k])
            print(f"\n`{name}` ({n} dispatches, per-dispatch averages)\n\n| counter | value |\n|---|---:|")
            for c, v in sorted(dd.items()):
                print(f"| {c} | {v / n:.4g} |")


if __name__ == "__main__":
    main(sys.argv[1])
