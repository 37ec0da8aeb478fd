"""Summarise a rocprofv3 results database (kernel trace / PMC) as text.

Usage: python -m cpipe_amd.profsum RESULTS.db [--pmc]
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def _short(name: str) -> str:
    return name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]


def kernel_stats(db: str) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    # rocprofv3 stores durations in microseconds in this view
    out = [f"{'kernel':<60} {'calls':>6} {'total_ms':>10} {'avg_ms':>9} {'pct':>6}"]
    for name, n, tot, avg, pct in rows:
        out.append(f"{_short(name)[:60]:<60} {n:>6} {tot / 1e3:>10.3f} {avg / 1e3:>9.3f} {pct:>6.2f}")
    return "\n".join(out)


def pmc(db: str) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, dispatch_id, counter_name, sum(value) from counters_collection "
                     "group by dispatch_id, counter_name").fetchall()
    per = defaultdict(lambda: defaultdict(list))
    for k, _, cn, v in rows:
        per[_short(k)][cn].append(v)
    out = []
    for k, d in per.items():
        vals = ", ".join(f"{cn}={sum(v) / len(v):.4g}" for cn, v in sorted(d.items()))
        out.append(f"{k[:60]}: {vals}")
    return "\n".join(out)


if __name__ == "__main__":
    db = sys.argv[1]
    print(pmc(db) if "--pmc" in sys.argv else kernel_stats(db))
