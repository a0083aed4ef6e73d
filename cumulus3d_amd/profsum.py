"""Summaries of rocprofv3 output databases (SQLite, ROCm 7): per-kernel launch counts and times of
a --kernel-trace run, and per-kernel means of the counters of a --pmc run.

  python -m cumulus3d_amd.profsum trace <dir>/run_results.db [top]
  python -m cumulus3d_amd.profsum pmc <dir>/run_results.db [kernel-substring]
"""
from __future__ import annotations

import collections
import sqlite3
import sys


def _q(db, sql):
    con = sqlite3.connect(db)
    try:
        return list(con.execute(sql))
    finally:
        con.close()


def _cols(db, table):
    return [r[1] for r in _q(db, f"pragma table_info('{table}')")]


def trace(db, top=15):
    rows = _q(db, "select name, count(*), sum(end - start), avg(end - start) from kernels group by name "
                  "order by sum(end - start) desc")
    tot = sum(r[2] for r in rows) or 1
    out = ["| kernel | launches | total ms | average ms | share |", "|---|---|---|---|---|"]
    for name, n, s, a in rows[:top]:
        out.append(f"| `{name[:90]}` | {n} | {s / 1e6:.3f} | {a / 1e6:.4f} | {100 * s / tot:.2f} % |")
    return "\n".join(out)


def pmc(db, match=""):
    c = _cols(db, "counters_collection")
    ki, ni, vi = c.index("kernel_name"), c.index("counter_name"), c.index("value")
    agg = collections.defaultdict(list)
    for r in _q(db, "select * from counters_collection"):
        if match in r[ki]:
            agg[(r[ki][:90], r[ni])].append(r[vi])
    out = ["| kernel | counter | dispatches | mean per dispatch |", "|---|---|---|---|"]
    for (k, n), v in sorted(agg.items()):
        out.append(f"| `{k}` | {n} | {len(v)} | {sum(v) / len(v):.6g} |")
    return "\n".join(out)


def main():
    if len(sys.argv) < 3 or sys.argv[1] not in ("trace", "pmc"):
        print(__doc__, file=sys.stderr)
        sys.exit(2)
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 15))
    else:
        print(pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else ""))


if __name__ == "__main__":
    main()
