"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text.

    python -m raytracer_3d_amd.profsum kt   DB [DB ...]     # kernel time table
    python -m raytracer_3d_amd.profsum pmc  DB COUNTER KERNEL_SUBSTR [--copy-bytes B]

``kt`` prints, per kernel, calls / average / min / max duration (us) and the
register and LDS footprint of the first dispatch.  ``pmc`` prints the
per-dispatch average of COUNTER for the kernels whose name contains
KERNEL_SUBSTR and for the copy kernel (rt6d::k_copy), and -- given the copy's
byte count -- the factor that maps the counter onto true bytes.
"""
from __future__ import annotations

import argparse
import sqlite3
import statistics
import sys


def _rows(db, sql):
    con = sqlite3.connect(db)
    try:
        cur = con.execute(sql)
        cols = [d[0] for d in cur.description]
        return [dict(zip(cols, r)) for r in cur.fetchall()]
    finally:
        con.close()


def kernel_table(db):
    rows = _rows(db, "select name, duration, vgpr_count, accum_vgpr_count, sgpr_count, lds_size, grid_x, "
                     "workgroup_x from kernels")
    by = {}
    for r in rows:
        by.setdefault(r["name"], []).append(r)
    out = []
    for name, rs in by.items():
        d = [r["duration"] / 1e3 for r in rs]
        f = rs[0]
        out.append(dict(name=name, calls=len(rs), avg_us=statistics.fmean(d), min_us=min(d), max_us=max(d),
                        total_us=sum(d), vgpr=f["vgpr_count"], agpr=f["accum_vgpr_count"], sgpr=f["sgpr_count"],
                        lds=f["lds_size"], grid=f["grid_x"], block=f["workgroup_x"]))
    out.sort(key=lambda r: -r["total_us"])
    return out


def counter_avg(db, counter, substr):
    rows = _rows(db, "select kernel_name, dispatch_id, counter_name, value from counters_collection")
    per = {}
    for r in rows:
        if r["counter_name"] != counter or substr not in r["kernel_name"]:
            continue
        per.setdefault(r["dispatch_id"], 0.0)
        per[r["dispatch_id"]] += r["value"]
    vals = list(per.values())
    return (statistics.fmean(vals) if vals else None), len(vals)


def short(name, width=100):
    return name if len(name) <= width else name[: width - 3] + "..."


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kt")
    k.add_argument("db", nargs="+")
    p = sub.add_parser("pmc")
    p.add_argument("db")
    p.add_argument("counter")
    p.add_argument("kernel")
    p.add_argument("--copy-bytes", type=float, default=0.0)
    a = ap.parse_args(argv)
    if a.cmd == "kt":
        for db in a.db:
            print(f"# {db}")
            print(f"{'calls':>5} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'vgpr':>4} {'sgpr':>4} {'lds':>6}  kernel")
            for r in kernel_table(db):
                print(f"{r['calls']:>5} {r['avg_us']:>10.1f} {r['min_us']:>10.1f} {r['max_us']:>10.1f} "
                      f"{r['vgpr']:>4} {r['sgpr']:>4} {r['lds']:>6}  {short(r['name'])}")
        return 0
    v, n = counter_avg(a.db, a.counter, a.kernel)
    c, nc = counter_avg(a.db, a.counter, "k_copy")
    print(f"{a.counter} (KiB per dispatch): kernel~'{a.kernel}' {v} over {n} dispatches; copy {c} over {nc}")
    if a.copy_bytes and c:
        f = a.copy_bytes / (c * 1024.0)
        print(f"copy of {a.copy_bytes:.0f} bytes -> scale {f:.4f}; kernel bytes per dispatch {v * 1024.0 * f:.4e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
