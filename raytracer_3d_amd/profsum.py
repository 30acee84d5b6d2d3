"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text.

    python -m raytracer_3d_amd.profsum kt   DB [DB ...]     # kernel time table
    python -m raytracer_3d_amd.profsum pmc  DB COUNTER KERNEL_SUBSTR [--copy-bytes B]

``kt`` prints, per kernel, calls / average / min / max duration (us) and the
register and LDS footprint of the first dispatch.  ``pmc`` prints the
per-dispatch average of COUNTER for the kernels whose name contains
KERNEL_SUBSTR and for the copy kernel (rt6d::k_copy), and -- given the copy's
byte count -- the factor that maps the counter onto true bytes.

``step`` writes the summary of one step kernel from three runs of the same
command (--kernel-trace --stats, --pmc FETCH_SIZE, --pmc WRITE_SIZE): the
kernel table, the HBM bytes per dispatch (each counter scaled by what it
reads for the copy kernel of known size) against the algorithmic bytes, and
optionally the entry of a traffic JSON file keyed by workload and tag:

    python -m raytracer_3d_amd.profsum step --kt KT.db --fetch F.db --write W.db \
        --kernel k_step_colx2 --cells 56623104 --bytes-per-cell 16 --copy-bytes 1073741824 \
        --title "..." --command "..." --out SUMMARY.txt --json traffic.json --workload W --tag T
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


def step_summary(a) -> str:
    kt = kernel_table(a.kt)
    mine = [r for r in kt if a.kernel in r["name"]]
    if not mine:
        raise SystemExit(f"no kernel matching {a.kernel!r} in {a.kt}")
    k = mine[0]
    fv, fn = counter_avg(a.fetch, "FETCH_SIZE", a.kernel)
    fc, _ = counter_avg(a.fetch, "FETCH_SIZE", "k_copy")
    wv, wn = counter_avg(a.write, "WRITE_SIZE", a.kernel)
    wc, _ = counter_avg(a.write, "WRITE_SIZE", "k_copy")
    fscale = a.copy_bytes / (fc * 1024.0)      # the copy reads copy_bytes
    wscale = a.copy_bytes / (wc * 1024.0)      # and writes copy_bytes
    fetch_b, write_b = fv * 1024.0 * fscale, wv * 1024.0 * wscale
    alg = a.cells * a.bytes_per_cell
    lines = [f"# {a.title}", "", "Commands (each its own run; summaries by python -m raytracer_3d_amd.profsum step ...):"]
    lines += [f"  {c}" for c in a.command]
    lines += ["", "Kernel trace (per dispatch):", f"{'calls':>5} {'avg_us':>9} {'min_us':>9} {'max_us':>9}  kernel"]
    for r in kt[: a.top]:
        lines.append(f"{r['calls']:>5} {r['avg_us']:>9.1f} {r['min_us']:>9.1f} {r['max_us']:>9.1f}  {short(r['name'])}")
    lines += ["", f"HBM bytes per dispatch of the step kernel (counters in KiB, scaled by the copy kernel:",
              f"rt6d::k_copy moving {a.copy_bytes:.0f} bytes reads {fc:.0f} KiB FETCH_SIZE -> scale {fscale:.3f}, "
              f"writes {wc:.0f} KiB WRITE_SIZE -> scale {wscale:.3f}):",
              f"  FETCH_SIZE {fv:.0f} KiB over {fn} dispatches -> {fetch_b / 1e6:.1f} MB",
              f"  WRITE_SIZE {wv:.0f} KiB over {wn} dispatches -> {write_b / 1e6:.1f} MB",
              f"  traffic {(fetch_b + write_b) / 1e6:.1f} MB per launch; algorithmic {a.bytes_per_cell:g} B/cell x "
              f"{a.cells} = {alg / 1e6:.1f} MB ({(fetch_b + write_b) / alg:.2f}x)",
              f"  algorithmic rate at the traced average {k['avg_us']:.1f} us: {alg / (k['avg_us'] * 1e3):.0f} GB/s "
              f"({alg / (k['avg_us'] * 1e3) / 80.0:.1f}% of 8 TB/s); HBM rate {(fetch_b + write_b) / (k['avg_us'] * 1e3):.0f} GB/s"]
    if a.json:
        import json
        import os
        t = {}
        if os.path.exists(a.json):
            with open(a.json) as f:
                t = json.load(f)
        t.setdefault(a.workload, {})[a.tag] = {
            "kernel": k["name"].split("(")[0].replace("void ", "").replace("rt6d::", ""),
            "bytes_per_launch": fetch_b + write_b, "fetch_bytes": fetch_b, "write_bytes": write_b,
            "fetch_scale": fscale, "write_scale": wscale, "algorithmic_bytes": float(alg),
            "avg_launch_us": k["avg_us"]}
        with open(a.json, "w") as f:
            json.dump(t, f, indent=1)
            f.write("\n")
    return "\n".join(lines) + "\n"


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
    st = sub.add_parser("step")
    st.add_argument("--kt", required=True)
    st.add_argument("--fetch", required=True)
    st.add_argument("--write", required=True)
    st.add_argument("--kernel", required=True)
    st.add_argument("--cells", type=int, required=True)
    st.add_argument("--bytes-per-cell", type=float, required=True)
    st.add_argument("--copy-bytes", type=float, required=True)
    st.add_argument("--title", required=True)
    st.add_argument("--command", action="append", default=[])
    st.add_argument("--top", type=int, default=8)
    st.add_argument("--out")
    st.add_argument("--json")
    st.add_argument("--workload", default="")
    st.add_argument("--tag", default="")
    a = ap.parse_args(argv)
    if a.cmd == "step":
        text = step_summary(a)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text)
        print(text, end="")
        return 0
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
