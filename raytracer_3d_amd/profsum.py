"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text.

    python -m raytracer_3d_amd.profsum kt   DB [DB ...]     # kernel time table
    python -m raytracer_3d_amd.profsum pmc  DB COUNTER KERNEL_SUBSTR [--copy-bytes B]

``kt`` prints, per kernel, calls / average / min / max duration (us) and the
register and LDS footprint of the first dispatch.  ``pmc`` prints the
per-dispatch average of COUNTER for the kernels whose name contains
KERNEL_SUBSTR and for the copy kernel (rt6d::k_copy), and -- given the copy's
byte count -- the factor that maps the counter onto true bytes.

``step`` writes the summary of one step kernel from the runs of the same
command (--kernel-trace --stats; --pmc FETCH_SIZE TCC_EA0_RDREQ_DRAM_sum;
--pmc WRITE_SIZE TCC_EA0_WRREQ_DRAM_sum; optionally a pass of busy counters
SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_SALU
SQ_INSTS_SMEM GRBM_GUI_ACTIVE TA_BUSY_avr TD_TD_BUSY_sum): the kernel table,
the HBM bytes per dispatch -- each counter scaled by what it reads for the
copy kernel of known size, FETCH_SIZE/WRITE_SIZE and the EA requests to DRAM
side by side -- against the algorithmic bytes, the unit busy figures, and
optionally the entry of a traffic JSON file keyed by workload and tag:

    python -m raytracer_3d_amd.profsum step --kt KT.db --fetch F.db --write W.db [--busy B.db] \
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


def _scaled(db, counter, kernel, copy_bytes):
    """(bytes per kernel dispatch, scale, raw per-dispatch value, dispatches):
    the counter's per-dispatch average mapped onto bytes by the copy kernel,
    which reads and writes copy_bytes."""
    v, n = counter_avg(db, counter, kernel)
    c, _ = counter_avg(db, counter, "k_copy")
    if v is None or not c:
        return None, None, v, n
    scale = copy_bytes / c
    return v * scale, scale, v, n


def counter_max(db, counter, substr):
    """Per-dispatch maximum of a counter over its instances (one per XCD for
    GRBM_GUI_ACTIVE: every XCD counts the dispatch's active cycles), averaged
    over the dispatches."""
    # pmc_events keeps one row per counter instance (counters_collection sums them)
    rows = _rows(db, "select name, dispatch_id, counter_name, counter_value from pmc_events")
    per = {}
    for r in rows:
        if r["counter_name"] != counter or substr not in r["name"]:
            continue
        per[r["dispatch_id"]] = max(per.get(r["dispatch_id"], 0.0), r["counter_value"])
    vals = list(per.values())
    return statistics.fmean(vals) if vals else None


def busy_figures(db, kernel, cus):
    """Unit activity of the kernel's dispatches from the busy pass, as
    rocprofv3's derived metrics define it: a unit's busy cycles over the
    dispatch's active cycles, GRBM_GUI_ACTIVE taken as its maximum over the
    XCDs (reduce(GRBM_GUI_ACTIVE, max)), not their sum."""
    g = {c: counter_avg(db, c, kernel)[0] for c in ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES",
                                                     "SQ_BUSY_CYCLES", "SQ_INSTS_SALU", "SQ_INSTS_SMEM",
                                                     "TA_BUSY_avr", "TD_TD_BUSY_sum")}
    g["GRBM_GUI_ACTIVE"] = counter_max(db, "GRBM_GUI_ACTIVE", kernel)
    out = {"raw": g}
    gui = g["GRBM_GUI_ACTIVE"]
    if gui:
        if g["SQ_ACTIVE_INST_VALU"] is not None:
            out["valu_busy"] = g["SQ_ACTIVE_INST_VALU"] / cus / gui
        if g["TA_BUSY_avr"] is not None:
            out["ta_busy"] = g["TA_BUSY_avr"] / gui
        if g["TD_TD_BUSY_sum"] is not None:
            out["td_busy"] = g["TD_TD_BUSY_sum"] / cus / gui
    return out


def step_summary(a) -> str:
    kt = kernel_table(a.kt)
    mine = [r for r in kt if a.kernel in r["name"]]
    if not mine:
        raise SystemExit(f"no kernel matching {a.kernel!r} in {a.kt}")
    k = mine[0]
    # FETCH_SIZE / WRITE_SIZE are in KiB; the copy kernel reads and writes copy_bytes
    fetch_b, fscale, fv, fn = _scaled(a.fetch, "FETCH_SIZE", a.kernel, a.copy_bytes / 1024.0)
    write_b, wscale, wv, wn = _scaled(a.write, "WRITE_SIZE", a.kernel, a.copy_bytes / 1024.0)
    fetch_b, write_b = fetch_b * 1024.0, write_b * 1024.0
    # EA requests destined for DRAM, mapped onto bytes by the same copy
    rd_b, rscale, rv, _ = _scaled(a.fetch, "TCC_EA0_RDREQ_DRAM_sum", a.kernel, a.copy_bytes)
    wr_b, wrscale, wrv, _ = _scaled(a.write, "TCC_EA0_WRREQ_DRAM_sum", a.kernel, a.copy_bytes)
    alg = a.cells * a.bytes_per_cell
    us = k["avg_us"]
    lines = [f"# {a.title}", "", "Commands (each its own run; summaries by python -m raytracer_3d_amd.profsum step ...):"]
    lines += [f"  {c}" for c in a.command]
    lines += ["", "Kernel trace (per dispatch):", f"{'calls':>5} {'avg_us':>9} {'min_us':>9} {'max_us':>9} {'vgpr':>4} "
              f"{'sgpr':>4} {'lds':>6}  kernel"]
    for r in kt[: a.top]:
        lines.append(f"{r['calls']:>5} {r['avg_us']:>9.1f} {r['min_us']:>9.1f} {r['max_us']:>9.1f} {r['vgpr']:>4} "
                     f"{r['sgpr']:>4} {r['lds']:>6}  {short(r['name'])}")
    lines += ["", "HBM bytes per dispatch of the step kernel (each counter scaled by what it reads for",
              f"rt6d::k_copy moving {a.copy_bytes:.0f} bytes each way):",
              f"  FETCH_SIZE {fv:.0f} KiB over {fn} dispatches, scale {fscale:.3f} -> {fetch_b / 1e6:.1f} MB",
              f"  WRITE_SIZE {wv:.0f} KiB over {wn} dispatches, scale {wscale:.3f} -> {write_b / 1e6:.1f} MB"]
    if rd_b is not None and wr_b is not None:
        lines += [f"  TCC_EA0_RDREQ_DRAM {rv:.0f} requests, {rscale:.1f} B each -> {rd_b / 1e6:.1f} MB",
                  f"  TCC_EA0_WRREQ_DRAM {wrv:.0f} requests, {wrscale:.1f} B each -> {wr_b / 1e6:.1f} MB"]
    lines += [f"  traffic (FETCH+WRITE) {(fetch_b + write_b) / 1e6:.1f} MB per launch"
              + (f", DRAM requests {(rd_b + wr_b) / 1e6:.1f} MB" if rd_b is not None and wr_b is not None else ""),
              f"  algorithmic {a.bytes_per_cell:g} B/cell x {a.cells} = {alg / 1e6:.1f} MB; traffic / algorithmic "
              f"{(fetch_b + write_b) / alg:.2f}x" + (f" (DRAM {(rd_b + wr_b) / alg:.2f}x)" if rd_b is not None and
                                                      wr_b is not None else ""),
              f"  algorithmic rate at the traced average {us:.1f} us: {alg / (us * 1e3):.0f} GB/s "
              f"({alg / (us * 1e3) / 80.0:.1f}% of 8 TB/s); HBM rate {(fetch_b + write_b) / (us * 1e3):.0f} GB/s"]
    copy_us = a.copy_us or next((r["avg_us"] for r in kt if "k_copy" in r["name"]), 0.0)
    if copy_us:
        cb = 2 * a.copy_bytes / (copy_us * 1e3)
        lines.append(f"  copy kernel {copy_us:.1f} us -> {cb:.0f} GB/s moved; the step's algorithmic rate is "
                     f"{alg / (us * 1e3) / cb * 100:.1f}% of that ceiling")
    busy = busy_figures(a.busy, a.kernel, a.cus) if a.busy else None
    if busy:
        g = busy["raw"]
        lines += ["", f"Unit activity per dispatch ({a.cus} CUs, GRBM_GUI_ACTIVE {g['GRBM_GUI_ACTIVE']:.0f} cycles, "
                      "the maximum over the XCDs):"]
        for key, label in (("valu_busy", "VALU busy (SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE, rocprofv3's VALUBusy)"),
                           ("ta_busy", "TA busy (TA_BUSY_avr / GRBM_GUI_ACTIVE)"),
                           ("td_busy", "TD busy (TD_TD_BUSY_sum / CUs / GRBM_GUI_ACTIVE)")):
            if key in busy:
                lines.append(f"  {label}: {100 * busy[key]:.1f}%")
        lines.append("  " + ", ".join(f"{c} {v:.4g}" for c, v in g.items() if v is not None))
    if a.json:
        import json
        import os
        t = {}
        if os.path.exists(a.json):
            with open(a.json) as f:
                t = json.load(f)
        ent = {"kernel": k["name"].split("(")[0].replace("void ", "").replace("rt6d::", ""),
               "bytes_per_launch": fetch_b + write_b, "fetch_bytes": fetch_b, "write_bytes": write_b,
               "fetch_scale": fscale, "write_scale": wscale, "algorithmic_bytes": float(alg),
               "avg_launch_us": us, "vgpr": k["vgpr"], "lds": k["lds"]}
        if rd_b is not None and wr_b is not None:
            ent.update(dram_bytes_per_launch=rd_b + wr_b, dram_read_bytes=rd_b, dram_write_bytes=wr_b,
                       dram_read_bytes_per_request=rscale, dram_write_bytes_per_request=wrscale)
        if busy:
            ent.update({k2: v for k2, v in busy.items() if k2 != "raw"})
        t.setdefault(a.workload, {})[a.tag] = ent
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
    st.add_argument("--copy-us", type=float, default=0.0,
                    help="average of the copy kernel (default: its row of the kernel trace)")
    st.add_argument("--busy", help="database of the busy-counter pass")
    st.add_argument("--cus", type=int, default=256)
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
