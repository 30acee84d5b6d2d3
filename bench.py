#!/usr/bin/env python3
"""Throughput of the 6D wave-energy step on one MI355X.

Workload (default): the production 6D grid nx=ny=nz=24, nk=nl=nm=16
(56,623,104 cells) under N(z) = 5.2e-3 exp(z/1300) s^-1, U(z) = 0.2
tanh((z+500)/300) m/s, a Garrett-Munk-shaped spectrum with a seeded lognormal
(x,y) modulation, dt at CFL 0.5, binary64.  A "step" is one integrate
(for_src/integrate.f90:4-271) over the whole grid with E resident in HBM.

Prints ONE JSON line (rank 0).  ``value`` is 6D cell-updates/s of the plain
binary64 step as the library runs it on this input.  ``roofline`` is that
step kernel's algorithmic HBM bytes per launch over its average launch
duration (HIP events on the library's stream around the K timed launches)
against the 8 TB/s MI355X peak, and also against the copy ceiling measured
in the same run (the library's full-grid 16-byte copy kernel); ``traffic`` is
the HBM bytes per launch from the committed rocprofv3 counter passes.
``cpu_baseline`` times the reference Fortran (oracle/_ref, MPI ranks on the
host cores) on a bounded sample of the same workload.

Extras (skip with --no-extras): the FP32 form and the general path, the
32^3 x 16^3 grid, the FP32 error report of the 24^3 and 32^3 grids after 500
steps, the Fortran drop-in driver on the same workload, the 96-member
batched ensembles (every member also timed alone, same clock), and the
serving scenario (Poisson 40 req/s for 120 s, 200 GB budget).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SETTLE_SECONDS = 0.25
COPY_BYTES = 1 << 30


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _settle(m, dt, seconds=SETTLE_SECONDS):
    """Untimed steps for about `seconds` so that the GPU has left its idle
    clocks before the warm-up (a production run steps for minutes); returns
    the number of steps taken."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        m.integrate(10, dt=dt)
        m.synchronize()
        n += 10
    return n


def _time_steps(m, dt, steps, warmup):
    """W untimed steps, then EXACTLY K timed steps bracketed by device
    synchronisation; returns (wall seconds, event ms of the K launches)."""
    if warmup:
        m.integrate(warmup, dt=dt)
    m.synchronize()
    t0 = time.perf_counter()
    ev_ms = m.integrate_timed(steps, dt=dt)
    m.synchronize()
    t1 = time.perf_counter()
    return t1 - t0, ev_ms


def _pe_grid(n, npes):
    """A legal PE grid for the reference decomposition (parallel_mpi.f90:14-28):
    factors of 2 dealt round-robin over x, y (any split) and k, l (even
    splits only), so the blocks stay as cubic as the factors allow."""
    grid = [1] * 6
    order = [0, 1, 3, 4, 2, 5]
    left = npes
    stuck = 0
    i = 0
    while left > 1 and left % 2 == 0 and stuck < len(order):
        ax = order[i % len(order)]
        i += 1
        nb = grid[ax] * 2
        if n[ax] % nb == 0 and n[ax] // nb >= 2:
            grid[ax] = nb
            left //= 2
            stuck = 0
        else:
            stuck += 1
    return grid if left == 1 else None


def cpu_baseline(case, dt, steps, ranks=(16, 8, 4, 2), label="24^3x16^3"):
    """The reference Fortran (oracle/_ref/ref_harness) on the host cores,
    MPI ranks over a 6D PE grid (the first count of `ranks` that the host
    and the grid allow); returns the cpu_baseline object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from raytracer_3d_amd import cases

    if not os.path.exists(O.REF_BIN):
        return None
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    tries = []
    for npes in ranks:
        if npes <= ncpu:
            g = _pe_grid(case.n, npes)
            if g:
                tries.append((npes, g))
    tries.append((1, None))
    with tempfile.TemporaryDirectory() as tmp:
        cf = os.path.join(tmp, "bench.case")
        cases.write_case_file(case, cf, nsteps=steps, dt=dt)
        last_err = None
        for npes, g in tries:
            try:
                _log(f"cpu baseline: reference Fortran, {npes} MPI rank(s), grid {g}, {steps} steps")
                secs = O.run_reference(cf, "-", np_grid=g, timeout=900)
                return {
                    "value": case.cells * steps / secs,
                    "unit": "cell-updates/s",
                    "cores": npes,
                    "kind": "reference",
                    "sample": f"{steps} integrate steps of the same {label} binary64 workload, "
                              f"MPI_Wtime around the loop, {npes} rank(s) PE grid {g}, host nproc {ncpu}",
                    "seconds": secs,
                }
            except Exception as e:  # noqa: BLE001 -- try the next rank count
                last_err = str(e)[-300:]
                _log(f"cpu baseline with {npes} ranks failed: {last_err}")
    return {"value": None, "unit": "cell-updates/s", "cores": 0, "kind": "reference",
            "sample": "failed: " + (last_err or "")}


def _traffic(workload, tag):
    """The committed rocprofv3 counter figures of the step kernel
    (profiles/traffic.json): HBM bytes per launch from the EA requests to
    DRAM and from FETCH_SIZE + WRITE_SIZE, each scaled by the copy kernel."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get(workload, {}).get(tag)
    except Exception:  # noqa: BLE001
        return None


def _step_kernel(precision, option, uxy, handover=False):
    """Name of the kernel one step launches for a resolved kernel option
    (rt6d_query RT6D_Q_KERNEL), compiled for all advected axes active;
    handover: RT6D_Q_X_HANDOVER (the row-aligned wave-local kernel)."""
    T = precision
    # binary64's uniform-spacing wave-local kernels are the 128-VGPR entry
    # (eight template parameters); the others carry a ninth, SB: binary32's
    # wave-local kernel with two z cells per thread branches on face signs
    w4 = uxy and T == "double"
    k = "k_step_colx2_w4" if w4 else "k_step_colx2"

    def wl(zc):
        sb = "" if w4 else (", true" if (uxy and T == "float" and zc == 2) else ", false")
        return (f"{k}<{T}, false, {zc}, true, true, {'true' if uxy else 'false'}, false, "
                f"{'true' if handover else 'false'}{sb}>")

    names = {
        15: f"k_step_colx2<{T}, false, 1, true, false, false, false, false, false>",
        16: f"k_step_colx2<{T}, false, 2, true, false, false, false, false, false>",
        21: wl(1), 22: wl(2),
        18: "k_step_colq<false, 1, true>", 19: "k_step_colq<false, 2, true>",
        9: f"k_step_col<{T}, false, false, 1, true, 1>", 10: f"k_step_col<{T}, false, false, 2, true, 1>",
        4: f"k_step_col<{T}, false, true, 1, true, 0>", 5: f"k_step_col<{T}, false, true, 2, true, 0>",
    }
    return names.get(option, f"kernel option {option}")


def fortran_dropin(case, dt, steps, precision="fp64"):
    """The Fortran driver (fortran/rt6d_driver: the reference's set-up order
    and time loop over the drop-in integrate) on the same workload; its own
    system_clock around the loop."""
    drv = os.path.join(ROOT, "raytracer_3d_amd", "fortran", "rt6d_driver")
    if not os.path.exists(drv):
        return None
    from raytracer_3d_amd import cases
    with tempfile.TemporaryDirectory() as tmp:
        cf = os.path.join(tmp, "bench.case")
        cases.write_case_file(case, cf, nsteps=steps, dt=dt)
        cmd = [drv, cf] + (["fp32"] if precision == "fp32" else [])
        p = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        return {"error": (p.stdout + p.stderr)[-400:]}
    mt = re.search(r"main loop\s+([0-9.Ee+-]+)\s*s,\s*([0-9.Ee+-]+)\s*cell-updates/s", p.stdout)
    if not mt:
        return {"error": p.stdout[-400:]}
    secs = float(mt.group(1))
    return {"value": float(mt.group(2)), "seconds": secs, "steps": steps, "ms_per_step": 1e3 * secs / steps,
            "how": "rt6d_driver CASE: set_parameter .. calc_parameter, then int(runlen/dt) integrate calls "
                   "(main.f90:51-79), system_clock around the loop and the final synchronize"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=300)
    ap.add_argument("--serve-seconds", type=float, default=120.0,
                    help="arrival window of the serving extra (the scenario is 120 s; 0 skips it)")
    ap.add_argument("--error-steps", type=int, default=500, help="steps of the FP32 error report (0 skips it)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the 6D step runs one resident grid per GPU and this build has "
              "no multi-GPU path (a sharded step would split m into one slab per GPU with a 2-plane halo exchange); "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from raytracer_3d_amd import cases, rt6d

    _log("building the 24^3 x 16^3 workload")
    case = cases.gm_case()
    cells = case.cells
    workload = "gm_24x24x24_16x16x16"

    def run(precision, path, kernel=0, c=None):
        c = case if c is None else c
        m = rt6d.Model.from_case(c, precision=precision, path=path, kernel=kernel)
        m.calc_parameter()
        dt = m.cfl_dt(0.5)
        settle = _settle(m, dt)
        wall, ev_ms = _time_steps(m, dt, args.steps, args.warmup)
        res = {
            "value": c.cells * args.steps / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "launch_ms": ev_ms / args.steps,
            "bytes_per_cell": m.bytes_per_cell,
            "homogeneous": m.homogeneous,
            "kernel_option": m.kernel,
            "uniform_xy": m.query(rt6d.Q_UNIFORM_XY),
            "x_handover": m.query(rt6d.Q_X_HANDOVER),
            "dt": dt,
            "settle_steps": settle,
        }
        res["achieved_GBs"] = res["bytes_per_cell"] * c.cells / (res["launch_ms"] * 1e-3) / 1e9
        m.close()
        return res

    _log("copy ceiling: 1 GiB device copy, 16 B per thread")
    copy_ms = rt6d.device_copy_timed(COPY_BYTES, reps=20)
    copy_gbs = 2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9

    _log("binary64 step")
    main_res = run(rt6d.FP64, 0)
    extras = {}
    if not args.no_extras:
        _log("FP32 form")
        extras["fp32"] = run(rt6d.FP32, 0)
        extras["fp32"]["speedup_vs_fp64"] = extras["fp32"]["value"] / main_res["value"]
        _log("binary64 and FP32 steps, general path (6D coefficient fields in HBM)")
        extras["fp64_general"] = run(rt6d.FP64, 1)
        extras["fp32_general"] = run(rt6d.FP32, 1)
        _log("binary64 step, per-cell kernel")
        extras["fp64_percell"] = run(rt6d.FP64, 0, kernel=1)
        # the driver's loop starts on a cold clock (no settle phase): 3000
        # steps keep the clock ramp a small part of its time
        _log("Fortran drop-in driver, 3000 steps of the same workload")
        extras["fortran_dropin_fp64"] = fortran_dropin(case, main_res["dt"], 3000)
        extras["fortran_dropin_fp32"] = fortran_dropin(case, main_res["dt"], 3000, "fp32")
        if args.error_steps > 0:
            from raytracer_3d_amd import precision
            _log(f"FP32 error report, 24^3 x 16^3, {args.error_steps} steps")
            extras["fp32_error_24cube"] = precision.fp32_report(
                case, steps=args.error_steps, checkpoints=[1, 100, args.error_steps], log=_log)
        _log("32^3 x 16^3 (134M cells): binary64 step and FP32 form")
        big = cases.gm_case(n=(32, 32, 32, 16, 16, 16))
        extras["fp64_32cube"] = run(rt6d.FP64, 0, c=big)
        extras["fp32_32cube"] = run(rt6d.FP32, 0, c=big)
        extras["fp32_32cube"]["speedup_vs_fp64"] = extras["fp32_32cube"]["value"] / extras["fp64_32cube"]["value"]
        for k in ("fp64_32cube", "fp32_32cube"):
            extras[k]["cells"] = big.cells
        if args.error_steps > 0:
            _log(f"FP32 error report, 32^3 x 16^3, {args.error_steps} steps")
            extras["fp32_error_32cube"] = precision.fp32_report(
                big, steps=args.error_steps, checkpoints=[1, 100, args.error_steps], end_fields=False, log=_log)
        del big
        _log("config-0 packet 32x4x32x16x4x16 (4.2M cells): binary64 step, reference Fortran on 4 ranks")
        c0 = cases.packet_case()
        extras["config0_packet_fp64"] = run(rt6d.FP64, 0, c=c0)
        extras["config0_packet_fp64"]["cells"] = c0.cells
        if not args.no_cpu_baseline:
            extras["config0_packet_fp64"]["cpu_baseline_np4"] = cpu_baseline(
                c0, extras["config0_packet_fp64"]["dt"], 20, ranks=(4, 2), label="config-0 packet 32x4x32x16x4x16")
        from raytracer_3d_amd import serve
        for prec, tag in ((rt6d.FP64, "ensemble96_fp64"), (rt6d.FP32, "ensemble96_fp32")):
            _log(f"batched ensemble, 96 members x {args.ensemble_steps} steps ({tag[-4:]}), every member also alone")
            extras[tag] = serve.ensemble_benchmark(96, args.ensemble_steps, prec, log=_log)
        if args.serve_seconds > 0:
            _log(f"serving: Poisson 40 req/s for {args.serve_seconds:g} s, 200 GB budget")
            extras["serving"] = serve.serve_benchmark(40.0, args.serve_seconds, 200.0, baseline_seconds=10.0,
                                                      log=_log)

    kname = _step_kernel("double", main_res["kernel_option"], main_res["uniform_xy"] == 3, main_res["x_handover"] == 1)
    ttag = "fp64_auto" if main_res["homogeneous"] else "fp64_general"
    achieved = main_res["achieved_GBs"]
    tr = _traffic(workload, ttag)
    traffic = None
    if tr:
        traffic = tr.get("dram_bytes_per_launch", tr.get("bytes_per_launch"))
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": kname,
        "algorithmic_bytes_per_cell": main_res["bytes_per_cell"],
        "launch_ms": round(main_res["launch_ms"], 4),
        "copy_ceiling_GBs": round(copy_gbs, 1),
        "frac_of_copy": round(achieved / copy_gbs, 4),
        "traffic_source": ("profiles/traffic.json " + ttag + ": TCC_EA0_RDREQ_DRAM + TCC_EA0_WRREQ_DRAM per dispatch, "
                           "scaled by the copy kernel; FETCH_SIZE + WRITE_SIZE there too") if tr else None,
        "traffic_fetch_write": tr.get("bytes_per_launch") if tr else None,
    }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case, main_res["dt"], args.cpu_steps)

    out = {
        "metric": "6D cell-updates/s FP64 & FP32 at 24^3x16^3; batched p50/p99 latency",
        "value": main_res["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "n": list(case.n), "cells": cells, "cfl": 0.5,
                   "dt": main_res["dt"], "boundaries": "open x,y; reflecting z (reference semantics)",
                   "path": "homogeneous" if main_res["homogeneous"] else "general",
                   "clock_settle_steps": main_res["settle_steps"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if extras:
        out["extras"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
