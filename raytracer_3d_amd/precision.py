"""Error of the FP32 form against the binary64 step.

The FP32 form carries E and the limited fluxes in binary32 with binary32
arithmetic; the coefficient fields come from the binary64 calc_parameter
rounded once, dt is chosen in binary64 and the energy sums of diagnose are
compensated binary64 sums.  This module runs one problem in both forms side
by side and reports, at checkpoints:

* per field (E, and at the end dE, meanfl and the six face-flux fields) the
  relative L2 error and the largest absolute error relative to max|field|;
* the energy budget: the relative difference of the total energy (E times
  the cell volume, summed), and each run's step budget residual (diagnose's
  eint + mflint + boundary-flux integrals, relative to the sum of the terms'
  magnitudes);
* the superbee branch flips: faces whose limiter branch differs between the
  two runs (rt6d_branch_flips).

The binary64 run is the library's plain step, which the parity tests pin to
the reference Fortran at relative L2 <= 1e-12 per field.

    python -m raytracer_3d_amd.precision --n 24 24 24 16 16 16 --steps 500
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import cases, rt6d

FIELDS_END = ["dE", "meanfl", "flux_x", "flux_y", "flux_z", "flux_k", "flux_l", "flux_m"]


def _err(a32, a64):
    d = a32 - a64
    n64 = float(np.linalg.norm(a64))
    mx = float(np.max(np.abs(a64))) if a64.size else 0.0
    return {"rel_l2": float(np.linalg.norm(d)) / n64 if n64 else float(np.linalg.norm(d)),
            "max_abs_rel": float(np.max(np.abs(d))) / mx if mx else float(np.max(np.abs(d)))}


def total_energy(E, case) -> float:
    """Sum of E times the cell volume (the product of the six T spacings)."""
    v = np.asarray(E, dtype=np.float64).reshape(case.n, order="F")
    for d in reversed(range(6)):
        v = np.tensordot(v, np.asarray(case.spacing[d], dtype=np.float64), axes=([d], [0]))
    return float(v)


def fp32_report(case, steps=500, checkpoints=None, cfl=0.5, end_fields=True, log=None):
    if checkpoints is None:
        checkpoints = sorted({min(steps, c) for c in (1, 10, 100, 200, 300, 400, 500, steps)})
    m64 = rt6d.Model.from_case(case, precision=rt6d.FP64)
    m32 = rt6d.Model.from_case(case, precision=rt6d.FP32)
    out = {"n": list(case.n), "cells": case.cells, "steps": steps, "checkpoints": []}
    try:
        m64.calc_parameter()
        m32.calc_parameter()
        dt = m64.cfl_dt(cfl)
        out["dt"] = dt
        out["homogeneous"] = [m64.homogeneous, m32.homogeneous]
        te0 = total_energy(case.E0, case) if case.E0 is not None else total_energy(m64.download("E"), case)
        out["total_energy_0"] = te0
        done = 0
        t_gpu = [0.0, 0.0]
        for cp in checkpoints:
            k = cp - done
            if k > 0:
                t_gpu[0] += m64.integrate_timed(k, dt=dt)
                t_gpu[1] += m32.integrate_timed(k, dt=dt)
                done = cp
            E64 = m64.download("E")
            E32 = m32.download("E")
            d64 = m64.diagnose(dt=dt)
            d32 = m32.diagnose(dt=dt)
            faces, flips = rt6d.branch_flips(m64, m32)
            te64, te32 = total_energy(E64, case), total_energy(E32, case)
            terms64 = abs(d64["eint"]) + abs(d64["mflint"]) + sum(abs(v) for k_, v in d64.items()
                                                                    if k_.startswith("bfl_"))
            rec = {"step": cp, "E": _err(E32, E64),
                   "total_energy_drift": (te32 - te64) / te64,
                   "total_energy_change_fp64": (te64 - te0) / te0,
                   "eint_rel_diff": (d32["eint"] - d64["eint"]) / abs(d64["eint"]) if d64["eint"] else None,
                   "budget_residual_rel_fp64": d64["residual"] / terms64 if terms64 else None,
                   "budget_residual_rel_fp32": d32["residual"] / terms64 if terms64 else None,
                   "flips": int(sum(flips)), "faces": int(sum(faces)),
                   "flip_fraction": sum(flips) / max(1, sum(faces)),
                   "flips_per_axis": dict(zip(cases.AXES, flips))}
            out["checkpoints"].append(rec)
            if log:
                log(f"step {cp}: E rel L2 {rec['E']['rel_l2']:.3e}, energy drift {rec['total_energy_drift']:.3e}, "
                    f"flips {rec['flips']} of {rec['faces']}")
        if end_fields:
            out["end_fields"] = {}
            for f in FIELDS_END:
                out["end_fields"][f] = _err(m32.download(f), m64.download(f))
        out["ms_per_step"] = {"fp64": t_gpu[0] / max(1, done), "fp32": t_gpu[1] / max(1, done)}
    finally:
        m64.close()
        m32.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs=6, default=[24, 24, 24, 16, 16, 16])
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--no-end-fields", action="store_true")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    case = cases.gm_case(n=tuple(a.n))
    print(f"[precision] built {case.n} in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    rep = fp32_report(case, steps=a.steps, end_fields=not a.no_end_fields,
                      log=lambda s: print("[precision] " + s, file=sys.stderr, flush=True))
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
