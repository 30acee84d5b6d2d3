"""Time one step configuration of the library (for profiling under rocprofv3).

    python -m raytracer_3d_amd.stepbench --workload gm --precision fp64 \
        --path auto --kernel auto --steps 20 --warmup 3 [--copy-mib 1024]

Workloads: ``gm`` (24^3 x 16^3 Garrett-Munk-shaped spectrum, the production
size), ``gm32`` (32^3 x 16^3), ``packet`` (32x4x32x16x4x16 wave packet).
``--copy-mib`` first runs a plain device-to-device copy of that size through
librt6d's own copy kernel, the reference point for the HBM byte counters.
Prints one JSON line with the event-timed launch average.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import cases, rt6d

WORKLOADS = {
    "gm": lambda: cases.gm_case(),
    "gm32": lambda: cases.gm_case(n=(32, 32, 32, 16, 16, 16)),
    "packet": lambda: cases.packet_case(),
}

# RT6D_OPT_KERNEL options by name
KERNELS = {"auto": 0, "percell": 1, "col1": 4, "col2": 5, "col4": 6, "coly1": 9, "coly2": 10, "coly4": 11,
           "colx1": 15, "colx2": 16, "colx4": 17, "colq1": 18, "colq2": 19, "colq4": 20,
           "colw1": 21, "colw2": 22, "colw4": 23}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="gm", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--path", default="auto", choices=["auto", "general"])
    ap.add_argument("--kernel", default="auto", choices=sorted(KERNELS, key=KERNELS.get))
    ap.add_argument("--exact", action="store_true")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=0)
    a = ap.parse_args(argv)
    if a.copy_mib:
        ms = rt6d.device_copy_timed(a.copy_mib << 20)
        print(json.dumps({"copy_bytes": a.copy_mib << 20, "copy_ms": ms}), flush=True)
    case = WORKLOADS[a.workload]()
    prec = rt6d.FP64 if a.precision == "fp64" else rt6d.FP32
    kern = KERNELS[a.kernel]
    m = rt6d.Model.from_case(case, precision=prec, path=1 if a.path == "general" else 0, kernel=kern,
                             exact=a.exact)
    m.calc_parameter()
    dt = m.cfl_dt(0.5)
    m.integrate(a.warmup, dt=dt)
    m.synchronize()
    t0 = time.perf_counter()
    ms = m.integrate_timed(a.steps, dt=dt)
    t1 = time.perf_counter()
    out = {
        "workload": a.workload, "n": list(case.n), "precision": a.precision, "path": a.path, "kernel": a.kernel,
        "kernel_option": m.kernel, "x_handover": m.query(rt6d.Q_X_HANDOVER),
        "exact": a.exact, "homogeneous": m.homogeneous, "steps": a.steps, "launch_ms": ms / a.steps,
        "wall_ms_per_step": 1e3 * (t1 - t0) / a.steps, "bytes_per_cell": m.bytes_per_cell,
        "cell_updates_per_s": case.cells * a.steps / (t1 - t0),
    }
    out["achieved_GBs"] = out["bytes_per_cell"] * case.cells / (out["launch_ms"] * 1e-3) / 1e9
    print(json.dumps(out), flush=True)
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
