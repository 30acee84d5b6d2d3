"""The reference's time loop on a Python host: set-up, restart, snapshots.

run() keeps the order of for_src/main.f90:36-81:
  set-up (grids, background, initial E; the case supplies them),
  calc_parameter, read_restart (when asked and a restart file exists),
  enditt = itt + number of steps, init_diag (snap3D.cdf / snap6D.cdf),
  then per step: set_forcing (the device hook of rt6d_set_forcing, set once),
  integrate, diagnose when mod(itt, int(snapint/dt)) == 0 or itt == 0,
  itt = itt + 1; finally write_restart.
Between two diagnose points no host work is needed, so those steps go to
the device as one rt6d_step call (the same steps, in the same order).

    python -m raytracer_3d_amd.driver CASE_FILE --snapint 3000 --restart --dir run1
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import cases, restart as _restart, rt6d, snapshot


def run(case, nsteps: int | None = None, runlen: float | None = None, snapint: float | None = None,
        precision=rt6d.FP64, directory: str = ".", use_restart: bool = False, write_6d: bool = True,
        forcing: tuple | None = None, cfl: float = 0.5, log=None) -> dict:
    """Integrate `case` as main.f90 does. The step count is nsteps, else
    int(runlen/dt) (main.f90:51), else case.nsteps. dt is case.dt, else the CFL
    choice. forcing: (kind, field) for rt6d_set_forcing. Returns itt, the
    steps taken, the step-loop seconds and rate, and the snapshot count."""
    os.makedirs(directory, exist_ok=True)
    say = log or (lambda s: None)
    m = rt6d.Model.from_case(case, precision=precision)
    try:
        m.calc_parameter()
        dt = case.dt if case.dt is not None else m.cfl_dt(cfl)
        m.dt = dt
        if forcing is not None:
            m.set_forcing(*forcing)
        itt = 0
        rfile = os.path.join(directory, _restart.restart_file_name(0))
        if use_restart:
            if os.path.exists(rfile):
                say(f"reading restart file {rfile}")
                itt = _restart.load(m, rfile)
            else:
                say(f"found no restart file {rfile}")
        steps = nsteps if nsteps is not None else (int(runlen / dt) if runlen is not None else case.nsteps)
        enditt = itt + steps
        snapsteps = max(1, int(snapint / dt)) if snapint else 0
        snaps = snapshot.Snapshots(m, case, directory, write_6d=write_6d) if snapsteps else None
        say(f"integrating from step {itt} to {enditt} with dt = {dt:.6g} s"
            + (f", snapshots every {snapsteps} steps" if snapsteps else ""))
        records = 0
        loop_s = 0.0
        while itt < enditt:
            # the next step after which diagnose runs (main.f90:74), or the end
            if snapsteps:
                nxt = itt if (itt == 0 or itt % snapsteps == 0) else (itt // snapsteps + 1) * snapsteps
                last = min(nxt, enditt - 1)
            else:
                last = enditt - 1
            t0 = time.perf_counter()
            m.integrate(last - itt + 1, dt=dt)
            m.synchronize()
            loop_s += time.perf_counter() - t0
            itt = last
            if snapsteps and (itt == 0 or itt % snapsteps == 0):
                d = snaps.diagnose(itt, dt=dt)
                records += 1
                say(f"snapshot at itt = {itt} ({itt * dt:.2f} s): eint {d['eint']:.6e} mflint {d['mflint']:.6e} "
                    f"residual {d['residual']:.3e}")
            itt += 1
        if use_restart:
            say(f"writing restart file {rfile} at itt = {itt}")
            _restart.save(m, rfile, itt)
        E = m.download("E")
    finally:
        m.close()
    cells = case.cells
    return {"itt": itt, "steps": steps, "dt": dt, "loop_s": loop_s,
            "cell_updates_per_s": cells * steps / loop_s if loop_s > 0 else None,
            "snapshots": records, "E": E}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("case", help="case file (cases.write_case_file format)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--runlen", type=float, default=None, help="seconds to integrate (main.f90's runlen)")
    ap.add_argument("--snapint", type=float, default=None, help="seconds between snapshots (main.f90's snapint)")
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--restart", action="store_true", help="read_restart before and write_restart after the loop")
    ap.add_argument("--dir", default=".")
    ap.add_argument("--no-6d", action="store_true", help="write snap3D.cdf only")
    ap.add_argument("--out", default=None, help="write the final interior E as a raw float64 stream")
    a = ap.parse_args(argv)
    case = cases.read_case_file(a.case)
    res = run(case, nsteps=a.steps, runlen=a.runlen, snapint=a.snapint,
              precision=rt6d.FP32 if a.fp32 else rt6d.FP64, directory=a.dir, use_restart=a.restart,
              write_6d=not a.no_6d, log=lambda s: print(" " + s, flush=True))
    if a.out:
        np.asarray(res["E"], dtype="<f8").tofile(a.out)
    res.pop("E")
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
