"""ctypes bindings of the C oracle (oracle/rt6d_oracle.c) and a runner for the
reference harness (oracle/_ref/ref_harness).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "ref_harness")

FIELD = {"E": 0, "dE": 1, "meanfl": 2, "omega_i": 3, "omega_e": 4,
         "xdot": 5, "ydot": 6, "zdot": 7, "kdot": 8, "ldot": 9, "mdot": 10,
         "flux_x": 11, "flux_y": 12, "flux_z": 13, "flux_k": 14, "flux_l": 15, "flux_m": 16,
         "lambda_x": 17, "lambda_y": 18, "lambda_z": 19}

DIAG_NAMES = (["eint", "mflint", "mfl_max", "umax", "vmax", "wmax", "kmax", "lmax", "mmax"]
              + ["bfl_xs", "bfl_xe", "bfl_ys", "bfl_ye", "bfl_zs", "bfl_ze",
                 "bfl_ks", "bfl_ke", "bfl_ls", "bfl_le", "bfl_ms", "bfl_me"]
              + [f"cfl_{a}" for a in "xyzklm"] + ["residual"])


def build(quiet: bool = True) -> str:
    """Compile liboracle.so with -ffp-contract=off (see Makefile)."""
    r = subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        lib.or_create.argtypes = [ip, C.c_int, C.c_int]
        lib.or_create.restype = vp
        lib.or_destroy.argtypes = [vp]
        lib.or_set_params.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        lib.or_set_spacing.argtypes = [vp, C.c_int, dp]
        lib.or_calc_grid.argtypes = [vp]
        lib.or_get_grid.argtypes = [vp, C.c_int, dp, dp, dp, dp]
        lib.or_set_grid.argtypes = [vp, C.c_int, dp, dp]
        lib.or_set_background.argtypes = [vp, dp, dp, dp]
        lib.or_calc_parameter.argtypes = [vp]
        lib.or_fill_ghosts.argtypes = [vp]
        lib.or_superbee_fluxes.argtypes = [vp, C.c_double]
        lib.or_integrate.argtypes = [vp, C.c_double, C.c_int]
        lib.or_diagnose.argtypes = [vp, C.c_double, dp, dp, dp, C.POINTER(dp)]
        lib.or_get_field.argtypes = [vp, C.c_int, C.c_int, dp]
        lib.or_get_field.restype = C.c_int
        lib.or_set_field.argtypes = [vp, C.c_int, C.c_int, dp]
        lib.or_set_field.restype = C.c_int
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """The reference state (halo'd arrays) and routines, restated in C."""

    def __init__(self, case):
        self.lib = load()
        self.n = tuple(case.n)
        arr = (C.c_int * 6)(*self.n)
        self.h = self.lib.or_create(arr, int(case.reflect_up), int(case.reflect_down))
        self.lib.or_set_params(self.h, case.f0, case.sign, *case.fixed)
        for ax in range(6):
            s = np.ascontiguousarray(case.spacing[ax], dtype=np.float64)
            self.lib.or_set_spacing(self.h, ax, _dp(s))
        self.lib.or_calc_grid(self.h)
        self.lib.or_set_background(self.h, *(_dp(np.ascontiguousarray(a)) for a in (case.Nsqr, case.U, case.V)))
        self.set("E", case.E0)
        self.dt = case.dt

    def close(self):
        if self.h:
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self, axis):
        h = self.n[axis] + 4
        t, dt, u, du = (np.zeros(h) for _ in range(4))
        self.lib.or_get_grid(self.h, axis, _dp(t), _dp(dt), _dp(u), _dp(du))
        return t, dt, u, du

    def calc_parameter(self):
        self.lib.or_calc_parameter(self.h)

    def integrate(self, nsteps=1, dt=None):
        self.lib.or_integrate(self.h, self.dt if dt is None else dt, nsteps)

    def superbee_fluxes(self, dt=None):
        self.lib.or_fill_ghosts(self.h)
        self.lib.or_superbee_fluxes(self.h, self.dt if dt is None else dt)

    def get(self, name, faces=None):
        fid = FIELD[name]
        if faces is None:
            faces = 5 <= fid <= 16
        if faces:
            d = (fid - 5) % 6
            size = int(np.prod(self.n)) // self.n[d] * (self.n[d] + 1)
        else:
            size = int(np.prod(self.n))
        out = np.zeros(size)
        rc = self.lib.or_get_field(self.h, fid, 1 if faces else 0, _dp(out))
        if rc:
            raise ValueError(f"or_get_field({name}) = {rc}")
        return out

    def set(self, name, values, faces=False):
        v = np.ascontiguousarray(values, dtype=np.float64)
        rc = self.lib.or_set_field(self.h, FIELD[name], 1 if faces else 0, _dp(v))
        if rc:
            raise ValueError(f"or_set_field({name}) = {rc}")

    def diagnose(self, dt=None, fields3d=False):
        out = np.zeros(28)
        n3 = self.n[0] * self.n[1] * self.n[2]
        if fields3d:
            Ei, Mi = np.zeros(n3), np.zeros(n3)
            D = [np.zeros(n3) for _ in range(6)]
            arr = (C.POINTER(C.c_double) * 6)(*[_dp(a) for a in D])
            self.lib.or_diagnose(self.h, self.dt if dt is None else dt, _dp(out), _dp(Ei), _dp(Mi), arr)
            return dict(zip(DIAG_NAMES, out.tolist())), {"E_int": Ei, "meanfl_int": Mi, "diss": np.stack(D)}
        self.lib.or_diagnose(self.h, self.dt if dt is None else dt, _dp(out), None, None, None)
        return dict(zip(DIAG_NAMES, out.tolist()))


def cfl_dt_from(vel: dict, n, spacing, cfl=0.5) -> float:
    """dt = cfl / max_cell sum_d max(|v_d left|, |v_d right|)/dxt_d(c) from
    face velocity arrays (numpy; the rule rt6d_cfl_dt uses on the device)."""
    tot = np.zeros(tuple(n), order="F")
    for d, a in enumerate("xyzklm"):
        if n[d] <= 1:
            continue
        shp = list(n)
        shp[d] += 1
        v = np.abs(vel[f"{a}dot"].reshape(shp, order="F"))
        sl_l = [slice(None)] * 6
        sl_r = [slice(None)] * 6
        sl_l[d] = slice(0, n[d])
        sl_r[d] = slice(1, n[d] + 1)
        m = np.maximum(v[tuple(sl_l)], v[tuple(sl_r)])
        bshape = [1] * 6
        bshape[d] = n[d]
        tot += m / np.asarray(spacing[d]).reshape(bshape)
    return cfl / tot.max()


def run_reference(case_file: str, out_file: str = "-", np_grid=None, mpirun: str = "/opt/conda/bin/mpirun",
                  timeout: int = 600) -> float:
    """Run oracle/_ref/ref_harness (single rank, or mpirun with a 6D PE grid);
    returns the seconds it reports for its integrate loop."""
    if not os.path.exists(REF_BIN):
        raise RuntimeError("reference harness not built (oracle/build_ref.sh)")
    env = dict(os.environ)
    if np_grid is None or int(np.prod(np_grid)) == 1:
        cmd = [REF_BIN, case_file, out_file]
    else:
        npes = int(np.prod(np_grid))
        cmd = [mpirun, "-np", str(npes), REF_BIN, case_file, out_file] + [str(v) for v in np_grid]
    def _stack_unlimited():
        # the reference keeps 6D automatic arrays (calc_parameter.f90:7-8) on the stack
        import resource
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env,
                       preexec_fn=_stack_unlimited)
    if r.returncode != 0:
        raise RuntimeError(f"reference run failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    for line in r.stdout.splitlines():
        if line.startswith("RT6D_REF"):
            return float(line.split("seconds=")[1])
    raise RuntimeError("reference run printed no timing line: " + r.stdout[-2000:])
