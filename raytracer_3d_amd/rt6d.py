"""Python host side of the 6D step: ctypes bindings of include/rt6d.h and a
``Model`` whose methods carry the reference routine names
(for_src/main.f90:36-79 order):

    m = Model.from_case(case)        # set_parameter, allocate_main_module,
                                     # set_grid + calc_grid, set_initial_conditions
    m.calc_parameter()               # calc_parameter.f90:2
    m.integrate(nsteps)              # integrate.f90:4, nsteps times
    m.superbee_fluxes()              # superbee_fluxes.f90:4
    d = m.diagnose()                 # diagnose.f90:45-329 reductions

Every call goes to the HIP library librt6d.so; there is no CPU fallback.  A
missing or unloadable library raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import cases as _cases

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "librt6d.so")

FP64, FP32 = 0, 1
FIELD = {"E": 0, "dE": 1, "meanfl": 2, "omega_i": 3, "omega_e": 4,
         "xdot": 5, "ydot": 6, "zdot": 7, "kdot": 8, "ldot": 9, "mdot": 10,
         "flux_x": 11, "flux_y": 12, "flux_z": 13, "flux_k": 14, "flux_l": 15, "flux_m": 16,
         "lambda_x": 17, "lambda_y": 18, "lambda_z": 19, "lambda": 20}
LAYOUT_INTERIOR, LAYOUT_FACES, LAYOUT_HALO = 0, 1, 2
OPT_EXACT, OPT_PATH, OPT_BLOCK, OPT_KERNEL, OPT_SYNC = 1, 2, 3, 4, 5
TILES_WAVE_LOCAL, TILES_WAVE_LOCAL_1Z, TILES_PAIRS = 0, 1, 2
EQ_LAUNCHES, EQ_TASKS, EQ_X_HANDOVER = 1, 2, 3
Q_HOMOGENEOUS, Q_DEVICE_BYTES, Q_CELLS, Q_STEPS, Q_BYTES_PER_CELL, Q_KERNEL, Q_UNIFORM_XY = 1, 2, 3, 4, 5, 6, 7
Q_X_HANDOVER = 8

# every symbol include/rt6d.h declares
EXPORTS = ("rt6d_version", "rt6d_last_error", "rt6d_device_count", "rt6d_create", "rt6d_destroy",
           "rt6d_build_grid", "rt6d_set_grid", "rt6d_set_background", "rt6d_calc_parameter", "rt6d_cfl_dt",
           "rt6d_upload_E", "rt6d_step", "rt6d_step_timed", "rt6d_fluxes", "rt6d_download", "rt6d_diagnose",
           "rt6d_set_option", "rt6d_query", "rt6d_synchronize", "rt6d_copy_bandwidth",
           "rt6d_ensemble_create", "rt6d_ensemble_destroy", "rt6d_ensemble_add", "rt6d_ensemble_remove",
           "rt6d_ensemble_step", "rt6d_ensemble_set_schedule", "rt6d_ensemble_tune", "rt6d_ensemble_info",
           "rt6d_init_separable", "rt6d_branch_flips", "rt6d_ensemble_step_async", "rt6d_ensemble_wait",
           "rt6d_set_forcing", "rt6d_diagnose_noalloc", "rt6d_ensemble_set_tiles", "rt6d_ensemble_query")


class Diag(C.Structure):
    _fields_ = [("eint", C.c_double), ("mflint", C.c_double), ("mfl_max", C.c_double),
                ("vmax", C.c_double * 6), ("bfl", C.c_double * 12), ("cfl", C.c_double * 6),
                ("residual", C.c_double)]

    def as_dict(self) -> dict:
        names = ["bfl_xs", "bfl_xe", "bfl_ys", "bfl_ye", "bfl_zs", "bfl_ze",
                 "bfl_ks", "bfl_ke", "bfl_ls", "bfl_le", "bfl_ms", "bfl_me"]
        d = {"eint": self.eint, "mflint": self.mflint, "mfl_max": self.mfl_max, "residual": self.residual}
        for i, a in enumerate("uvwklm"):
            d[f"{a}max"] = self.vmax[i]
        for i, a in enumerate(_cases.AXES):
            d[f"cfl_{a}"] = self.cfl[i]
        for i, nm in enumerate(names):
            d[nm] = self.bfl[i]
        return d


_lib = None


def load_library(path: str | None = None):
    """Load librt6d.so (once) and declare its signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"librt6d.so not built ({p}); run python -m raytracer_3d_amd.build")
    lib = C.CDLL(p)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    sig = {
        "rt6d_version": ([], C.c_char_p),
        "rt6d_last_error": ([], C.c_char_p),
        "rt6d_device_count": ([], C.c_int),
        "rt6d_create": ([ip, C.c_int, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
        "rt6d_destroy": ([vp], C.c_int),
        "rt6d_build_grid": ([C.c_int, C.c_int, dp, C.c_double, dp, dp, dp, dp], C.c_int),
        "rt6d_set_grid": ([vp, C.c_int, dp, dp], C.c_int),
        "rt6d_set_background": ([vp, dp, dp, dp, C.c_double, C.c_double], C.c_int),
        "rt6d_calc_parameter": ([vp], C.c_int),
        "rt6d_cfl_dt": ([vp, C.c_double, dp], C.c_int),
        "rt6d_upload_E": ([vp, dp, C.c_int], C.c_int),
        "rt6d_step": ([vp, C.c_double, C.c_int], C.c_int),
        "rt6d_step_timed": ([vp, C.c_double, C.c_int, dp], C.c_int),
        "rt6d_fluxes": ([vp, C.c_double], C.c_int),
        "rt6d_download": ([vp, C.c_int, C.c_int, dp], C.c_int),
        "rt6d_diagnose": ([vp, C.c_double, C.POINTER(Diag), dp, dp, dp], C.c_int),
        "rt6d_diagnose_noalloc": ([vp, C.c_double, C.POINTER(Diag), dp, dp, dp], C.c_int),
        "rt6d_set_option": ([vp, C.c_int, C.c_int], C.c_int),
        "rt6d_query": ([vp, C.c_int, C.POINTER(C.c_longlong)], C.c_int),
        "rt6d_synchronize": ([vp], C.c_int),
        "rt6d_copy_bandwidth": ([C.c_longlong, C.c_int, dp], C.c_int),
        "rt6d_ensemble_create": ([C.c_int, C.POINTER(vp)], C.c_int),
        "rt6d_ensemble_destroy": ([vp], C.c_int),
        "rt6d_ensemble_add": ([vp, vp, C.c_double, ip], C.c_int),
        "rt6d_ensemble_remove": ([vp, C.c_int], C.c_int),
        "rt6d_ensemble_step": ([vp, C.c_int, dp], C.c_int),
        "rt6d_ensemble_set_schedule": ([vp, C.c_int, C.c_int], C.c_int),
        "rt6d_ensemble_tune": ([vp, C.c_int, ip, ip, ip], C.c_int),
        "rt6d_ensemble_set_tiles": ([vp, C.c_int], C.c_int),
        "rt6d_ensemble_info": ([vp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], C.c_int),
        "rt6d_ensemble_query": ([vp, C.c_int, C.POINTER(C.c_longlong)], C.c_int),
        "rt6d_init_separable": ([vp, dp, C.c_double, C.c_double, C.c_ulonglong, C.c_ulonglong], C.c_int),
        "rt6d_branch_flips": ([vp, vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], C.c_int),
        "rt6d_ensemble_step_async": ([vp, C.c_int], C.c_int),
        "rt6d_ensemble_wait": ([vp, dp], C.c_int),
        "rt6d_set_forcing": ([vp, C.c_int, dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if path is None:
        _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class RT6DError(RuntimeError):
    pass


def _check(lib, rc: int, what: str):
    if rc != 0:
        msg = lib.rt6d_last_error().decode(errors="replace")
        raise RT6DError(f"{what} failed ({rc}): {msg}")


def device_count() -> int:
    return load_library().rt6d_device_count()


def device_copy_timed(nbytes: int, reps: int = 10) -> float:
    """Average ms of a plain device copy of nbytes (rt6d_copy_bandwidth)."""
    lib = load_library()
    ms = C.c_double()
    _check(lib, lib.rt6d_copy_bandwidth(int(nbytes), reps, C.byref(ms)), "rt6d_copy_bandwidth")
    return ms.value


def branch_flips(m64: "Model", m32: "Model"):
    """(faces, flips) per axis between a binary64 and a binary32 model (rt6d_branch_flips)."""
    lib = load_library()
    faces, flips = (C.c_longlong * 6)(), (C.c_longlong * 6)()
    _check(lib, lib.rt6d_branch_flips(m64._h, m32._h, faces, flips), "rt6d_branch_flips")
    return list(faces), list(flips)


def build_grid(axis: int, spacing, fixed: float = 0.0):
    lib = load_library()
    s = np.ascontiguousarray(spacing, dtype=np.float64)
    N = s.size + 4
    t, dt, u, du = (np.zeros(N) for _ in range(4))
    _check(lib, lib.rt6d_build_grid(axis, s.size, _dptr(s), fixed, _dptr(t), _dptr(dt), _dptr(u), _dptr(du)),
           "rt6d_build_grid")
    return t, dt, u, du


class Model:
    """One resident integration on the current HIP device."""

    def __init__(self, n, reflect_up=True, reflect_down=True, precision=FP64):
        self.lib = load_library()
        self.n = tuple(int(v) for v in n)
        self.precision = precision
        self.cells = int(np.prod(self.n))
        self._h = C.c_void_p()
        arr = (C.c_int * 6)(*self.n)
        _check(self.lib, self.lib.rt6d_create(arr, int(reflect_up), int(reflect_down), precision,
                                              C.byref(self._h)), "rt6d_create")
        self.dt = None

    # -- lifetime
    def close(self):
        if self._h:
            self.lib.rt6d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- setup (set_grid/calc_grid, set_initial_conditions)
    def set_grid(self, axis: int, t, dt):
        t = np.ascontiguousarray(t, dtype=np.float64)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        if t.size != self.n[axis] + 4 or dt.size != self.n[axis] + 4:
            raise ValueError("grid arrays must hold n+4 values")
        _check(self.lib, self.lib.rt6d_set_grid(self._h, axis, _dptr(t), _dptr(dt)), "rt6d_set_grid")

    def set_background(self, Nsqr, U, V, f0=1e-4, sign=1.0):
        n3 = (self.n[0] + 4) * (self.n[1] + 4) * (self.n[2] + 4)
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (Nsqr, U, V)]
        if any(a.size != n3 for a in arrs):
            raise ValueError("background arrays must be halo'd 3D")
        _check(self.lib, self.lib.rt6d_set_background(self._h, *(_dptr(a) for a in arrs), f0, sign),
               "rt6d_set_background")

    def upload_E(self, E, layout=LAYOUT_INTERIOR):
        E = np.ascontiguousarray(E, dtype=np.float64).ravel()
        _check(self.lib, self.lib.rt6d_upload_E(self._h, _dptr(E), layout), "rt6d_upload_E")

    def init_separable(self, factors, peak: float, noise: float, seed: int, member: int):
        """E on the device: peak*fx*fy*fz*fk*fl*fm*(1+noise*(2u-1)) (rt6d_init_separable)."""
        f = np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=np.float64) for a in factors]))
        if f.size != sum(self.n):
            raise ValueError("factors must hold n0+...+n5 values")
        _check(self.lib, self.lib.rt6d_init_separable(self._h, _dptr(f), peak, noise, seed, member),
               "rt6d_init_separable")

    def set_forcing(self, kind: int, field=None):
        """Bottom forcing before every step: 0 none, 1 flux injection, 2 relaxation (rt6d_set_forcing)."""
        if kind == 0:
            _check(self.lib, self.lib.rt6d_set_forcing(self._h, 0, None), "rt6d_set_forcing")
            return
        f = np.ascontiguousarray(field, dtype=np.float64).ravel()
        if f.size != self.cells // self.n[2]:
            raise ValueError("forcing field must hold nx*ny*nk*nl*nm values")
        _check(self.lib, self.lib.rt6d_set_forcing(self._h, kind, _dptr(f)), "rt6d_set_forcing")

    @classmethod
    def from_case(cls, case: "_cases.Case", precision=FP64, exact=False, path=0, block=None,
                  kernel=0) -> "Model":
        m = cls(case.n, case.reflect_up, case.reflect_down, precision)
        if exact:
            m.set_option(OPT_EXACT, 1)
        if path:
            m.set_option(OPT_PATH, path)
        if block:
            m.set_option(OPT_BLOCK, block)
        if kernel:
            m.set_option(OPT_KERNEL, kernel)
        for ax, (t, dt) in enumerate(case.grids()):
            m.set_grid(ax, t, dt)
        m.set_background(case.Nsqr, case.U, case.V, case.f0, case.sign)
        if case.E0 is None and "init" in case.meta:
            i = case.meta["init"]
            m.init_separable(i["factors"], i["peak"], i["noise"], i["seed"], i["member"])
        else:
            m.upload_E(case.E0)
        m.dt = case.dt
        return m

    # -- the hot path, reference names
    def calc_parameter(self):
        _check(self.lib, self.lib.rt6d_calc_parameter(self._h), "rt6d_calc_parameter")

    def cfl_dt(self, cfl=0.5) -> float:
        out = C.c_double()
        _check(self.lib, self.lib.rt6d_cfl_dt(self._h, cfl, C.byref(out)), "rt6d_cfl_dt")
        return out.value

    def integrate(self, nsteps: int = 1, dt: float | None = None):
        dt = self.dt if dt is None else dt
        _check(self.lib, self.lib.rt6d_step(self._h, dt, nsteps), "rt6d_step")

    def integrate_timed(self, nsteps: int, dt: float | None = None) -> float:
        dt = self.dt if dt is None else dt
        ms = C.c_double()
        _check(self.lib, self.lib.rt6d_step_timed(self._h, dt, nsteps, C.byref(ms)), "rt6d_step_timed")
        return ms.value

    def superbee_fluxes(self, dt: float | None = None):
        dt = self.dt if dt is None else dt
        _check(self.lib, self.lib.rt6d_fluxes(self._h, dt), "rt6d_fluxes")

    def diagnose(self, dt: float | None = None, fields3d: bool = False, noalloc: bool = False):
        """rt6d_diagnose, or with noalloc rt6d_diagnose_noalloc (the same
        kernels on the library's per-GPU scratch, no device allocation)."""
        dt = self.dt if dt is None else dt
        d = Diag()
        name = "rt6d_diagnose_noalloc" if noalloc else "rt6d_diagnose"
        fn = getattr(self.lib, name)
        if fields3d:
            n3 = self.n[0] * self.n[1] * self.n[2]
            Ei, Mi, Di = np.zeros(n3), np.zeros(n3), np.zeros(6 * n3)
            _check(self.lib, fn(self._h, dt, C.byref(d), _dptr(Ei), _dptr(Mi), _dptr(Di)), name)
            return d.as_dict(), {"E_int": Ei, "meanfl_int": Mi, "diss": Di.reshape(6, n3)}
        _check(self.lib, fn(self._h, dt, C.byref(d), None, None, None), name)
        return d.as_dict()

    def download(self, field: str, layout: int | None = None) -> np.ndarray:
        fid = FIELD[field]
        if layout is None:
            layout = LAYOUT_FACES if (5 <= fid <= 16) else LAYOUT_INTERIOR
        if layout == LAYOUT_FACES:
            d = (fid - 5) % 6
            size = _cases.face_count(self.n, d)
        elif layout == LAYOUT_HALO:
            size = int(np.prod([v + 4 for v in self.n]))
        else:
            size = self.cells
        out = np.zeros(size)
        _check(self.lib, self.lib.rt6d_download(self._h, fid, layout, _dptr(out)), f"rt6d_download({field})")
        return out

    def set_option(self, opt: int, value: int):
        _check(self.lib, self.lib.rt6d_set_option(self._h, opt, value), "rt6d_set_option")

    def query(self, what: int) -> int:
        v = C.c_longlong()
        _check(self.lib, self.lib.rt6d_query(self._h, what, C.byref(v)), "rt6d_query")
        return v.value

    @property
    def homogeneous(self) -> bool:
        return bool(self.query(Q_HOMOGENEOUS))

    @property
    def bytes_per_cell(self) -> float:
        return self.query(Q_BYTES_PER_CELL) / 1000.0

    @property
    def kernel(self) -> int:
        """The step kernel option rt6d_step runs (the auto choice resolved)."""
        return self.query(Q_KERNEL)

    def synchronize(self):
        _check(self.lib, self.lib.rt6d_synchronize(self._h), "rt6d_synchronize")


class Ensemble:
    """Many independent integrations advanced together (rt6d_ensemble_*).

    Members are Models set up as for a single run (calc_parameter done, E
    uploaded).  While a member, a Model is stepped only through the ensemble;
    download/diagnose work on it between ensemble steps.
    """

    def __init__(self, precision=FP64):
        self.lib = load_library()
        self.precision = precision
        self._h = C.c_void_p()
        _check(self.lib, self.lib.rt6d_ensemble_create(precision, C.byref(self._h)), "rt6d_ensemble_create")
        self.members: dict[int, Model] = {}

    def close(self):
        if self._h:
            self.lib.rt6d_ensemble_destroy(self._h)
            self._h = C.c_void_p()
            self.members.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, model: Model, dt: float | None = None) -> int:
        dt = model.dt if dt is None else dt
        slot = C.c_int()
        _check(self.lib, self.lib.rt6d_ensemble_add(self._h, model._h, dt, C.byref(slot)), "rt6d_ensemble_add")
        self.members[slot.value] = model
        return slot.value

    def remove(self, slot: int) -> Model:
        _check(self.lib, self.lib.rt6d_ensemble_remove(self._h, slot), "rt6d_ensemble_remove")
        return self.members.pop(slot)

    def step(self, nsteps: int = 1) -> float:
        """Advance every member nsteps steps; returns the device time in ms."""
        ms = C.c_double()
        _check(self.lib, self.lib.rt6d_ensemble_step(self._h, nsteps, C.byref(ms)), "rt6d_ensemble_step")
        return ms.value

    def step_async(self, nsteps: int = 1):
        """Enqueue nsteps steps and return (see wait)."""
        _check(self.lib, self.lib.rt6d_ensemble_step_async(self._h, nsteps), "rt6d_ensemble_step_async")

    def wait(self) -> float:
        """Wait for enqueued steps; returns their device time in ms."""
        ms = C.c_double()
        _check(self.lib, self.lib.rt6d_ensemble_wait(self._h, C.byref(ms)), "rt6d_ensemble_wait")
        return ms.value

    def set_schedule(self, blocks_per_cu: int, order: int = 0):
        _check(self.lib, self.lib.rt6d_ensemble_set_schedule(self._h, blocks_per_cu, order),
               "rt6d_ensemble_set_schedule")

    def set_tiles(self, tiles: int):
        """TILES_WAVE_LOCAL, TILES_WAVE_LOCAL_1Z or TILES_PAIRS (rt6d_ensemble_set_tiles)."""
        _check(self.lib, self.lib.rt6d_ensemble_set_tiles(self._h, tiles), "rt6d_ensemble_set_tiles")

    def tune(self, trial_steps: int = 4):
        """rt6d_ensemble_tune: returns (blocks_per_cu, order, tiles) of the fastest trial."""
        b, o, t = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib, self.lib.rt6d_ensemble_tune(self._h, trial_steps, C.byref(b), C.byref(o), C.byref(t)),
               "rt6d_ensemble_tune")
        return b.value, o.value, t.value

    def info(self):
        m, c, s = C.c_int(), C.c_longlong(), C.c_longlong()
        _check(self.lib, self.lib.rt6d_ensemble_info(self._h, C.byref(m), C.byref(c), C.byref(s)),
               "rt6d_ensemble_info")
        return {"members": m.value, "cells": c.value, "steps": s.value}

    def query(self, what: int) -> int:
        """rt6d_ensemble_query: EQ_LAUNCHES (kernel launches per step), EQ_TASKS or
        EQ_X_HANDOVER (members whose step passes x faces between lanes)."""
        v = C.c_longlong()
        _check(self.lib, self.lib.rt6d_ensemble_query(self._h, what, C.byref(v)), "rt6d_ensemble_query")
        return v.value
