"""Snapshot files snap3D.cdf and snap6D.cdf in the reference's schema.

The reference writes its diagnostics through the netCDF Fortran library:

* init_snap3D_cdf (for_src/netcdf_io_3D.f90:1-260) creates snap3D.cdf;
* init_snap6D_cdf_single (for_src/netcdf_io_6D.f90:240-656) creates
  snap6D.cdf;
* diagnose (for_src/diagnose.f90:335-530) appends one record to each file
  per snapshot.

Both files are netCDF "64-bit offset" files (CDF-2, NF_64BIT_OFFSET), with
NOFILL and the record dimension Time. This module writes the same
dimensions, variables, attributes and records. Names, definition order, the
24- and 16-character blank-padded long_name / units attributes and
time_origin all follow the reference. Data arrays are the reference's
Fortran arrays, so the file's dimension order is the Fortran order
reversed: (Time, zt, yt, xt) for the 3D fields, (Time, mt, lt, kt, zt, yt,
xt) for E.

The netCDF library is not part of this build. The writer below implements
the published CDF-2 layout (header, then the fixed-size variables, then the
records, all big-endian) with the standard library and numpy, and appends a
record by writing its slabs and then the new record count. Any netCDF
reader opens the files; the tests read them back with scipy.io.netcdf_file.

    snaps = Snapshots(model, case, directory=".", write_6d=True)
    model.integrate(1); snaps.diagnose(itt=0)
"""
from __future__ import annotations

import os
import struct

import numpy as np

from . import cases

NC_DIMENSION, NC_VARIABLE, NC_ATTRIBUTE = 10, 11, 12
NC_CHAR, NC_DOUBLE = 2, 6
_VSIZE_MAX = 2**32 - 4


def _pad4(n: int) -> int:
    return (n + 3) & ~3


def _name(s: str) -> bytes:
    b = s.encode("ascii")
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _attrs(attrs: list) -> bytes:
    if not attrs:
        return b"\0" * 8
    out = struct.pack(">ii", NC_ATTRIBUTE, len(attrs))
    for key, val in attrs:
        if isinstance(val, str):
            b = val.encode("ascii")
            out += _name(key) + struct.pack(">ii", NC_CHAR, len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))
        else:
            v = np.atleast_1d(np.asarray(val, dtype=">f8"))
            out += _name(key) + struct.pack(">ii", NC_DOUBLE, v.size) + v.tobytes()
    return out


def char_att(text: str, width: int) -> str:
    """A Fortran CHARACTER*width value as ncaptc stores it: blank-padded."""
    if len(text) > width:
        raise ValueError(f"{text!r} is longer than {width} characters")
    return text.ljust(width)


class CdfVar:
    def __init__(self, name, dims, attrs):
        self.name, self.dims, self.attrs = name, tuple(dims), list(attrs)
        self.begin = 0
        self.vsize = 0
        self.shape = ()
        self.record = False


class CdfFile:
    """A netCDF CDF-2 file with doubles only. Define the dimensions, variables
    and global attributes, then call create(). After that, fill the
    fixed-size variables with put() and add records with append().

    dims: (name, length) pairs; length None marks the record dimension.
    """

    def __init__(self, path: str):
        self.path = path
        self.dims: list[tuple[str, int | None]] = []
        self.vars: list[CdfVar] = []
        self.gatts: list = []
        self.numrecs = 0
        self.recsize = 0

    def dim(self, name: str, length: int | None):
        if length is None and any(d[1] is None for d in self.dims):
            raise ValueError("one record dimension only")
        self.dims.append((name, length))

    def var(self, name: str, dims, long_name: str | None = None, units: str | None = None, extra=()):
        attrs = []
        if long_name is not None:
            attrs.append(("long_name", char_att(long_name, 24)))
        if units is not None:
            attrs.append(("units", char_att(units, 16)))
        attrs.extend(extra)
        self.vars.append(CdfVar(name, dims, attrs))

    def _dim_index(self, name: str) -> int:
        for i, (nm, _) in enumerate(self.dims):
            if nm == name:
                return i
        raise KeyError(name)

    def _layout(self) -> bytes:
        lens = dict(self.dims)
        for v in self.vars:
            v.record = any(lens[d] is None for d in v.dims)
            if v.record and lens[v.dims[0]] is not None:
                raise ValueError(f"{v.name}: the record dimension must come first")
            v.shape = tuple(lens[d] for d in v.dims if lens[d] is not None)
            v.vsize = _pad4(8 * int(np.prod(v.shape, dtype=np.int64)))

        def header(begins) -> bytes:
            h = b"CDF\x02" + struct.pack(">i", self.numrecs)
            h += struct.pack(">ii", NC_DIMENSION, len(self.dims))
            for nm, ln in self.dims:
                h += _name(nm) + struct.pack(">i", 0 if ln is None else ln)
            h += _attrs(self.gatts)
            h += struct.pack(">ii", NC_VARIABLE, len(self.vars))
            for v, b in zip(self.vars, begins):
                h += _name(v.name) + struct.pack(">i", len(v.dims))
                h += b"".join(struct.pack(">i", self._dim_index(d)) for d in v.dims)
                h += _attrs(v.attrs) + struct.pack(">i", NC_DOUBLE)
                h += struct.pack(">I", min(v.vsize, 2**32 - 1)) + struct.pack(">q", b)
            return h

        size = len(header([0] * len(self.vars)))     # begin fields have fixed width
        off = size
        for v in self.vars:
            if not v.record:
                v.begin = off
                off += v.vsize
        rec_vars = [v for v in self.vars if v.record]
        for i, v in enumerate(rec_vars):
            if v.vsize > _VSIZE_MAX and i != len(rec_vars) - 1:
                raise ValueError(f"{v.name}: a record slab above 4 GiB is allowed only for the last record "
                                 "variable of a 64-bit offset file")
            v.begin = off
            off += v.vsize
        self.recsize = sum(v.vsize for v in rec_vars)
        self._rec_begin = rec_vars[0].begin if rec_vars else off
        return header([v.begin for v in self.vars])

    def create(self):
        h = self._layout()
        with open(self.path, "wb") as f:
            f.write(h)

    def _get(self, name: str) -> CdfVar:
        for v in self.vars:
            if v.name == name:
                return v
        raise KeyError(name)

    @staticmethod
    def _write(f, off: int, data: np.ndarray, nbytes: int):
        a = np.ascontiguousarray(data, dtype=">f8")
        if a.nbytes != nbytes:
            raise ValueError(f"{a.nbytes} bytes given, {nbytes} expected")
        f.seek(off)
        a.tofile(f)

    def put(self, name: str, data):
        """Write a fixed-size variable. data: its values with the file's last
        dimension fastest (a Fortran array's flat storage)."""
        v = self._get(name)
        if v.record:
            raise ValueError(f"{name} is a record variable")
        with open(self.path, "r+b") as f:
            self._write(f, v.begin, np.asarray(data).reshape(-1, order="A"), 8 * int(np.prod(v.shape)))

    def append(self, values: dict):
        """Add one record; values maps every record variable to its data."""
        rec = self.numrecs
        with open(self.path, "r+b") as f:
            for v in self.vars:
                if not v.record:
                    continue
                if v.name not in values:
                    raise KeyError(f"record variable {v.name} missing")
                self._write(f, v.begin + rec * self.recsize, np.asarray(values[v.name]).reshape(-1, order="A"),
                            8 * int(np.prod(v.shape, dtype=np.int64)))
            # the record count last, so that a reader never sees half a record
            f.seek(4)
            f.write(struct.pack(">i", rec + 1))
        self.numrecs = rec + 1
        return self.numrecs


def _interior3(a: np.ndarray, n) -> np.ndarray:
    """Interior of a halo'd (onx = 2) 3D Fortran array, flat, x fastest."""
    h = np.asarray(a, dtype=np.float64).reshape([v + 2 * cases.ONX for v in n[:3]], order="F")
    s = slice(cases.ONX, -cases.ONX)
    return h[s, s, s].reshape(-1, order="F")


def _axis_grid(case, ax: int):
    """Interior (t, u, dt) of one axis as calc_grid makes them."""
    fx = case.fixed[ax - 3] if ax >= 3 else 0.0
    t, d, u, _ = cases.calc_grid(ax, case.spacing[ax], fx)
    s = slice(cases.ONX, cases.ONX + case.n[ax])
    return t[s], u[s], d[s]


def _faces_to_cells(flat: np.ndarray, n, d: int) -> np.ndarray:
    """A face array (n_d + 1 entries along axis d, faces 0..n_d) restricted to
    the reference's indices 1..n_d (the right face of each cell), flat."""
    shape = list(n)
    shape[d] += 1
    a = np.asarray(flat).reshape(shape, order="F")
    sl = [slice(None)] * 6
    sl[d] = slice(1, None)
    return a[tuple(sl)].reshape(-1, order="F")


BFLUX = ("xs", "xe", "ys", "ye", "ks", "ke", "ls", "le", "ms", "me")
DISS = ("diss_ks", "diss_ke", "diss_ls", "diss_le", "diss_ms", "diss_me")
TIME_ORIGIN = ("time_origin", "01-JAN-1900 00:00:00")


def define_snap3d(path: str, case) -> CdfFile:
    """init_snap3D_cdf (for_src/netcdf_io_3D.f90:1-175 definitions, :176-260
    the grid and background data)."""
    nx, ny, nz = case.n[:3]
    f = CdfFile(path)
    for nm, ln in (("xt", nx), ("xu", nx), ("yt", ny), ("yu", ny), ("zt", nz), ("zu", nz), ("Time", None)):
        f.dim(nm, ln)
    f.var("Time", ["Time"], "Time ", "seconds", extra=[TIME_ORIGIN])
    for a in "xyz":
        f.var(f"{a}t", [f"{a}t"], "distance", "m")
        f.var(f"{a}u", [f"{a}u"], "distance", "m")
    for nm, ln, un in (("Nsqr", "buoyancy frequency", "1/s"), ("U", "mean flow", "m/s"), ("V", "mean flow", "m/s")):
        f.var(nm, ["zt", "yt", "xt"], ln, un)
    f.var("E", ["Time", "zt", "yt", "xt"], "energy density", "m^2/s^2")
    for nm in ("meanfl",) + DISS:
        f.var(nm, ["Time", "zt", "yt", "xt"], "energy change", "m^2/s^3")
    for b in BFLUX:
        side = "min" if b[1] == "s" else "max"
        f.var(f"bflux_{b}", ["Time"], f"Energy flux across {side} {b[0]}", " ")
    f.var("mflint", ["Time"], "mean flow energy flux", " ")
    f.gatts = [("f0", case.f0), ("k_fixed", case.fixed[0]), ("l_fixed", case.fixed[1]), ("m_fixed", case.fixed[2])]
    f.create()
    f.put("Nsqr", _interior3(case.Nsqr, case.n))
    f.put("U", _interior3(case.U, case.n))
    f.put("V", _interior3(case.V, case.n))
    for ax, a in enumerate("xyz"):
        t, u, _ = _axis_grid(case, ax)
        f.put(f"{a}t", t)
        f.put(f"{a}u", u)
    return f


def define_snap6d(path: str, case, model) -> CdfFile:
    """init_snap6D_cdf_single (for_src/netcdf_io_6D.f90:240-398 definitions,
    :399-656 the grid, the group velocities and the frequencies)."""
    f = CdfFile(path)
    for ax, a in enumerate(cases.AXES):
        f.dim(f"{a}t", case.n[ax])
        f.dim(f"{a}u", case.n[ax])
    f.dim("Time", None)
    f.var("Time", ["Time"], "Time ", "seconds", extra=[TIME_ORIGIN])
    for ax, a in enumerate(cases.AXES):
        ln, un = ("distance", "m") if ax < 3 else ("wavenumber", "1/m")
        f.var(f"{a}t", [f"{a}t"], ln, un)
        f.var(f"{a}u", [f"{a}u"], ln, un)
        f.var(f"d{a}t", [f"{a}u"], ln, un)
    tdims = [f"{a}t" for a in reversed(cases.AXES)]
    vel = (("hor. group velocity", "m/s"), ("hor. group velocity", "m/s"), ("vertical group velocity", "m/s"),
           ("hor. wavenr. change", "1/ms"), ("hor. wavenr. change", "1/ms"), ("vert. wavenr. change", "1/ms"))
    for ax, a in enumerate(cases.AXES):
        dims = list(tdims)
        dims[5 - ax] = f"{a}u"
        f.var(f"{a}dot", dims, *vel[ax])
    f.var("omega_i", tdims, "intrinsic frequency", "1/s")
    f.var("omega_e", tdims, "extrinsic frequency", "1/s")
    f.var("E", ["Time"] + tdims, "energy density", "m^5/s^2")
    f.var("meanfl", ["Time"] + tdims, "energy change", "m^5/s^3")
    f.create()
    for ax, a in enumerate(cases.AXES):
        t, u, d = _axis_grid(case, ax)
        f.put(f"{a}t", t)
        f.put(f"{a}u", u)
        f.put(f"d{a}t", d)
    for ax, a in enumerate(cases.AXES):
        f.put(f"{a}dot", _faces_to_cells(model.download(f"{a}dot"), case.n, ax))
    f.put("omega_i", model.download("omega_i"))
    f.put("omega_e", model.download("omega_e"))
    return f


def per_pe_6d_name(n) -> str:
    """init_snap6D_cdf_many's file name (for_src/netcdf_io_6D.f90:10-13) for
    the one block that covers the whole grid: i5 fields, blanks made '0'."""
    parts = []
    for tag, ln in zip(("i", "j", "k", "l", "n", "m"), n):
        parts.append(f"{tag}={1:5d}:{ln:5d}")
    return ("snap6D_" + "_".join(parts) + ".cdf").replace(" ", "0")


class Snapshots:
    """init_diag (for_src/diagnose.f90:2-13) and the file part of diagnose
    (for_src/diagnose.f90:335-530) for one Model.

    write_6d: also keep snap6D.cdf (E and meanfl, 7D); the reference always
    writes it, but at production sizes one record is 2 x 8 bytes per cell.
    """

    def __init__(self, model, case, directory: str = ".", write_6d: bool = True, per_pe_name: bool = False):
        os.makedirs(directory, exist_ok=True)
        self.model = model
        self.case = case
        self.snap3d = define_snap3d(os.path.join(directory, "snap3D.cdf"), case)
        self.snap6d = None
        if write_6d:
            name = per_pe_6d_name(case.n) if per_pe_name else "snap6D.cdf"
            self.snap6d = define_snap6d(os.path.join(directory, name), case, model)

    def diagnose(self, itt: int, dt: float | None = None) -> dict:
        """Reductions on the device and one record in each file at Time =
        itt * dt. Returns the budget dict of Model.diagnose."""
        dt = self.model.dt if dt is None else dt
        d, f3 = self.model.diagnose(dt=dt, fields3d=True)
        time = np.array([itt * dt])
        if self.snap6d is not None:
            self.snap6d.append({"Time": time, "E": self.model.download("E"),
                                "meanfl": self.model.download("meanfl")})
        rec = {"Time": time, "E": f3["E_int"], "meanfl": f3["meanfl_int"], "mflint": np.array([d["mflint"]])}
        for i, nm in enumerate(DISS):
            rec[nm] = f3["diss"][i]
        for b in BFLUX:
            rec[f"bflux_{b}"] = np.array([d[f"bfl_{b}"]])
        self.snap3d.append(rec)
        return d
