"""Restart files in the reference's record layout, for Python hosts.

write_restart / read_restart (for_src/restart.f90:1-93) keep one
unformatted sequential file per PE, restart_pe=<my_pe, i9, blanks as 0>.dta,
with four records:
  1. nx, ny, nz, nk, nl, nm, itt            (7 x int32)
  2. xs_pe, xe_pe, ys_pe, ye_pe, zs_pe, ze_pe (6 x int32)
  3. ks_pe, ke_pe, ls_pe, le_pe, ms_pe, me_pe (6 x int32)
  4. E, dE, meanfl over their halo'd extents (n_d + 4 per axis, binary64)
and write the step counter to a text file 'ritt'. Records are framed as
gfortran (the reference's compiler, make.inc) writes unformatted sequential
files: 4-byte little-endian markers around subrecords of at most
2,147,483,639 bytes.  A record longer than that (the field record from
prod(n+4) > 89.5M, e.g. 24^3 x 16^3) is split: a negative leading marker
says the record continues in the next subrecord, a negative trailing
marker that the subrecord continues a previous one.
One device holds the whole grid, so the PE bounds are 1..n_d.

The Fortran layer (raytracer_3d_amd/fortran/rt6d_hot.f90) writes and reads
the same files; the tests read each side's files with the other.

Only E is needed to continue a run: dE and meanfl are by-products of the
last step, which the next step recomputes.
"""
from __future__ import annotations

import os

import numpy as np

from . import rt6d

ONX = 2


def restart_file_name(my_pe: int = 0) -> str:
    """restart_file_name (for_src/restart.f90:7-8)."""
    return f"restart_pe={my_pe:9d}.dta".replace(" ", "0")


# gfortran's default -fmax-subrecord-length
MAX_SUBRECORD = 2147483639


def _marker(v: int) -> bytes:
    return np.array([v], dtype="<i4").tobytes()


def _record(f, parts, max_subrecord: int = MAX_SUBRECORD):
    """Write one record whose payload is the concatenation of `parts`
    (bytes-like, 1-D uint8 views), split into subrecords of at most
    max_subrecord bytes."""
    views = [memoryview(p).cast("B") for p in parts]
    total = sum(v.nbytes for v in views)
    left, first = total, True
    pi, off = 0, 0
    while True:
        sub = min(left, max_subrecord)
        more = left > sub
        f.write(_marker(-sub if more else sub))
        need = sub
        while need:
            while views[pi].nbytes - off == 0:
                pi, off = pi + 1, 0
            take = min(need, views[pi].nbytes - off)
            f.write(views[pi][off:off + take])
            off += take
            need -= take
        f.write(_marker(sub if first else -sub))
        left -= sub
        first = False
        if not more:
            return


def _read_records(f, path):
    """All records of a gfortran unformatted sequential file, each as one
    uint8 array (subrecords joined)."""
    f.seek(0, 2)
    size = f.tell()
    f.seek(0)
    recs = []
    while f.tell() < size:
        pieces, first = [], True
        while True:
            at = f.tell()
            head = f.read(4)
            if len(head) < 4:
                raise ValueError(f"{path}: truncated record marker at byte {at}")
            lead = int(np.frombuffer(head, "<i4")[0])
            ln = abs(lead)
            if at + 8 + ln > size:
                raise ValueError(f"{path}: broken record framing at byte {at}")
            buf = np.empty(ln, dtype=np.uint8)
            if f.readinto(memoryview(buf)) != ln:
                raise ValueError(f"{path}: short read at byte {at}")
            trail = int(np.frombuffer(f.read(4), "<i4")[0])
            if abs(trail) != ln or (trail < 0) == first:
                raise ValueError(f"{path}: broken record framing at byte {at}")
            pieces.append(buf)
            first = False
            if lead >= 0:
                break
        recs.append(pieces[0] if len(pieces) == 1 else np.concatenate(pieces))
    return recs


def _halo(interior: np.ndarray | None, n) -> np.ndarray:
    """A halo'd Fortran-order array with the interior filled (ghosts 0)."""
    H = [v + 2 * ONX for v in n]
    a = np.zeros(H, dtype="<f8", order="F")
    if interior is not None:
        inner = tuple(slice(ONX, ONX + v) for v in n)
        a[inner] = np.asarray(interior, dtype=np.float64).reshape(n, order="F")
    return a


def write_restart(path: str, n, itt: int, E, dE=None, meanfl=None, ritt: bool = True,
                  max_subrecord: int = MAX_SUBRECORD) -> None:
    """Write a restart file. E, dE, meanfl: interior arrays (Fortran order,
    flat or shaped); dE/meanfl default to zeros.  max_subrecord: the
    subrecord length limit (gfortran's by default)."""
    n = tuple(int(v) for v in n)
    ints = lambda v: [np.array(v, dtype="<i4")]
    with open(path, "wb") as f:
        _record(f, ints(list(n) + [int(itt)]), max_subrecord)
        _record(f, ints([1, n[0], 1, n[1], 1, n[2]]), max_subrecord)
        _record(f, ints([1, n[3], 1, n[4], 1, n[5]]), max_subrecord)
        # one halo'd array at a time, as flat Fortran-order views
        _record(f, [_halo(a, n).reshape(-1, order="F") for a in (E, dE, meanfl)], max_subrecord)
    if ritt:
        with open(os.path.join(os.path.dirname(os.path.abspath(path)), "ritt"), "w") as f:
            f.write(f" {int(itt):11d}\n")


def read_restart(path: str, n=None) -> dict:
    """Read a restart file; with n given, the dimensions must match (as
    read_restart requires). Returns n, itt, the PE bounds and the interior
    E, dE, meanfl (flat, Fortran order)."""
    with open(path, "rb") as f:
        recs = _read_records(f, path)
    if len(recs) != 4:
        raise ValueError(f"{path}: {len(recs)} records, expected 4")
    head = np.frombuffer(recs[0].tobytes(), "<i4")
    if head.size != 7:
        raise ValueError(f"{path}: first record has {head.size} integers, expected 7")
    dims, itt = tuple(int(v) for v in head[:6]), int(head[6])
    if n is not None and tuple(n) != dims:
        raise ValueError(f"read from restart dimensions {dims} do not match dimensions {tuple(n)}")
    bounds = [int(v) for r in recs[1:3] for v in np.frombuffer(r.tobytes(), "<i4")]
    expect = [1, dims[0], 1, dims[1], 1, dims[2], 1, dims[3], 1, dims[4], 1, dims[5]]
    if bounds != expect:
        raise ValueError(f"read from restart PE boundaries {bounds} do not match {expect}")
    H = [v + 2 * ONX for v in dims]
    nh = int(np.prod(H))
    big = recs[3].view("<f8") if recs[3].nbytes % 8 == 0 else np.zeros(0)
    if big.size != 3 * nh:
        raise ValueError(f"{path}: field record holds {big.size} values, expected {3 * nh}")
    inner = tuple(slice(ONX, ONX + v) for v in dims)
    out = {"n": dims, "itt": itt, "bounds": bounds}
    for k, name in enumerate(("E", "dE", "meanfl")):
        out[name] = big[k * nh:(k + 1) * nh].reshape(H, order="F")[inner].ravel(order="F").copy()
    return out


def save(model: "rt6d.Model", path: str, itt: int, ritt: bool = True) -> None:
    """write_restart for a device context: E is the current state; dE and
    meanfl are the last step's by-products (zeros before the first step)."""
    E = model.download("E")
    try:
        dE, mf = model.download("dE"), model.download("meanfl")
    except rt6d.RT6DError:
        dE = mf = None
    write_restart(path, model.n, itt, E, dE, mf, ritt=ritt)


def load(model: "rt6d.Model", path: str) -> int:
    """read_restart for a device context: uploads E and returns itt."""
    r = read_restart(path, model.n)
    model.upload_E(r["E"])
    return r["itt"]
