"""The oracle's diagnose (oracle/rt6d_oracle.c or_diagnose) against an
independent numpy restatement of for_src/diagnose.f90:45-313, evaluated on
the fields the compiled reference dumped after one integrate (tests/golden).

diagnose.f90 needs netCDF and does not build here, so these are the pins of
the quantities only it defines:
  * the velocity maxima umax..mmax over the interior index range 1..n of the
    face arrays (:46-68), and mfl_max with its real*4 1e-12 (:54);
  * eint, mflint and the twelve boundary-flux sums, as sequential sums in
    the reference's loop order (:46-68, :102-277): numpy's add.accumulate is
    a sequential scan, so the sums are reproduced bit for bit;
  * the 3D E_int, meanfl_int (:84-98) and diss_ks..diss_me fields
    (:180-275), accumulated over the wavenumber axes in loop order;
  * the CFL numbers (:294-313) and the residual (:327-328).
Each restated formula follows the Fortran text, not the C oracle, and
planted slips (face 0 in the maxima, the wrong k face in diss_ks, a dropped
spacing in a boundary sum) must change the result.
"""
import numpy as np
import pytest

from conftest import golden_case, golden_data, golden_index

import oracle as O
from raytracer_3d_amd import cases

ENTRIES = golden_index()
IDS = [e["stem"] for e in ENTRIES]
EPS_R4 = float(np.float32(1e-12))


def _spacings(g, n):
    """dxt(1..n) .. dmt(1..n) from the golden grids (rows t, dt, u, du; halo 2)."""
    return [g[f"grid_{a}"][1][2:n[d] + 2] for d, a in enumerate(cases.AXES)]


def _faces(g, name, n, d):
    shp = list(n)
    shp[d] += 1
    return g[name].reshape(shp, order="F")


def _cell(g, name, n):
    return g[name].reshape(n, order="F")


def _b(v, d):
    """1D array v broadcast along axis d of a 6D array."""
    shp = [1] * 6
    shp[d] = v.size
    return v.reshape(shp)


def _seqsum(a):
    """Sequential sum in Fortran loop order (first index fastest)."""
    flat = np.asarray(a).ravel(order="F")
    return float(np.add.accumulate(flat)[-1]) if flat.size else 0.0


def restate_diagnose(g, n, dt, slip=None, vel=None):
    """diagnose.f90:45-313 in numpy on reference-dumped fields.
    slip: None, or one planted mistake ('faces0', 'diss_face1', 'bfl_drop_dz')."""
    nx, ny, nz, nk, nl, nm = n
    sp = _spacings(g, n)
    dxt, dyt, dzt, dkt, dlt, dmt = sp
    E, dE, mf = (_cell(g, nm_, n) for nm_ in ("E1", "dE1", "meanfl1"))
    out = {}
    # fxa = dxt(i)*dyt(j)*dzt(k)*dkt(l)*dlt(n)*dmt(m), left to right (:52)
    fxa = _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5)
    out["eint"] = _seqsum(dE * fxa)
    out["mflint"] = _seqsum(mf * fxa)
    out["mfl_max"] = float(np.max(np.abs(mf) / np.maximum(EPS_R4, E)))
    lo = 0 if slip == "faces0" else 1
    vmax = []
    for d, a in enumerate(cases.AXES):
        v = vel[a] if vel is not None else _faces(g, f"{a}dot", n, d)
        sl = [slice(None)] * 6
        sl[d] = slice(lo, n[d] + 1)
        vmax.append(float(np.max(np.abs(v[tuple(sl)]))))
    for k_, v in zip(("umax", "vmax", "wmax", "kmax", "lmax", "mmax"), vmax):
        out[k_] = v
    # E_int, meanfl_int (:84-98): loop m, n, l outermost-to-innermost over the
    # wavenumber axes, for every (i,j,k)
    Ei = np.zeros((nx, ny, nz))
    Mi = np.zeros((nx, ny, nz))
    for m in range(nm):
        for nn in range(nl):
            for l in range(nk):
                Ei = Ei + E[:, :, :, l, nn, m] * dlt[nn] * dmt[m] * dkt[l]
                Mi = Mi + mf[:, :, :, l, nn, m] * dkt[l] * dlt[nn] * dmt[m]
    out["E_int"], out["meanfl_int"] = Ei, Mi
    F = [_faces(g, f"flux_{a}1", n, d) for d, a in enumerate(cases.AXES)]
    # x, y, z boundary sums (:102-170): the face is fixed, the other five
    # spacings multiply left to right, the loops run over the rest in order
    def face(Fd, d, f):
        sl = [slice(None)] * 6
        sl[d] = slice(f, f + 1)
        return Fd[tuple(sl)]
    fx0, fxn = face(F[0], 0, 0), face(F[0], 0, nx)
    out["bfl_xs"] = _seqsum(fx0 * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5))
    out["bfl_xe"] = _seqsum(fxn * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5))
    fy0, fyn = face(F[1], 1, 0), face(F[1], 1, ny)
    out["bfl_ys"] = _seqsum(fy0 * _b(dxt, 0) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5))
    out["bfl_ye"] = _seqsum(fyn * _b(dxt, 0) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5))
    fz0, fzn = face(F[2], 2, 0), face(F[2], 2, nz)
    out["bfl_zs"] = _seqsum(fz0 * _b(dxt, 0) * _b(dyt, 1) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5))
    if slip == "bfl_drop_dz":
        out["bfl_zs"] = _seqsum(fz0 * _b(dxt, 0) * _b(dyt, 1) * _b(dlt, 4) * _b(dmt, 5))
    out["bfl_ze"] = _seqsum(fzn * _b(dxt, 0) * _b(dyt, 1) * _b(dkt, 3) * _b(dlt, 4) * _b(dmt, 5))
    # k, l, m boundary sums and the diss fields (:180-275)
    kf0 = 1 if slip == "diss_face1" else 0
    fk0, fkn = face(F[3], 3, kf0), face(F[3], 3, nk)
    out["bfl_ks"] = _seqsum(fk0 * _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dlt, 4) * _b(dmt, 5))
    out["bfl_ke"] = _seqsum(fkn * _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dlt, 4) * _b(dmt, 5))
    fl0, fln = face(F[4], 4, 0), face(F[4], 4, nl)
    out["bfl_ls"] = _seqsum(fl0 * _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dmt, 5))
    out["bfl_le"] = _seqsum(fln * _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dmt, 5))
    fm0, fmn = face(F[5], 5, 0), face(F[5], 5, nm)
    out["bfl_ms"] = _seqsum(fm0 * _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4))
    out["bfl_me"] = _seqsum(fmn * _b(dxt, 0) * _b(dyt, 1) * _b(dzt, 2) * _b(dkt, 3) * _b(dlt, 4))
    diss = np.zeros((6, nx, ny, nz))
    for m in range(nm):
        for nn in range(nl):
            diss[0] = diss[0] + fk0[:, :, :, 0, nn, m] * dlt[nn] * dmt[m]
            diss[1] = diss[1] + fkn[:, :, :, 0, nn, m] * dlt[nn] * dmt[m]
    for m in range(nm):
        for l in range(nk):
            diss[2] = diss[2] + fl0[:, :, :, l, 0, m] * dkt[l] * dmt[m]
            diss[3] = diss[3] + fln[:, :, :, l, 0, m] * dkt[l] * dmt[m]
    for nn in range(nl):
        for l in range(nk):
            diss[4] = diss[4] + fm0[:, :, :, l, nn, 0] * dkt[l] * dlt[nn]
            diss[5] = diss[5] + fmn[:, :, :, l, nn, 0] * dkt[l] * dlt[nn]
    out["diss"] = diss
    # CFL numbers (:294-313): max over the T cells of vmax*dt/d?t(i)
    for d, a in enumerate(cases.AXES):
        out[f"cfl_{a}"] = float(np.max(vmax[d] * dt / sp[d]))
    b = out
    out["residual"] = (b["eint"] + b["mflint"] + b["bfl_xe"] + b["bfl_ye"] + b["bfl_ze"] + b["bfl_ke"]
                       + b["bfl_le"] + b["bfl_me"] - b["bfl_xs"] - b["bfl_ys"] - b["bfl_zs"]
                       - b["bfl_ks"] - b["bfl_ls"] - b["bfl_ms"])
    return out


SCALARS = (["eint", "mflint", "mfl_max", "umax", "vmax", "wmax", "kmax", "lmax", "mmax"]
           + [f"bfl_{a}{s}" for a in "xyzklm" for s in "se"]
           + [f"cfl_{a}" for a in "xyzklm"] + ["residual"])


def _oracle_after_one_step(entry):
    case = golden_case(entry)
    o = O.Oracle(case)
    o.calc_parameter()
    o.integrate(1)
    return case, o


@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_oracle_diagnose_bitwise_against_restatement(entry):
    case, o = _oracle_after_one_step(entry)
    g = golden_data(entry)
    od, o3 = o.diagnose(fields3d=True)
    ref = restate_diagnose(g, tuple(case.n), case.dt)
    for k in SCALARS:
        assert od[k] == ref[k], (k, od[k], ref[k])
    n3 = tuple(case.n[:3])
    assert np.array_equal(o3["E_int"].reshape(n3, order="F"), ref["E_int"])
    assert np.array_equal(o3["meanfl_int"].reshape(n3, order="F"), ref["meanfl_int"])
    for q in range(6):
        assert np.array_equal(o3["diss"][q].reshape(n3, order="F"), ref["diss"][q]), q


@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_velocity_maxima_exclude_face_zero(entry):
    """umax..mmax run over i = 1..n of the face arrays (diagnose.f90:46-68).
    With face 0 of every active axis made the largest, the oracle must still
    equal the restatement over 1..n and must differ from a slip that
    includes face 0."""
    case, o = _oracle_after_one_step(entry)
    g = golden_data(entry)
    n = tuple(case.n)
    vel = {}
    for d, a in enumerate(cases.AXES):
        v = _faces(g, f"{a}dot", n, d).copy()
        sl = [slice(None)] * 6
        sl[d] = 0
        v[tuple(sl)] = 10.0 * (np.max(np.abs(v)) + 1.0)
        vel[a] = v
        o.set(f"{a}dot", v.ravel(order="F"), faces=True)
    od = o.diagnose()
    ref = restate_diagnose(g, n, case.dt, vel=vel)
    bad = restate_diagnose(g, n, case.dt, slip="faces0", vel=vel)
    for a, k in zip(cases.AXES, ("umax", "vmax", "wmax", "kmax", "lmax", "mmax")):
        assert od[k] == ref[k], k
        assert bad[k] != od[k], k
    for a in cases.AXES:
        assert od[f"cfl_{a}"] == ref[f"cfl_{a}"]


@pytest.mark.parametrize("slip,what", [("diss_face1", "diss"), ("bfl_drop_dz", "bfl_zs")])
def test_planted_diagnose_slips_are_caught(slip, what):
    entry = next(e for e in ENTRIES if e["stem"] == "random_all_axes")
    case, o = _oracle_after_one_step(entry)
    g = golden_data(entry)
    od, o3 = o.diagnose(fields3d=True)
    bad = restate_diagnose(g, tuple(case.n), case.dt, slip=slip)
    if what == "diss":
        n3 = tuple(case.n[:3])
        assert not np.array_equal(o3["diss"][0].reshape(n3, order="F"), bad["diss"][0])
        assert od["bfl_ks"] != bad["bfl_ks"]
    else:
        assert od[what] != bad[what]


@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_golden_dt_is_the_cfl_rule_on_reference_velocities(entry):
    """Each fixture's dt is cfl_dt_from (the rule rt6d_cfl_dt applies on the
    device) at CFL 0.5 on the reference's own velocities."""
    case = golden_case(entry)
    g = golden_data(entry)
    vel = {f"{a}dot": g[f"{a}dot"] for a in cases.AXES}
    assert O.cfl_dt_from(vel, case.n, case.spacing, 0.5) == float(g["dt"])
