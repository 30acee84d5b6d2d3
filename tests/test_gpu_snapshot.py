"""Snapshot files written from a device run (the reference's main loop:
integrate, then diagnose every snapint, for_src/main.f90:64-78).

The 3D record's E must be the wavenumber integral of the 6D record's E,
E_int(i,j,k) = sum E(i,j,k,l,n,m) dkt(l) dlt(n) dmt(m)
(for_src/diagnose.f90:84-100), both read back from the files; the 6D
record must equal the device state at the time it was written.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from raytracer_3d_amd import cases, rt6d, snapshot

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [rt6d.FP64, rt6d.FP32], ids=["fp64", "fp32"])
def test_snapshots_from_device_run(tmp_path, precision, gpu_available):
    case = cases.packet_case(n=(8, 4, 6, 4, 4, 6), name="packet_snap")
    with rt6d.Model.from_case(case, precision=precision) as m:
        m.calc_parameter()
        m.dt = m.cfl_dt(0.5)
        snaps = snapshot.Snapshots(m, case, directory=str(tmp_path), write_6d=True)
        every = 3
        E_written = []
        budgets = []
        for itt in range(7):
            m.integrate(1)
            if itt % every == 0:
                budgets.append(snaps.diagnose(itt))
                E_written.append(m.download("E"))
        xdot = m.download("xdot")
    with netcdf_file(str(tmp_path / "snap6D.cdf"), "r", mmap=False) as f6, \
            netcdf_file(str(tmp_path / "snap3D.cdf"), "r", mmap=False) as f3:
        np.testing.assert_allclose(f6.variables["Time"][:], [0.0, 3 * m.dt, 6 * m.dt], rtol=0)
        np.testing.assert_array_equal(f3.variables["Time"][:], f6.variables["Time"][:])
        E6 = f6.variables["E"][:]
        for r, E in enumerate(E_written):
            assert np.array_equal(E6[r].reshape(-1), E)
        dk, dl, dm = (f6.variables[v][:] for v in ("dkt", "dlt", "dmt"))
        w = dm[:, None, None] * dl[None, :, None] * dk[None, None, :]    # (m, l, k)
        E3 = f3.variables["E"][:]
        for r in range(3):
            Eint = np.tensordot(w, E6[r], axes=([0, 1, 2], [0, 1, 2]))   # (z, y, x)
            scale = np.max(np.abs(Eint))
            assert scale > 0
            assert np.max(np.abs(E3[r] - Eint)) <= 1e-12 * scale
            assert f3.variables["mflint"][r] == budgets[r]["mflint"]
            assert f3.variables["bflux_xs"][r] == budgets[r]["bfl_xs"]
        X = f6.variables["xdot"][:]
        faces = xdot.reshape((9, 4, 6, 4, 4, 6), order="F")
        assert np.array_equal(X, faces[1:].T)
        assert np.array_equal(f6.variables["omega_i"][:].reshape(-1), m_omega(case, precision))


def m_omega(case, precision):
    with rt6d.Model.from_case(case, precision=precision) as m:
        m.calc_parameter()
        return m.download("omega_i")
