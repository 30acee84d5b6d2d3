"""The Python host's time loop (raytracer_3d_amd.driver): main.f90's order
and cadence with snapshots and restart files.

Snapshots are written after integrate when mod(itt, int(snapint/dt)) == 0 or
itt == 0 (for_src/main.f90:74); Time = itt * dt. A run split by a restart
equals the uninterrupted run bit for bit, and the restarted run's snapshot
file starts afresh, as the reference's init_diag re-creates it.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from raytracer_3d_amd import cases, driver, restart, rt6d

pytestmark = pytest.mark.gpu


def _case():
    c = cases.packet_case(n=(8, 4, 6, 4, 4, 6), name="pydriver")
    with rt6d.Model.from_case(c) as m:
        m.calc_parameter()
        c.dt = m.cfl_dt(0.5)
    return c


def test_cadence_snapshots_and_final_state(tmp_path, gpu_available):
    c = _case()
    res = driver.run(c, nsteps=7, snapint=3 * c.dt + 1e-9 * c.dt, directory=str(tmp_path))
    assert res["itt"] == 7 and res["steps"] == 7 and res["snapshots"] == 3
    with rt6d.Model.from_case(c) as m:
        m.calc_parameter()
        m.integrate(7, dt=c.dt)
        assert np.array_equal(res["E"], m.download("E"))
    with netcdf_file(str(tmp_path / "snap3D.cdf"), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["Time"][:], np.array([0, 3, 6]) * c.dt)
    with netcdf_file(str(tmp_path / "snap6D.cdf"), "r", mmap=False) as f:
        assert f.variables["E"].shape[0] == 3
        assert np.array_equal(f.variables["E"][2].reshape(-1), res["E"])   # the last snapshot is after step 6


def test_restart_splits_the_run(tmp_path, gpu_available):
    c = _case()
    full = driver.run(c, nsteps=7, directory=str(tmp_path / "full"))
    d = str(tmp_path / "split")
    a = driver.run(c, nsteps=4, snapint=3 * c.dt * (1 + 1e-9), directory=d, use_restart=True)
    assert a["itt"] == 4 and a["snapshots"] == 2            # itt 0 and 3
    r = restart.read_restart(f"{d}/{restart.restart_file_name(0)}", c.n)
    assert r["itt"] == 4
    b = driver.run(c, nsteps=3, snapint=3 * c.dt * (1 + 1e-9), directory=d, use_restart=True)
    assert b["itt"] == 7 and b["snapshots"] == 1            # itt 6 only
    assert np.array_equal(b["E"], full["E"])
    with netcdf_file(f"{d}/snap3D.cdf", "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["Time"][:], [6 * c.dt])
