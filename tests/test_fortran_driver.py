"""The Fortran host side (raytracer_3d_amd/fortran): the reference's driver
order and routine names, calling the GPU through iso_c_binding."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_case, golden_data, golden_index, rel_l2

from raytracer_3d_amd import cases

DRIVER = os.path.join(ROOT, "raytracer_3d_amd", "fortran", "rt6d_driver")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (amdflang absent)")


@needs_driver
def test_driver_without_device_fails_cleanly(tmp_path):
    from raytracer_3d_amd import rt6d
    if rt6d.device_count() > 0:
        pytest.skip("a device is present")
    entry = golden_index()[0]
    case = golden_case(entry)
    cf = str(tmp_path / "c.case")
    cases.write_case_file(case, cf, nsteps=1)
    r = subprocess.run([DRIVER, cf], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "no HIP device" in r.stdout + r.stderr


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("entry", golden_index(), ids=lambda e: e["stem"])
def test_driver_matches_reference(entry, tmp_path, gpu_available):
    case = golden_case(entry)
    g = golden_data(entry)
    cf = str(tmp_path / "c.case")
    of = str(tmp_path / "E.out")
    cases.write_case_file(case, cf, nsteps=int(g["nsteps"]))
    r = subprocess.run([DRIVER, cf, "out=" + of, "snap=2"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "snap3D.cdf").exists() and (tmp_path / "snap6D.cdf").exists()
    E = np.fromfile(of, dtype="<f8")
    assert rel_l2(E, g["EN"]) <= 1e-12
    res = [float(x) for x in re.findall(r"residual :\s*(\S+)", r.stdout)]
    tot = [float(x) for x in re.findall(r"total dE\s*:\s*(\S+)", r.stdout)]
    assert res and len(res) == len(tot)
    for a, b in zip(res, tot):
        assert abs(a) <= 1e-6 * (abs(b) + 1e-300) or abs(a) < 1e-25


@needs_driver
@pytest.mark.gpu
def test_driver_fp32_form(tmp_path, gpu_available):
    entry = golden_index()[-1]
    case = golden_case(entry)
    g = golden_data(entry)
    cf = str(tmp_path / "c.case")
    of = str(tmp_path / "E.out")
    cases.write_case_file(case, cf, nsteps=int(g["nsteps"]))
    r = subprocess.run([DRIVER, cf, "out=" + of, "fp32"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert rel_l2(np.fromfile(of, dtype="<f8"), g["EN"]) <= 1e-5


def _fortran_records(path):
    """Records of a Fortran unformatted sequential file (4-byte length markers)."""
    data = open(path, "rb").read()
    out, i = [], 0
    while i < len(data):
        n = int(np.frombuffer(data, "<i4", 1, i)[0])
        out.append(data[i + 4:i + 4 + n])
        assert int(np.frombuffer(data, "<i4", 1, i + 4 + n)[0]) == n
        i += 8 + n
    return out


@needs_driver
@pytest.mark.gpu
def test_driver_restart_resumes_bit_for_bit(tmp_path, gpu_available):
    """write_restart/read_restart in the record layout of for_src/restart.f90:
    3 steps + restart + 2 steps equals 5 steps, and the file holds the
    dimensions, itt, the PE bounds and the halo'd E, dE, meanfl."""
    from raytracer_3d_amd import rt6d
    entry = golden_index()[0]
    case = golden_case(entry)
    n = case.n
    cf3, cf2 = str(tmp_path / "c3.case"), str(tmp_path / "c2.case")
    cases.write_case_file(case, cf3, nsteps=3)
    cases.write_case_file(case, cf2, nsteps=2)
    r = subprocess.run([DRIVER, cf3, "out=E3.out", "restart"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rec = _fortran_records(tmp_path / "restart_pe=000000000.dta")
    assert len(rec) == 4
    assert list(np.frombuffer(rec[0], "<i4")) == list(n) + [3]
    assert list(np.frombuffer(rec[1], "<i4")) == [1, n[0], 1, n[1], 1, n[2]]
    assert list(np.frombuffer(rec[2], "<i4")) == [1, n[3], 1, n[4], 1, n[5]]
    H = [v + 4 for v in n]
    big = np.frombuffer(rec[3], "<f8")
    assert big.size == 3 * int(np.prod(H))
    E_h, dE_h = (big[i * int(np.prod(H)):(i + 1) * int(np.prod(H))].reshape(H, order="F") for i in range(2))
    inner = tuple(slice(2, 2 + v) for v in n)
    E3 = np.fromfile(tmp_path / "E3.out", dtype="<f8")
    assert np.array_equal(E_h[inner].ravel(order="F"), E3)
    with rt6d.Model.from_case(case) as m:
        m.calc_parameter()
        m.integrate(3)
        assert np.array_equal(E3, m.download("E"))
        assert np.array_equal(dE_h[inner].ravel(order="F"), m.download("dE"))
        m.integrate(2)
        E5 = m.download("E")
    r = subprocess.run([DRIVER, cf2, "out=E5.out", "restart"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "reading restart file" in r.stdout
    assert np.array_equal(np.fromfile(tmp_path / "E5.out", dtype="<f8"), E5)
    assert list(np.frombuffer(_fortran_records(tmp_path / "restart_pe=000000000.dta")[0], "<i4"))[-1] == 5


@needs_driver
@pytest.mark.gpu
def test_restart_files_cross_between_fortran_and_python(tmp_path, gpu_available):
    """A restart file written by the Fortran driver reads in Python
    (raytracer_3d_amd.restart) and resumes a device context; one written from
    Python resumes the Fortran driver. Both continuations equal the
    uninterrupted run bit for bit."""
    from raytracer_3d_amd import restart, rt6d
    entry = golden_index()[0]
    case = golden_case(entry)
    cf3, cf2 = str(tmp_path / "c3.case"), str(tmp_path / "c2.case")
    cases.write_case_file(case, cf3, nsteps=3)
    cases.write_case_file(case, cf2, nsteps=2)
    with rt6d.Model.from_case(case) as m:
        m.calc_parameter()
        m.integrate(3)
        E3 = m.download("E")
        m.integrate(2)
        E5 = m.download("E")
    # Fortran writes, Python reads and resumes
    r = subprocess.run([DRIVER, cf3, "out=E3.out", "restart"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    f = tmp_path / restart.restart_file_name(0)
    got = restart.read_restart(str(f), case.n)
    assert got["itt"] == 3 and np.array_equal(got["E"], E3)
    with rt6d.Model.from_case(case) as m:
        m.calc_parameter()
        assert restart.load(m, str(f)) == 3
        m.integrate(2)
        assert np.array_equal(m.download("E"), E5)
        restart.save(m, str(f), 5)      # Python writes at step 5 over the Fortran file
    got = restart.read_restart(str(f), case.n)
    assert got["itt"] == 5 and np.array_equal(got["E"], E5)
    # Python writes at step 3, Fortran reads and runs the last 2 steps
    restart.write_restart(str(f), case.n, 3, E3)
    r = subprocess.run([DRIVER, cf2, "out=E5.out", "restart"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "reading restart file" in r.stdout
    assert np.array_equal(np.fromfile(tmp_path / "E5.out", dtype="<f8"), E5)


@needs_driver
@pytest.mark.gpu
def test_driver_snapshots_equal_the_python_hosts(tmp_path, gpu_available):
    """With diagnose every 3 steps the Fortran driver writes snap3D.cdf and
    snap6D.cdf (fortran/rt6d_netcdf.f90); the Python host's time loop
    (raytracer_3d_amd.driver) writes the same files byte for byte, and they
    read back with scipy in the reference's schema."""
    from scipy.io import netcdf_file
    from raytracer_3d_amd import driver
    entry = golden_index()[0]
    case = golden_case(entry)
    cf = str(tmp_path / "c.case")
    cases.write_case_file(case, cf, nsteps=7)
    fdir, pdir = tmp_path / "fortran", tmp_path / "python"
    fdir.mkdir()
    r = subprocess.run([DRIVER, cf, "snap=3"], capture_output=True, text=True, timeout=120, cwd=fdir)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "time steps in file snap3D.cdf :      3" in r.stdout
    res = driver.run(case, nsteps=7, snapint=3 * case.dt * (1 + 1e-9), directory=str(pdir))
    assert res["snapshots"] == 3
    for name in ("snap3D.cdf", "snap6D.cdf"):
        a, b = (fdir / name).read_bytes(), (pdir / name).read_bytes()
        assert len(a) == len(b), name
        assert a == b, f"{name}: first difference at byte {next(i for i in range(len(a)) if a[i] != b[i])}"
    with netcdf_file(str(fdir / "snap3D.cdf"), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["Time"][:], np.array([0, 3, 6]) * case.dt)
        assert f.variables["E"].dimensions == ("Time", "zt", "yt", "xt")
    with netcdf_file(str(fdir / "snap6D.cdf"), "r", mmap=False) as f:
        assert np.array_equal(f.variables["E"][2].reshape(-1), res["E"])
