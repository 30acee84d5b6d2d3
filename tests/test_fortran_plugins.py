"""Config plug-ins of the reference form through the Fortran host.

* every reference plug-in that needs no netCDF (for_config/gm1.f90,
  test1D.f90, test2D.f90, brava.f90, reebus_plot1.f90) compiles and links
  unchanged against this package's main_module and drop-in routines
  (CPU, only where the reference tree is present);
* a plug-in of the same form (tests/plugins/file_plugin.f90: the four hooks
  over main_module) drives the golden cases through the GPU and matches the
  reference's vectors; its set_forcing changes E on the host, as
  for_config/lee1.f90 and critical_gm1.f90 do (the latter reading the host
  zdot), and the result equals the device forcing hooks; restart files with
  records split into gfortran subrecords cross between the hosts.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_case, golden_data, golden_index, rel_l2

from raytracer_3d_amd import cases

REF_CONFIG = "/root/reference/for_config"
FDIR = os.path.join(ROOT, "raytracer_3d_amd", "fortran")
PLUGIN_EXE = os.path.join(ROOT, "tests", "plugins", "_build", "model_file_plugin.x")
have_flang = shutil.which("amdflang") is not None or os.path.exists("/opt/rocm/bin/amdflang")
needs_plugin = pytest.mark.skipif(not os.path.exists(PLUGIN_EXE), reason="test plug-in driver not built")


@pytest.mark.skipif(not (os.path.isdir(REF_CONFIG) and have_flang), reason="reference tree or amdflang absent")
@pytest.mark.parametrize("name", ["gm1", "test1D", "test2D", "brava", "reebus_plot1", "lee1", "critical_gm1"])
def test_reference_plugin_compiles_unchanged(name, tmp_path):
    src = os.path.join(REF_CONFIG, name + ".f90")
    if name in ("lee1", "critical_gm1"):
        # their netCDF writers (lee1.f90:146-189, critical_gm1.f90:148-191)
        # include netcdf.inc, which this image does not have
        assert 'include "netcdf.inc"' in open(src).read()
        pytest.skip("needs netcdf.inc (no netCDF library in this image)")
    mpi = "1" if name == "brava" else "0"
    r = subprocess.run(["make", "-s", "-C", FDIR, "plugin", f"PLUGIN={src}", f"PLUGIN_DIR={tmp_path}", f"MPI={mpi}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / f"model_{name}.x").exists()


def _write_plugin_inputs(d, case, nsteps, snapsteps=0, forcing=0, field=None):
    n = case.n
    with open(d / "plugin.txt", "w") as f:
        f.write(" ".join(map(str, n)) + "\n")
        f.write(f"{nsteps} {snapsteps} {forcing}\n")
        f.write(" ".join(repr(float(v)) for v in (case.dt, case.f0, case.sign) + tuple(case.fixed)) + "\n")
        f.write(f"{int(case.reflect_up)} {int(case.reflect_down)}\n")
    np.concatenate([np.asarray(s, dtype="<f8") for s in case.spacing]).tofile(d / "spacing.bin")
    np.concatenate([np.asarray(a, dtype="<f8") for a in (case.Nsqr, case.U, case.V)]).tofile(d / "background.bin")
    np.asarray(case.E0, dtype="<f8").tofile(d / "E0.bin")
    if field is not None:
        np.asarray(field, dtype="<f8").tofile(d / "forcing.bin")


def _run(d, *args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([PLUGIN_EXE, *args], capture_output=True, text=True, timeout=120, cwd=d, env=e)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@needs_plugin
@pytest.mark.gpu
@pytest.mark.parametrize("entry", golden_index(), ids=lambda e: e["stem"])
def test_plugin_driver_matches_reference(entry, tmp_path, gpu_available):
    case = golden_case(entry)
    g = golden_data(entry)
    _write_plugin_inputs(tmp_path, case, int(g["nsteps"]), snapsteps=2)
    out = _run(tmp_path, "out=E.out")
    assert f"to {int(g['nsteps']):8d}" in out
    assert out.count("residual :") == 3          # steps 0, 2 and 4
    assert (tmp_path / "snap3D.cdf").exists() and (tmp_path / "snap6D.cdf").exists()
    assert rel_l2(np.fromfile(tmp_path / "E.out", dtype="<f8"), g["EN"]) <= 1e-12


@needs_plugin
@pytest.mark.gpu
@pytest.mark.parametrize("kind", [1, 2], ids=["flux_injection", "relaxation"])
def test_host_set_forcing_equals_device_forcing(kind, tmp_path, gpu_available):
    """set_forcing on the host E (the driver's default host_forcing: E to the
    host before it, back to the device after it) gives the device hook's
    result (rt6d_set_forcing, the same statements on the device)."""
    from raytracer_3d_amd import rt6d
    entry = next(e for e in golden_index() if e["stem"] == "gm_small")
    case = golden_case(entry)
    n = case.n
    n5 = n[0] * n[1] * n[3] * n[4] * n[5]
    u = cases.uniform(5, kind, np.arange(n5, dtype=np.uint64))
    # kind 1: a bottom flux that adds up to 1% of max E per step; kind 2: a
    # relaxation target of the order of E
    dz1 = float(case.spacing[2][0])
    field = (0.01 * np.max(case.E0) * dz1 / case.dt * u) if kind == 1 else (np.max(case.E0) * u)
    _write_plugin_inputs(tmp_path, case, 6, forcing=kind, field=field)
    _run(tmp_path, "out=E.out")
    got = np.fromfile(tmp_path / "E.out", dtype="<f8")
    with rt6d.Model.from_case(case) as m:
        m.calc_parameter()
        m.set_forcing(kind, field)
        m.integrate(6)
        ref = m.download("E")
    with rt6d.Model.from_case(case) as m:
        m.calc_parameter()
        m.integrate(6)
        unforced = m.download("E")
    assert rel_l2(got, unforced) > 1e-10, "the forcing had no effect"
    assert rel_l2(got, ref) <= 1e-14


@needs_plugin
@pytest.mark.gpu
def test_restart_subrecords_cross_between_hosts(tmp_path, gpu_available):
    """With the subrecord limit set to 1000 bytes the Fortran host splits the
    field record as gfortran does; Python reads it, and a file Python writes
    with another limit resumes the Fortran host: 3 + 2 steps equal 5."""
    from raytracer_3d_amd import restart, rt6d
    entry = next(e for e in golden_index() if e["stem"] == "gm_small")
    case = golden_case(entry)
    with rt6d.Model.from_case(case) as m:
        m.calc_parameter()
        m.integrate(3)
        E3 = m.download("E")
        m.integrate(2)
        E5 = m.download("E")
    _write_plugin_inputs(tmp_path, case, 3)
    _run(tmp_path, "out=E3.out", "restart", env={"RT6D_RESTART_MAX_SUBRECORD": "1000"})
    f = tmp_path / restart.restart_file_name(0)
    data = f.read_bytes()
    assert np.frombuffer(data[:4], "<i4")[0] == 28
    lead = np.frombuffer(data[100:104], "<i4")[0]   # after the three header records (36 + 32 + 32 bytes)
    assert lead == -1000
    got = restart.read_restart(str(f), case.n)
    assert got["itt"] == 3 and np.array_equal(got["E"], E3)
    restart.write_restart(str(f), case.n, 3, E3, max_subrecord=777)
    _write_plugin_inputs(tmp_path, case, 2)
    out = _run(tmp_path, "out=E5.out", "restart")
    assert "reading restart file" in out
    assert np.array_equal(np.fromfile(tmp_path / "E5.out", dtype="<f8"), E5)
