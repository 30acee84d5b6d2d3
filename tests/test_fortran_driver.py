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
    r = subprocess.run([DRIVER, cf, of, "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
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
    r = subprocess.run([DRIVER, cf, of, "0", "fp32"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert rel_l2(np.fromfile(of, dtype="<f8"), g["EN"]) <= 1e-5
