"""Host-side pieces of bench.py (no GPU): the kernel names it reports must be
the demangled names of the instantiations that the launchers in
raytracer_3d_amd/csrc/rt6d_kernels.h start, so that the roofline line and the
rocprofv3 summaries name the same kernel."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("precision,option,uxy,handover,expected", [
    ("double", 22, True, True, "k_step_colx2_w4<double, false, 2, true, true, true, false, true>"),
    ("double", 22, True, False, "k_step_colx2_w4<double, false, 2, true, true, true, false, false>"),
    ("float", 22, True, False, "k_step_colx2<float, false, 2, true, true, true, false, false, true>"),
    ("float", 22, True, True, "k_step_colx2<float, false, 2, true, true, true, false, true, true>"),
    ("float", 21, True, False, "k_step_colx2<float, false, 1, true, true, true, false, false, false>"),
    ("double", 22, False, False, "k_step_colx2<double, false, 2, true, true, false, false, false, false>"),
    ("double", 16, True, False, "k_step_colx2<double, false, 2, true, false, false, false, false, false>"),
    ("double", 10, True, False, "k_step_col<double, false, false, 2, true, 1>"),
])
def test_step_kernel_names(precision, option, uxy, handover, expected):
    assert bench._step_kernel(precision, option, uxy, handover) == expected


def test_profiles_name_the_reported_kernels():
    """The committed default-step summaries carry the names bench reports."""
    prof = os.path.join(ROOT, "profiles")
    for fname, name in [
        ("k_step_colx2wl_uxy_fp64_homog_gm24.txt", bench._step_kernel("double", 22, True, True)),
        ("k_step_colx2wl_uxy_fp32_homog_gm24.txt", bench._step_kernel("float", 22, True, False)),
        ("k_step_colx2wl_uxy_fp32_homog_gm32.txt", bench._step_kernel("float", 22, True, True)),
        ("k_step_col_fp64_general_gm24.txt", bench._step_kernel("double", 10, True, False)),
    ]:
        with open(os.path.join(prof, fname)) as f:
            text = f.read()
        assert name in text, (fname, name)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert re.search(r"one GPU only", r.stderr)
