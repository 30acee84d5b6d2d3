"""The FP32-form error report (raytracer_3d_amd.precision) on a small case:
every quantity present and inside the FP32 form's stated bounds."""
import numpy as np
import pytest

from raytracer_3d_amd import cases, precision

pytestmark = pytest.mark.gpu


def test_fp32_report_small(gpu_available):
    c = cases.gm_case(n=(8, 8, 8, 6, 6, 6))
    rep = precision.fp32_report(c, steps=30, checkpoints=[1, 10, 30])
    assert [r["step"] for r in rep["checkpoints"]] == [1, 10, 30]
    for r in rep["checkpoints"]:
        assert r["E"]["rel_l2"] <= 1e-5
        assert abs(r["total_energy_drift"]) <= 1e-5
        assert abs(r["budget_residual_rel_fp64"]) <= 1e-11
        assert abs(r["budget_residual_rel_fp32"]) <= 1e-4
        assert r["faces"] > 0 and 0 <= r["flips"] <= r["faces"]
    for f in precision.FIELDS_END:
        assert np.isfinite(rep["end_fields"][f]["rel_l2"])
    assert rep["end_fields"]["dE"]["rel_l2"] <= 1e-3
