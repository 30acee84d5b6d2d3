"""rt6d_diagnose_noalloc (the diagnose entry point that allocates nothing)
against rt6d_diagnose, through the C ABI: the same scalars and 3D fields bit
for bit, on two calls in a row where the second has a smaller grid (the
shared per-GPU scratch keeps the first call's contents), in both
precisions; the size limit of its 3D staging is an error, not a fault."""
import numpy as np
import pytest

from raytracer_3d_amd import cases, rt6d

pytestmark = pytest.mark.gpu


def _same(a, b):
    sa, fa = a
    sb, fb = b
    assert sa.keys() == sb.keys()
    for k in sa:
        assert sa[k] == sb[k] or (np.isnan(sa[k]) and np.isnan(sb[k])), (k, sa[k], sb[k])
    for k in ("E_int", "meanfl_int", "diss"):
        assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("precision", [rt6d.FP64, rt6d.FP32], ids=["fp64", "fp32"])
def test_noalloc_diagnose_equals_diagnose(precision, gpu_available):
    big = cases.gm_case(n=(24, 24, 24, 16, 16, 16))     # 221,184 blocks of cells, 13,824 columns
    small = cases.random_case((6, 5, 4, 4, 3, 4), seed=41)
    for case in (big, small):                           # the second call is smaller
        with rt6d.Model.from_case(case, precision=precision) as m:
            m.calc_parameter()
            dt = m.cfl_dt(0.5)
            m.integrate(3, dt=dt)
            bytes_before = m.query(rt6d.Q_DEVICE_BYTES)
            got = m.diagnose(dt=dt, fields3d=True, noalloc=True)
            assert m.query(rt6d.Q_DEVICE_BYTES) == bytes_before, "the no-alloc diagnose allocated"
            ref = m.diagnose(dt=dt, fields3d=True)
            _same(got, ref)
            assert m.diagnose(dt=dt, noalloc=True) == ref[0]   # scalars only
            m.integrate(1, dt=dt)
            _same(m.diagnose(dt=dt, fields3d=True, noalloc=True), m.diagnose(dt=dt, fields3d=True))


def test_noalloc_diagnose_limits(gpu_available):
    c = cases.random_case((48, 40, 36, 2, 2, 2), seed=42)   # 69,120 columns > 65,536
    with rt6d.Model.from_case(c) as m:
        m.calc_parameter()
        dt = m.cfl_dt(0.5)
        m.integrate(1, dt=dt)
        with pytest.raises(rt6d.RT6DError, match="exceeds"):
            m.diagnose(dt=dt, fields3d=True, noalloc=True)
        assert m.diagnose(dt=dt, noalloc=True) == m.diagnose(dt=dt)   # scalars have no size limit
        m.superbee_fluxes(dt=2 * dt)
        with pytest.raises(rt6d.RT6DError, match="coefficients"):
            m.diagnose(dt=dt, noalloc=True)
    with rt6d.Model.from_case(cases.random_case((4, 4, 4, 4, 4, 4), seed=43)) as m:
        m.calc_parameter()
        with pytest.raises(rt6d.RT6DError, match="no step"):
            m.diagnose(dt=1.0, noalloc=True)
