"""Bottom forcing hooks on the device (rt6d_set_forcing), the set_forcing
routines of the reference's config plug-ins applied before every step.

The reference's plug-in files (for_config/lee1.f90, critical_gm1.f90) need
netCDF and do not build here, so these hooks are pinned by their statements
restated in numpy (left-to-right binary64, as the Fortran is written)
combined with the oracle's integrate: forcing then step, repeated.
"""
import numpy as np
import pytest

from conftest import rel_l2

import oracle as O
from raytracer_3d_amd import cases, rt6d

pytestmark = pytest.mark.gpu


def _numpy_forcing(E, kind, F, zdot_faces, n, dt, dz1):
    """E (interior, Fortran order) after one set_forcing call."""
    nx, ny, nz, nk, nl, nm = n
    E6 = E.reshape(n, order="F").copy()
    F5 = F.reshape((nx, ny, nk, nl, nm), order="F")
    e = E6[:, :, 0, :, :, :]
    if kind == 1:
        E6[:, :, 0, :, :, :] = e + dt * F5 / dz1
    else:
        zf = zdot_faces.reshape((nx, ny, nz + 1, nk, nl, nm), order="F")[:, :, 1, :, :, :]
        w = np.maximum(0.0, zf)
        E6[:, :, 0, :, :, :] = e - dt * w * (e - F5) / dz1
    return E6.ravel(order="F")


@pytest.mark.parametrize("kind", [1, 2], ids=["flux_injection", "relaxation"])
@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
def test_forcing_then_step_matches_oracle(kind, exact, gpu_available):
    c = cases.random_case((5, 4, 6, 4, 3, 4), seed=41)
    rng = np.random.default_rng(5)
    n5 = c.cells // c.n[2]
    F = (2e-7 * rng.random(n5)) if kind == 1 else (1e-3 * (0.5 + rng.random(n5)))
    with rt6d.Model.from_case(c, exact=exact) as m:
        m.calc_parameter()
        dt = m.cfl_dt(0.5)
        zdot = m.download("zdot")
        dz1 = c.grids()[2][1][2]           # dzt(1): halo slot 2
        m.set_forcing(kind, F)
        m.integrate(3, dt=dt)
        got = m.download("E")
        # forcing off again: a further step is a plain step
        m.set_forcing(0)
        m.integrate(1, dt=dt)
        got4 = m.download("E")
    c.dt = dt
    o = O.Oracle(c)
    o.calc_parameter()
    assert np.array_equal(o.get("zdot"), zdot)
    for _ in range(3):
        o.set("E", _numpy_forcing(o.get("E"), kind, F, zdot, c.n, dt, dz1))
        o.integrate(1)
    ref = o.get("E")
    o.integrate(1)
    ref4 = o.get("E")
    o.close()
    if exact:
        assert np.array_equal(got, ref) or rel_l2(got, ref) <= 1e-15
    assert rel_l2(got, ref) <= 1e-12
    assert rel_l2(got4, ref4) <= 1e-12


def test_forcing_closed_forms(gpu_available):
    # relaxation towards the current bottom state changes nothing; injection of
    # zero flux changes nothing: both give the plain step bit for bit
    c = cases.packet_case(n=(6, 4, 6, 4, 4, 6))
    E0 = c.E0.reshape(c.n, order="F")
    with rt6d.Model.from_case(c) as plain, rt6d.Model.from_case(c) as relax, rt6d.Model.from_case(c) as zero:
        for m in (plain, relax, zero):
            m.calc_parameter()
        dt = plain.cfl_dt(0.5)
        relax.set_forcing(2, np.ascontiguousarray(E0[:, :, 0]).ravel(order="F"))
        zero.set_forcing(1, np.zeros(c.cells // c.n[2]))
        for m in (plain, relax, zero):
            m.integrate(1, dt=dt)
        e = plain.download("E")
        assert np.array_equal(relax.download("E"), e)
        assert np.array_equal(zero.download("E"), e)
        with pytest.raises(rt6d.RT6DError, match="kind"):
            plain.set_forcing(3, np.zeros(c.cells // c.n[2]))
        with rt6d.Ensemble(rt6d.FP64) as ens:
            with pytest.raises(rt6d.RT6DError, match="forcing"):
                ens.add(zero, dt=dt)
