"""Host-side case generation: seeded inputs, case-file round trip, grid checks."""
import numpy as np
import pytest

from raytracer_3d_amd import cases


def test_prng_is_deterministic_and_uniform():
    idx = np.arange(200000, dtype=np.uint64)
    a = cases.uniform(5, 1, idx)
    b = cases.uniform(5, 1, idx)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, cases.uniform(5, 2, idx))
    assert 0.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.005 and abs(a.var() - 1 / 12) < 0.002


def test_case_file_round_trip(tmp_path):
    c = cases.random_case((3, 4, 5, 2, 2, 4), seed=1)
    p = str(tmp_path / "c.case")
    cases.write_case_file(c, p, nsteps=7, dt=12.5)
    r = cases.read_case_file(p)
    assert r.n == c.n and r.dt == 12.5 and r.nsteps == 7
    for a, b in zip(r.spacing, c.spacing):
        assert np.array_equal(a, b)
    for a, b in ((r.Nsqr, c.Nsqr), (r.U, c.U), (r.V, c.V), (r.E0, c.E0)):
        assert np.array_equal(a, b)


def test_grid_check_rejects_growing_dz():
    sp = [np.ones(4)] * 2 + [np.array([10.0, 10.0, 20.0, 20.0])] + [np.ones(4)] * 3
    with pytest.raises(ValueError, match="axis z"):
        cases.check_grid(sp)


def test_gm_spectrum_normalisation():
    c = cases.gm_case(n=(4, 4, 8, 8, 8, 8))
    n = c.n
    E = c.E0.reshape(n, order="F")
    dk, dl, dm = (c.spacing[a][0] for a in (3, 4, 5))
    col = E.sum(axis=(3, 4, 5)) * dk * dl * dm
    assert col.mean() == pytest.approx(3e-3, rel=0.05)
    assert np.all(E > 0)


def test_member_sizes_from_the_stated_sets():
    for m in range(20):
        c = cases.member_case(99, m)
        assert all(v in (12, 16, 20, 24) for v in c.n[:3])
        assert all(v in (8, 12, 16) for v in c.n[3:])
