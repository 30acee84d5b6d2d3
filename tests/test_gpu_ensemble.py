"""Ensembles (rt6d_ensemble_*) against members run one at a time.

An ensemble runs the same column-kernel body as a single context's fast
path, so every member's E after any sequence of ensemble steps -- graph
replays of two steps, single launches for odd counts, members joining and
leaving, schedule changes, the autotuner -- must equal, bit for bit, the E
of the same case stepped the same number of times on its own.
"""
import numpy as np
import pytest

from raytracer_3d_amd import cases, rt6d

pytestmark = pytest.mark.gpu


def _cases():
    return [
        cases.packet_case(n=(8, 4, 6, 4, 4, 6), name="packet_even_nz"),           # homogeneous, ZC=2
        cases.packet_case(n=(6, 3, 5, 4, 3, 5), U0=0.2, Nscale=1300.0, name="packet_odd_nz"),  # ZC=1
        cases.random_case((6, 5, 4, 4, 3, 4), seed=21),                            # general (lambda_x,y, kdot)
        cases.random_case((5, 4, 6, 3, 2, 4), seed=22, horizontal=False, reflect_up=False),
    ]


def _model(case, precision, path=0):
    m = rt6d.Model.from_case(case, precision=precision, path=path)
    m.calc_parameter()
    m.dt = m.cfl_dt(0.5)
    return m


def _alone(case, precision, nsteps, path=0):
    with _model(case, precision, path) as m:
        m.integrate(nsteps)
        return m.download("E")


@pytest.mark.parametrize("precision", [rt6d.FP64, rt6d.FP32], ids=["fp64", "fp32"])
def test_ensemble_equals_individual_runs(precision, gpu_available):
    cs = _cases()
    ms = [_model(c, precision, path=1 if i == 3 else 0) for i, c in enumerate(cs)]
    assert ms[0].homogeneous and not ms[2].homogeneous and not ms[3].homogeneous
    with rt6d.Ensemble(precision) as ens:
        slots = [ens.add(m) for m in ms]
        ens.step(3)        # one graph (2 steps) + one single launch
        ens.step(4)        # two graph replays from odd parity
        ens.step(1)
        info = ens.info()
        assert info["members"] == 4 and info["steps"] == 8
        for c, m in zip(cs, ms):
            E = m.download("E")
            ref = _alone(c, precision, 8, path=1 if c is cs[3] else 0)
            assert np.array_equal(E, ref), f"{c.name}: max diff {np.max(np.abs(E - ref))}"
        # membership changes: member 1 leaves after 8 steps, a new one joins
        gone = ens.remove(slots[1])
        new_case = cases.random_case((4, 4, 5, 4, 2, 4), seed=23)
        new = _model(new_case, precision)
        ens.add(new)
        ens.set_schedule(2, 1)
        ens.step(5)
        ens.set_schedule(8, 0)
        ens.step(2)
        ens.set_schedule(0, 1)      # one block per task
        ens.step_async(3)
        ens.step_async(2)           # enqueued behind the first
        assert ens.wait() > 0.0
        for c, m, n in ((cs[0], ms[0], 20), (cs[2], ms[2], 20), (cs[3], ms[3], 20), (new_case, new, 12)):
            E = m.download("E")
            ref = _alone(c, precision, n, path=1 if c is cs[3] else 0)
            assert np.array_equal(E, ref), f"{c.name}: max diff {np.max(np.abs(E - ref))}"
        # the removed member is a normal context again and resumes where it stopped
        gone.integrate(2)
        assert np.array_equal(gone.download("E"), _alone(cs[1], precision, 10))
    for m in ms + [new]:
        m.close()


@pytest.mark.parametrize("precision", [rt6d.FP64, rt6d.FP32], ids=["fp64", "fp32"])
def test_ensemble_member_kernel_variants(precision, gpu_available):
    """Members that take each ensemble kernel variant of the homogeneous path
    under each tile choice -- the wave-local pair kernel (nx even and <= 32)
    with two or one z cells per thread, with uniform x/y spacings as scalars
    or per-lane spacings (non-uniform grid), and without the exchange the
    plain pairs, quads in binary32 (nx divisible by 4) and the column kernel
    (odd nx) -- step together and each equals its own run bit for bit,
    through the graph and the single launches, in both block modes."""
    cs = [cases.packet_case(n=(8, 8, 4, 4, 4, 4), name="nx8"),
          cases.packet_case(n=(28, 4, 6, 4, 3, 4), name="nx28"),
          cases.packet_case(n=(26, 3, 4, 3, 3, 4), name="nx26"),
          cases.packet_case(n=(32, 5, 5, 3, 4, 4), name="nx32_odd_nz"),
          cases.packet_case(n=(36, 4, 4, 3, 3, 4), name="nx36"),
          cases.packet_case(n=(7, 4, 4, 3, 3, 4), name="nx7"),
          cases.random_case((8, 6, 6, 4, 4, 4), seed=24, horizontal=False, name="nonuniform_xy")]
    ms = [_model(c, precision) for c in cs]
    assert all(m.homogeneous for m in ms)
    assert ms[-1].query(rt6d.Q_UNIFORM_XY) == 0 and ms[0].query(rt6d.Q_UNIFORM_XY) == 3
    steps = 0
    with rt6d.Ensemble(precision) as ens:
        for m in ms:
            ens.add(m)
        for tiles in (rt6d.TILES_WAVE_LOCAL, rt6d.TILES_WAVE_LOCAL_1Z, rt6d.TILES_PAIRS):
            ens.set_tiles(tiles)
            ens.set_schedule(0, 0)
            ens.step(3)
            ens.set_schedule(4, 1)
            ens.step(4)
            steps += 7
            for c, m in zip(cs, ms):
                E = m.download("E")
                ref = _alone(c, precision, steps)
                assert np.array_equal(E, ref), f"{c.name}, tiles {tiles}: max diff {np.max(np.abs(E - ref))}"
    for m in ms:
        m.close()


def test_ensemble_tune_and_state(gpu_available):
    cs = _cases()[:3]
    ms = [_model(c, rt6d.FP64) for c in cs]
    with rt6d.Ensemble(rt6d.FP64) as ens:
        for m in ms:
            ens.add(m)
        bpc, order, tiles = ens.tune(trial_steps=2)
        assert bpc in (0, 2, 4, 8) and order in (0, 1)
        assert tiles in (rt6d.TILES_WAVE_LOCAL, rt6d.TILES_WAVE_LOCAL_1Z, rt6d.TILES_PAIRS)
        steps = ens.info()["steps"]
        assert steps == 24 * (2 + 2)
        ens.step(1)
        for c, m in zip(cs, ms):
            assert np.array_equal(m.download("E"), _alone(c, rt6d.FP64, steps + 1))
        # a member is stepped only through its ensemble
        with pytest.raises(rt6d.RT6DError, match="ensemble member"):
            ms[0].integrate(1)
        with pytest.raises(rt6d.RT6DError, match="ensemble member"):
            ms[0].upload_E(cs[0].E0)
        # diagnose works between ensemble steps and matches an individual run
        d = ms[1].diagnose()
        with _model(cs[1], rt6d.FP64) as solo:
            solo.integrate(steps + 1)
            ds = solo.diagnose()
        assert d["eint"] == ds["eint"] and d["residual"] == ds["residual"]
    # destroying the ensemble hands members back
    ms[0].integrate(1)
    for m in ms:
        m.close()


def test_ensemble_rejects_bad_members(gpu_available):
    c = _cases()[0]
    m64 = _model(c, rt6d.FP64)
    m32 = _model(c, rt6d.FP32)
    mex = _model(c, rt6d.FP64)
    mex.set_option(rt6d.OPT_EXACT, 1)
    bare = rt6d.Model(c.n)
    with rt6d.Ensemble(rt6d.FP64) as ens:
        with pytest.raises(rt6d.RT6DError, match="precision"):
            ens.add(m32)
        with pytest.raises(rt6d.RT6DError, match="exact"):
            ens.add(mex)
        with pytest.raises(rt6d.RT6DError, match="calc_parameter"):
            ens.add(bare, dt=1.0)
        with pytest.raises(rt6d.RT6DError, match="dt"):
            ens.add(m64, dt=-1.0)
        s = ens.add(m64)
        with pytest.raises(rt6d.RT6DError, match="already"):
            ens.add(m64)
        # destroying a member context takes it out of the ensemble
        m64.close()
        assert ens.info()["members"] == 0
        with pytest.raises(rt6d.RT6DError, match="slot"):
            ens.remove(s)
        ens.step(2)   # empty ensemble: nothing to do
    for m in (m32, mex, bare):
        m.close()


@pytest.mark.parametrize("persistent", [False, True], ids=["dispatch", "persistent"])
def test_row_aligned_members_pass_x_faces_between_lanes(persistent, gpu_available):
    """Binary64 wave-local members that can all deal whole rows to their
    wavefronts take the x-face hand-over together (RT6D_EQ_X_HANDOVER); one
    member that cannot (24 x 16: whole rows would take a fourth wavefront
    per plane where 192 pairs fill three) turns it off for the ensemble, so
    that a step stays one launch.  Every member equals its own run, bit for
    bit, either way."""
    same = [cases.member_spec(77, i, n=(24, 24, 8, 8, 8, 8)) for i in range(3)]
    mixed = same + [cases.member_spec(77, 9, n=(24, 16, 8, 8, 8, 8))]
    for specs, handover in ((same, 3), (mixed, 0)):
        ms = [_model(s, rt6d.FP64) for s in specs]
        try:
            with rt6d.Ensemble(rt6d.FP64) as ens:
                for m in ms:
                    ens.add(m)
                ens.set_tiles(rt6d.TILES_WAVE_LOCAL)
                if persistent:
                    ens.set_schedule(4, 1)
                assert ens.query(rt6d.EQ_X_HANDOVER) == handover
                assert ens.query(rt6d.EQ_LAUNCHES) == 1
                ens.step(5)
                got = [m.download("E") for m in ms]
            for s, E in zip(specs, got):
                assert np.array_equal(E, _alone(s, rt6d.FP64, 5)), s.n
        finally:
            for m in ms:
                m.close()
