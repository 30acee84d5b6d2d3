"""Serving loop (raytracer_3d_amd.serve): continuous batching is invisible in
the results.

Requests with uneven grids, both precisions and different step counts
arrive while others are on the device, wait for HBM room, join and leave
their ensemble at different steps; each result must equal, bit for bit, the
same request run alone with plain single-integration steps.
"""
import numpy as np
import pytest

from raytracer_3d_amd import cases, rt6d, serve

pytestmark = pytest.mark.gpu

SHAPES = [(8, 6, 6, 4, 4, 6), (6, 8, 4, 4, 4, 4), (10, 4, 6, 6, 2, 4), (5, 5, 5, 4, 3, 5), (12, 6, 8, 4, 4, 4)]


def _requests(count, steps=(5, 40), spacing=0.02, seed=99):
    out = []
    for i in range(count):
        u = cases.uniform(seed, i, np.arange(3, dtype=np.uint64), stream=3)
        st = int(steps[0] + (steps[1] - steps[0]) * u[0])
        prec = rt6d.FP32 if u[1] < 0.75 else rt6d.FP64
        out.append(serve.Request(i, spacing * i, i, SHAPES[i % len(SHAPES)], st, prec, seed))
    return out


def _check_alone(reqs):
    for r in reqs:
        assert r.result is not None and np.isfinite(r.done) and r.done >= r.arrival
        assert r.result["steps"] == r.steps
        _, E = serve.run_alone(r)
        assert np.array_equal(r.result["E"], E), f"request {r.rid} ({r.n}, prec {r.precision}) differs"


def test_each_result_equals_the_request_alone(gpu_available):
    reqs = _requests(12)
    # room for about three members at a time: requests queue and join as others leave
    budget = 3 * max(serve.peak_bytes(r.n, r.precision) for r in reqs)
    srv = serve.Server(budget, chunk=4, keep_E=True, tune=False)
    try:
        srv.run(reqs)
        assert srv.used == pytest.approx(0.0, abs=1.0)
        assert srv.peak_used <= budget
    finally:
        srv.close()
    _check_alone(reqs)


@pytest.mark.parametrize("to_disk", [False, True], ids=["memory", "restart_files"])
def test_preempted_requests_resume_to_the_same_result(tmp_path, to_disk, gpu_available):
    """With room for two members and preemption at once, long requests are
    checkpointed, evicted and resumed; every result still equals the request
    run alone, and its step count is whole."""
    reqs = _requests(8, steps=(100, 200), spacing=0.0, seed=7)   # all queued at once
    budget = 2 * max(serve.peak_bytes(r.n, r.precision) for r in reqs)
    srv = serve.Server(budget, chunk=4, keep_E=True, tune=False, preempt_after=0.0,
                       checkpoint_dir=str(tmp_path) if to_disk else None)
    try:
        srv.run(reqs)
        assert srv.preemptions >= 1
        assert srv.used == pytest.approx(0.0, abs=1.0)
    finally:
        srv.close()
    assert any(r.preemptions for r in reqs)
    if to_disk:
        assert list(tmp_path.glob("restart_request*.dta"))
    _check_alone(reqs)


def test_tuned_mix_equals_alone(gpu_available):
    reqs = _requests(10, steps=(300, 400), spacing=0.0)
    for r in reqs:
        r.precision = rt6d.FP64   # one ensemble, so that members share it while the tuner runs
    srv = serve.Server(1e12, chunk=8, keep_E=True, tune=True, tune_min_members=2)
    try:
        srv.run(reqs)
        assert any(srv.tuned.values())
    finally:
        srv.close()
    _check_alone(reqs)


def test_resident_estimate_covers_the_device_bytes(gpu_available):
    for n in SHAPES:
        for prec in (rt6d.FP64, rt6d.FP32):
            m = serve.build_member(serve.Request(0, 0.0, 3, n, 1, prec))
            assert m.homogeneous
            got = m.query(rt6d.Q_DEVICE_BYTES)
            est = serve.resident_bytes(n, prec)
            m.close()
            assert got <= est <= 1.5 * got + (1 << 16), (n, prec, got, est)
            # a sheared background keeps the 6D fields
            c = cases.random_case(n, seed=5)
            with rt6d.Model.from_case(c, precision=prec) as m:
                m.calc_parameter()
                assert not m.homogeneous
                got = m.query(rt6d.Q_DEVICE_BYTES)
            est = serve.resident_bytes(n, prec, False)
            assert got <= est <= 1.5 * got + (1 << 16), (n, prec, got, est)


def test_separable_init_matches_host(gpu_available):
    spec = cases.member_spec(7, 3, n=(6, 5, 4, 4, 3, 5))
    i = spec.meta["init"]
    ref = cases.separable_E(i["factors"], i["peak"], i["noise"], i["seed"], i["member"])
    with rt6d.Model.from_case(spec) as m:
        assert np.array_equal(m.download("E"), ref)
    with rt6d.Model.from_case(spec, precision=rt6d.FP32) as m:
        assert np.array_equal(m.download("E"), ref.astype(np.float32).astype(np.float64))


def test_branch_flips_counts_faces_and_rare_flips(gpu_available):
    c = cases.random_case((6, 5, 4, 4, 3, 4), seed=31)
    m32 = rt6d.Model.from_case(c, precision=rt6d.FP32)
    # the binary64 model starts from the binary32 state: flips come from arithmetic alone
    c.E0 = c.E0.astype(np.float32).astype(np.float64)
    m64 = rt6d.Model.from_case(c)
    try:
        m64.calc_parameter()
        m32.calc_parameter()
        faces, flips = rt6d.branch_flips(m64, m32)
        assert faces == [c.cells] * 6          # every axis active on this case
        assert sum(flips) <= 1e-2 * sum(faces)
        dt = m64.cfl_dt(0.5)
        m64.integrate(20, dt=dt)
        m32.integrate(20, dt=dt)
        faces2, flips2 = rt6d.branch_flips(m64, m32)
        assert faces2 == faces and sum(flips2) <= 0.05 * sum(faces)
        with pytest.raises(rt6d.RT6DError, match="binary64 and a binary32"):
            rt6d.branch_flips(m32, m64)
    finally:
        m64.close()
        m32.close()
