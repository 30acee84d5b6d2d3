"""HIP path (librt6d.so through the C ABI) against the reference.

* golden vectors (written by the compiled reference Fortran): every field,
  relative L2 <= 1e-12 in binary64 (bitwise equality is reported for the
  EXACT arithmetic mode and asserted for calc_parameter's outputs);
* the C oracle at medium size over many steps;
* size-independent properties at the 24^3 x 16^3 production size.
"""
import copy
import math

import numpy as np
import pytest

from conftest import golden_case, golden_data, golden_index, rel_l2

import oracle as O
from raytracer_3d_amd import cases, rt6d

pytestmark = pytest.mark.gpu

TOL64 = 1e-12      # relative L2 per field, binary64 step
TOL32 = 1e-5       # relative L2 per field, FP32 form against the binary64 reference
ENTRIES = golden_index()
IDS = [e["stem"] for e in ENTRIES]
HOMOG_EXPECTED = {"random_homog_bg", "packet_small", "gm_small"}


def _model(case, **kw):
    m = rt6d.Model.from_case(case, **kw)
    m.calc_parameter()
    return m


@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_calc_parameter_bitwise(entry, gpu_available):
    case = golden_case(entry)
    g = golden_data(entry)
    with _model(case) as m:
        for a in cases.AXES:
            got = m.download(f"{a}dot")
            assert np.array_equal(got, g[f"{a}dot"]), f"{a}dot max diff {np.max(np.abs(got - g[a + 'dot']))}"
        for name in ("lambda_x", "lambda_y", "lambda_z", "omega_i", "omega_e"):
            assert np.array_equal(m.download(name), g[name]), name
        assert m.homogeneous == (entry["stem"] in HOMOG_EXPECTED)


@pytest.mark.parametrize("kernel", [0, 1, 4, 5, 6, 9, 10, 11, 15, 16, 17, 21, 22, 23],
                         ids=["auto", "percell", "col1", "col2", "col4", "coly1", "coly2", "coly4", "colx1", "colx2",
                              "colx4", "colw1", "colw2", "colw4"])
@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
@pytest.mark.parametrize("path", [0, 1], ids=["auto", "general"])
@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_integrate_matches_reference(entry, path, exact, kernel, gpu_available):
    case = golden_case(entry)
    g = golden_data(entry)
    if kernel in (5, 10, 16, 22) and case.n[2] % 2:
        pytest.skip("two z cells per thread need an even nz")
    if kernel in (6, 11, 17, 23) and case.n[2] % 4:
        pytest.skip("four z cells per thread need nz divisible by 4")
    with _model(case, exact=exact, path=path, kernel=kernel) as m:
        if path == 1:
            assert not m.homogeneous
        # bitwise equality is expected in EXACT mode where lambda_x = lambda_y = 0
        # (meanfl = E*lambda_z either way); otherwise the pre-summed lambda
        # rounds differently from E*lambda_x + E*lambda_y + E*lambda_z
        bitwise = exact and entry["stem"] in HOMOG_EXPECTED
        m.integrate(1)
        worst = 0.0
        for name in ["E", "dE", "meanfl"] + [f"flux_{a}" for a in cases.AXES]:
            got = m.download(name)
            err = rel_l2(got, g[name + "1"])
            worst = max(worst, err)
            assert err <= TOL64, (name, err)
            if bitwise:
                assert np.array_equal(got, g[name + "1"]), f"exact mode {name} not bitwise: {err:.3e}"
        m.integrate(int(g["nsteps"]) - 1)
        for name in ("E", "dE", "meanfl"):
            got = m.download(name)
            err = rel_l2(got, g[name + "N"])
            assert err <= TOL64, (name, err)
            if bitwise:
                assert np.array_equal(got, g[name + "N"]), f"exact mode {name} not bitwise: {err:.3e}"


@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_superbee_fluxes_of_current_state(entry, gpu_available):
    """rt6d_fluxes = superbee_fluxes on the current E (with integrate's ghosts)."""
    case = golden_case(entry)
    o = O.Oracle(case)
    o.calc_parameter()
    o.integrate(2)
    o.superbee_fluxes()
    with _model(case) as m:
        m.integrate(2)
        m.superbee_fluxes()
        for a in cases.AXES:
            assert rel_l2(m.download(f"flux_{a}"), o.get(f"flux_{a}")) <= TOL64, a


@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_fp32_form_against_fp64_reference(entry, gpu_available):
    case = golden_case(entry)
    g = golden_data(entry)
    with _model(case, precision=rt6d.FP32) as m:
        m.integrate(int(g["nsteps"]))
        err = rel_l2(m.download("E"), g["EN"])
        assert err <= TOL32, err


@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_diagnose_matches_oracle(entry, exact, gpu_available):
    case = golden_case(entry)
    o = O.Oracle(case)
    o.calc_parameter()
    o.integrate(3)
    od, o3 = o.diagnose(fields3d=True)
    with _model(case, exact=exact) as m:
        m.integrate(3)
        gd, g3 = m.diagnose(fields3d=True)
    terms = [od["eint"], od["mflint"]] + [od[k] for k in od if k.startswith("bfl_")]
    scale = sum(abs(t) for t in terms) or 1.0
    for k in ["eint", "mflint", "residual"] + [k for k in od if k.startswith("bfl_")]:
        assert abs(gd[k] - od[k]) <= 1e-12 * scale, (k, gd[k], od[k])
    for k in ["mfl_max", "umax", "vmax", "wmax", "kmax", "lmax", "mmax"] + [f"cfl_{a}" for a in "xyzklm"]:
        assert gd[k] == pytest.approx(od[k], rel=1e-14, abs=1e-300), k
    assert abs(gd["residual"]) <= 1e-11 * scale
    for k in ("E_int", "meanfl_int"):
        assert rel_l2(g3[k], o3[k]) <= 1e-13, k
    for q in range(6):
        a, b = g3["diss"][q], o3["diss"][q]
        assert np.max(np.abs(a - b)) <= 1e-13 * (np.max(np.abs(b)) + 1e-300), q


def test_step_before_setup_is_an_error(gpu_available):
    with rt6d.Model((4, 4, 4, 4, 4, 4)) as m:
        with pytest.raises(rt6d.RT6DError, match="calc_parameter"):
            m.integrate(1, dt=1.0)
        with pytest.raises(rt6d.RT6DError, match="grid"):
            m.calc_parameter()


def _medium_packet():
    c = cases.packet_case(n=(16, 4, 16, 8, 4, 8), seed=77, U0=0.15, Nscale=1300.0)
    return c


@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
def test_medium_packet_many_steps_against_oracle(exact, gpu_available):
    c = _medium_packet()
    with _model(c, exact=exact) as m:
        c.dt = m.cfl_dt(0.5)
        m.dt = c.dt
        assert m.homogeneous
        m.integrate(40)
        got = m.download("E")
    o = O.Oracle(c)
    o.calc_parameter()
    o.integrate(40)
    ref = o.get("E")
    err = rel_l2(got, ref)
    assert err <= TOL64, err
    if exact:
        assert np.array_equal(got, ref), err


def test_fp32_default_kernel_choice_is_invisible(gpu_available):
    """The binary32 homogeneous default is the wave-local pair kernel where nx
    is even and at most 32 (the plain pair kernel on grids too small to fill
    the device with one z cell per thread, the quad kernel for wider rows);
    all of them give the column kernel's values bit for bit, so ensembles
    and single contexts agree."""
    case = cases.packet_case(n=(16, 8, 8, 4, 4, 4), name="fp32_pairs")   # 64 pairs per plane
    outs = {}
    for kernel in (0, 5, 16, 17, 19, 21, 22, 23):
        with _model(case, precision=rt6d.FP32, kernel=kernel) as m:
            m.dt = m.cfl_dt(0.5)
            m.integrate(9)
            outs[kernel] = m.download("E")
    assert all(np.array_equal(outs[0], outs[k]) for k in (5, 16, 17, 19, 21, 22, 23))
    with _model(case) as m:
        m.integrate(9, dt=m.cfl_dt(0.5))
        assert rel_l2(outs[0], m.download("E")) <= TOL32


@pytest.mark.parametrize("n", [(16, 8, 8, 4, 4, 4), (12, 5, 6, 4, 3, 5), (8, 1, 4, 4, 4, 4), (4, 6, 4, 3, 4, 4),
                               (24, 24, 4, 4, 4, 4)], ids=lambda n: "x".join(map(str, n)))
def test_fp32_quad_kernel_is_bitwise_the_column_kernel(n, gpu_available):
    """The quad kernel (four x cells per thread, options 18/19/20) gives the
    column kernel's binary32 values bit for bit: one quad per row (nx = 4),
    ragged planes (padding lanes), inactive y, and whole-wave planes."""
    case = cases.packet_case(n=n, name="fp32_quads")
    outs = {}
    for kernel in (5, 18, 19, 20):
        with _model(case, precision=rt6d.FP32, kernel=kernel) as m:
            assert m.homogeneous
            m.dt = m.cfl_dt(0.5)
            m.integrate(9)
            outs[kernel] = m.download("E")
    for kernel in (18, 19, 20):
        assert np.array_equal(outs[kernel], outs[5]), kernel
    with _model(case) as m:
        m.integrate(9, dt=m.cfl_dt(0.5))
        assert rel_l2(outs[19], m.download("E")) <= TOL32


@pytest.mark.parametrize("precision", [rt6d.FP64, rt6d.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [(24, 24, 4, 4, 4, 4), (32, 8, 4, 4, 3, 4), (12, 5, 6, 4, 3, 5), (2, 7, 4, 3, 4, 4),
                               (30, 3, 4, 4, 4, 4), (16, 1, 4, 4, 4, 4), (4, 9, 4, 3, 3, 4), (64, 3, 4, 3, 3, 4),
                               (22, 22, 4, 4, 3, 4), (20, 13, 4, 3, 3, 4)],
                         ids=lambda n: "x".join(map(str, n)))
def test_wave_local_pair_kernel_is_bitwise_the_column_kernel(n, precision, gpu_available):
    """The pair kernel with x and y neighbours exchanged through wave-local
    LDS slots (options 21/22/23) gives the column kernel's values bit for
    bit: halos cut by the plane's ends, ragged last waves (padding lanes that
    load halo pairs), a single pair per row, inactive y, nx up to 32 (and
    wider rows, which fall back to the column kernel).  Where wavefronts
    start at row starts (64 a multiple of nx/2, or in binary64 whole rows
    dealt to wavefronts: 24, 12, 30, 20 wide) the x face between two pairs
    passes from lane to lane; at 22 x 22 whole rows would take one wavefront
    more per plane, so it is computed by both lanes."""
    case = cases.packet_case(n=n, name="wave_local")
    outs = {}
    for kernel in (5, 21, 22, 23):
        with _model(case, precision=precision, kernel=kernel) as m:
            assert m.homogeneous
            m.dt = m.cfl_dt(0.5)
            m.integrate(9)
            outs[kernel] = m.download("E")
    for kernel in (21, 22, 23):
        assert np.array_equal(outs[kernel], outs[5]), kernel


@pytest.mark.parametrize("kernel", [18, 19, 20], ids=["colq1", "colq2", "colq4"])
@pytest.mark.parametrize("entry", ENTRIES, ids=IDS)
def test_fp32_quad_kernel_against_reference(entry, kernel, gpu_available):
    case = golden_case(entry)
    g = golden_data(entry)
    outs = []
    for k in (kernel, 0):
        with _model(case, precision=rt6d.FP32, kernel=k) as m:
            m.integrate(int(g["nsteps"]))
            outs.append(m.download("E"))
    assert np.array_equal(outs[0], outs[1])
    assert rel_l2(outs[0], g["EN"]) <= TOL32


@pytest.mark.parametrize("n, precision, path, expected", [
    ((24, 24, 24, 16, 16, 16), rt6d.FP64, 0, 22),     # wave-local pairs, two z cells
    ((24, 24, 24, 16, 16, 16), rt6d.FP32, 0, 22),     # wave-local pairs
    ((32, 32, 32, 16, 16, 16), rt6d.FP32, 0, 22),     # wave-local pairs up to nx = 32
    ((12, 12, 12, 8, 8, 8), rt6d.FP32, 0, 16),        # too small to fill: plain pairs
    ((12, 12, 12, 8, 8, 8), rt6d.FP64, 0, 21),        # too small for two z cells: one
    ((24, 24, 23, 16, 16, 16), rt6d.FP32, 0, 21),     # odd nz: wave-local pairs, one z cell
    ((36, 36, 36, 8, 8, 8), rt6d.FP32, 0, 19),        # nx > 32: quads
    ((36, 36, 36, 8, 8, 8), rt6d.FP64, 0, 16),        # nx > 32: plain pairs
    ((23, 24, 24, 8, 8, 8), rt6d.FP64, 0, 5),         # odd nx: column kernel
    ((24, 24, 24, 8, 8, 8), rt6d.FP64, 1, 10),        # general path: LDS-staged y
], ids=["fp64-24", "fp32-24", "fp32-32", "fp32-12", "fp64-12", "fp32-oddz", "fp32-36", "fp64-36", "fp64-oddx",
        "general"])
def test_auto_kernel_choice(n, precision, path, expected, gpu_available):
    """RT6D_Q_KERNEL reports what the auto choice resolves to for the grid
    (the choice is invisible in the results, so it is checked here)."""
    case = cases.member_spec(3, 0, n=n)
    with rt6d.Model.from_case(case, precision=precision, path=path) as m:
        m.calc_parameter()
        assert m.homogeneous == (path == 0)
        assert m.kernel == expected
    with rt6d.Model.from_case(case, precision=precision, kernel=4) as m:
        m.calc_parameter()
        assert m.kernel == 4


@pytest.mark.parametrize("n, precision, exact, kernel, expected", [
    ((24, 24, 24, 16, 16, 16), rt6d.FP64, False, 0, 1),   # whole rows dealt: 5 rows of 12 pairs per wavefront
    ((24, 24, 24, 16, 16, 16), rt6d.FP32, False, 0, 0),   # binary32 keeps 64 pairs per wavefront
    ((32, 32, 32, 16, 16, 16), rt6d.FP32, False, 0, 1),   # 64 pairs are 4 rows anyway
    ((24, 24, 24, 16, 16, 16), rt6d.FP64, True, 0, 0),    # exact mode computes every face
    ((24, 24, 24, 16, 16, 16), rt6d.FP64, False, 16, 0),  # plain pairs
    ((22, 22, 24, 8, 8, 8), rt6d.FP64, False, 22, 0),     # whole rows would add a wavefront per plane
], ids=["fp64-24", "fp32-24", "fp32-32", "exact", "pairs", "fp64-22"])
def test_x_face_handover_query(n, precision, exact, kernel, expected, gpu_available):
    """RT6D_Q_X_HANDOVER: the wave-local pair kernel passes the x face between
    two pairs from lane to lane only where its wavefronts start at row starts
    (the results are those of the column kernel either way)."""
    case = cases.member_spec(3, 2, n=n)
    with rt6d.Model.from_case(case, precision=precision, exact=exact, kernel=kernel) as m:
        m.calc_parameter()
        assert m.query(rt6d.Q_X_HANDOVER) == expected


@pytest.mark.parametrize("n, path, option, expected", [
    ((24, 24, 23, 8, 8, 8), 0, 22, 4),     # two z cells per thread need nz even: column kernel, one
    ((24, 24, 24, 8, 8, 8), 1, 16, 5),     # pairs need the homogeneous path: column kernel, two
    ((24, 24, 24, 8, 8, 8), 1, 22, 5),
    ((36, 36, 36, 8, 8, 8), 0, 22, 5),     # the wave-local exchange needs nx <= 32
    ((36, 36, 36, 8, 8, 8), 0, 16, 16),    # plain pairs fit
    ((24, 24, 24, 8, 8, 8), 0, 19, 5),     # quads are binary32 only
], ids=["oddz-22", "general-16", "general-22", "wide-22", "wide-16", "fp64-19"])
def test_kernel_query_names_the_fallback(n, path, option, expected, gpu_available):
    """A forced option that does not fit the grid runs the dispatch's
    fallback, and RT6D_Q_KERNEL names that kernel; the results are the
    fallback's bits."""
    case = cases.member_spec(3, 1, n=n)
    outs = []
    for k in (option, expected):
        with rt6d.Model.from_case(case, path=path, kernel=k) as m:
            m.calc_parameter()
            assert m.kernel == expected, (k, m.kernel)
            m.integrate(3, dt=m.cfl_dt(0.5))
            outs.append(m.download("E"))
    assert np.array_equal(outs[0], outs[1])


# ----------------------------------------------------------- production size

@pytest.fixture(scope="module")
def gm_full():
    c = cases.gm_case()   # 24^3 x 16^3
    return c


def test_full_size_determinism_paths_and_budget(gm_full, gpu_available):
    c = gm_full
    with _model(c) as m:
        dt = m.cfl_dt(0.5)
        assert m.homogeneous
        m.integrate(10, dt=dt)
        e1 = m.download("E")
        d = m.diagnose(dt=dt)
    with _model(c) as m:
        m.integrate(10, dt=dt)
        assert np.array_equal(m.download("E"), e1), "not deterministic"
    with _model(c, path=1) as m:
        assert not m.homogeneous
        m.integrate(10, dt=dt)
        assert rel_l2(m.download("E"), e1) <= 1e-14
    with _model(c, exact=True) as m:
        m.integrate(10, dt=dt)
        ex = m.download("E")
        assert rel_l2(ex, e1) <= TOL64
    with _model(c, exact=True, kernel=1) as m:
        m.integrate(10, dt=dt)
        assert np.array_equal(m.download("E"), ex), "auto and per-cell kernels differ in exact mode"
    for kernel in (1, 4, 5, 6, 9, 10, 11, 15, 16, 17, 21, 22, 23):
        with _model(c, kernel=kernel) as m:
            m.integrate(10, dt=dt)
            assert rel_l2(m.download("E"), e1) <= 1e-14, kernel
        with _model(c, exact=True, kernel=kernel) as m:
            m.integrate(10, dt=dt)
            assert np.array_equal(m.download("E"), ex), f"kernel {kernel} differs in exact mode"
    terms = [d["eint"], d["mflint"]] + [d[k] for k in d if k.startswith("bfl_")]
    assert abs(d["residual"]) <= 1e-11 * sum(abs(t) for t in terms)
    assert np.all(np.isfinite(e1))


def test_full_size_fp32_form(gm_full, gpu_available):
    c = gm_full
    with _model(c) as m:
        dt = m.cfl_dt(0.5)
        m.integrate(20, dt=dt)
        e64 = m.download("E")
    with _model(c, precision=rt6d.FP32) as m:
        assert m.homogeneous
        m.integrate(20, dt=dt)
        e32 = m.download("E")
    assert rel_l2(e32, e64) <= TOL32
    # at this size the default binary32 kernel loads the far outer z planes
    # only where the face signs read them; the column kernel loads all four
    # (with the z reflection and with zero-gradient z ends)
    with _model(c, precision=rt6d.FP32, kernel=5) as m:
        m.integrate(20, dt=dt)
        assert np.array_equal(m.download("E"), e32)
    c_open = copy.copy(c)
    c_open.reflect_up = c_open.reflect_down = False
    outs = []
    for kernel in (0, 5):
        with _model(c_open, precision=rt6d.FP32, kernel=kernel) as m:
            m.integrate(20, dt=dt)
            outs.append(m.download("E"))
    assert np.array_equal(outs[0], outs[1])


# ----------------------------------------------------------- config-0 size

def test_config0_packet_full_size(gpu_available):
    """The 32x4x32x16x4x16 wave packet (4,194,304 cells), 200 steps at CFL 0.5:
    the exact mode equals the oracle bit for bit after 5 steps; after 200
    steps the fast binary64 step is within 1e-12 of the exact one, the FP32
    form within 1e-5, and the energy budget closes."""
    c = cases.packet_case()
    with _model(c, exact=True) as m:
        dt = m.cfl_dt(0.5)
        assert m.homogeneous
        m.integrate(5, dt=dt)
        e5 = m.download("E")
        m.integrate(195, dt=dt)
        ex = m.download("E")
    c.dt = dt
    o = O.Oracle(c)
    o.calc_parameter()
    o.integrate(5)
    assert np.array_equal(e5, o.get("E"))
    o.close()
    with _model(c) as m:
        m.integrate(200, dt=dt)
        fast = m.download("E")
        d = m.diagnose(dt=dt)
    assert rel_l2(fast, ex) <= TOL64
    terms = [d["eint"], d["mflint"]] + [d[k] for k in d if k.startswith("bfl_")]
    assert abs(d["residual"]) <= 1e-11 * sum(abs(t) for t in terms)
    with _model(c, precision=rt6d.FP32) as m:
        m.integrate(200, dt=dt)
        assert rel_l2(m.download("E"), ex) <= TOL32


def test_pair_kernel_beyond_2gib_fields(gpu_available):
    """48^3 x 16^3 (452,984,832 cells, 3.6 GB per binary64 field): the pair
    kernel addresses E one m plane at a time (226 MB), so the binary64 step
    keeps it; it gives the per-cell kernel's bits in both arithmetic modes,
    the budget closes, and FP32 (quad kernel, 1.8 GB fields) agrees to FP32
    accuracy."""
    spec = cases.member_spec(11, 2, n=(48, 48, 48, 16, 16, 16))
    out = {}
    dt = None
    for exact in (False, True):
        for kernel in (0, 1):
            with rt6d.Model.from_case(spec, kernel=kernel, exact=exact) as m64:
                m64.calc_parameter()
                assert m64.homogeneous and m64.query(rt6d.Q_CELLS) == 452984832
                assert m64.kernel == (16 if kernel == 0 else 1), m64.kernel
                dt = m64.cfl_dt(0.5) if dt is None else dt
                m64.integrate(4, dt=dt)
                if kernel == 0 and not exact:
                    d = m64.diagnose(dt=dt)
                out[exact, kernel] = m64.download("E")
        assert np.array_equal(out[exact, 0], out[exact, 1]), f"exact={exact}"
        del out[exact, 1]
    e64 = out[False, 0]
    assert rel_l2(out[True, 0], e64) <= TOL64
    terms = [d["eint"], d["mflint"]] + [d[k] for k in d if k.startswith("bfl_")]
    assert abs(d["residual"]) <= 1e-11 * sum(abs(t) for t in terms)
    assert np.all(np.isfinite(e64))
    with rt6d.Model.from_case(spec, precision=rt6d.FP32) as m32:
        m32.calc_parameter()
        assert m32.kernel == 19
        m32.integrate(4, dt=dt)
        e32 = m32.download("E")
    assert rel_l2(e32, e64) <= TOL32


@pytest.mark.parametrize("kernel", [2, 3, 7, 8, 12, 13, 14, 24, 25, 26, 27, -1])
def test_removed_kernel_options_are_rejected(kernel, gpu_available):
    """Options of kernels that were removed (and out-of-range options) are
    errors, not a silent fallback to another kernel."""
    with rt6d.Model((8, 4, 4, 4, 4, 4)) as m:
        with pytest.raises(rt6d.RT6DError, match="kernel option"):
            m.set_option(rt6d.OPT_KERNEL, kernel)
