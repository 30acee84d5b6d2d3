"""snap3D.cdf / snap6D.cdf schema and contents, read back with scipy.

The files must carry what init_snap3D_cdf (for_src/netcdf_io_3D.f90:1-260),
init_snap6D_cdf_single (for_src/netcdf_io_6D.f90:240-656) and diagnose
(for_src/diagnose.f90:335-530) write. The expected names, dimensions and
attribute strings below are read off those routines. The data are stand-in
fields whose values encode their Fortran indices, so a transposed or
shifted array shows up as a wrong value at a known index. The reader is
scipy's independent implementation of the netCDF classic formats.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from raytracer_3d_amd import cases, snapshot

N = (5, 4, 3, 4, 2, 3)


class FakeModel:
    """download/diagnose with index-coded values (no device)."""

    def __init__(self, n):
        self.n = n
        self.dt = 30.0
        self.calls = 0

    @staticmethod
    def _code(shape, base):
        idx = np.indices(shape)
        v = np.full(shape, float(base))
        for ax in range(len(shape)):
            v += idx[ax] * 10.0 ** ax
        return v.reshape(-1, order="F")

    def download(self, field):
        n = list(self.n)
        if field.endswith("dot"):
            d = cases.AXES.index(field[0])
            n[d] += 1
            return self._code(n, 1e7 * (d + 1))
        base = {"E": 2e7, "meanfl": 3e7, "omega_i": 4e7, "omega_e": 5e7}[field]
        return self._code(n, base + 1e6 * self.calls)

    def diagnose(self, dt=None, fields3d=False):
        self.calls += 1
        n3 = self.n[:3]
        d = {"mflint": 0.5 + self.calls}
        for i, a in enumerate(("xs", "xe", "ys", "ye", "zs", "ze", "ks", "ke", "ls", "le", "ms", "me")):
            d[f"bfl_{a}"] = 100.0 * self.calls + i
        f3 = {"E_int": self._code(n3, 1e3 * self.calls), "meanfl_int": self._code(n3, 1e3 * self.calls + 500),
              "diss": np.stack([self._code(n3, 1e4 * (i + 1) + self.calls) for i in range(6)])}
        return d, f3


def _case():
    c = cases.random_case(N, seed=5)
    c.fixed = (0.0, 0.0, 0.0)
    return c


def _open(path):
    return netcdf_file(path, "r", mmap=False)


def test_snap3d_schema_and_records(tmp_path):
    case = _case()
    m = FakeModel(N)
    s = snapshot.Snapshots(m, case, directory=str(tmp_path), write_6d=False)
    s.diagnose(itt=0, dt=30.0)
    s.diagnose(itt=7, dt=30.0)
    path = tmp_path / "snap3D.cdf"
    assert path.read_bytes()[:4] == b"CDF\x02"          # NF_64BIT_OFFSET
    with _open(str(path)) as f:
        assert f._dims == ["xt", "xu", "yt", "yu", "zt", "zu", "Time"]
        assert f.dimensions["Time"] is None and f.dimensions["xt"] == 5 and f.dimensions["zu"] == 3
        expect_vars = (["Time", "xt", "xu", "yt", "yu", "zt", "zu", "Nsqr", "U", "V", "E", "meanfl"]
                       + list(snapshot.DISS) + [f"bflux_{b}" for b in snapshot.BFLUX] + ["mflint"])
        assert list(f.variables) == expect_vars
        t = f.variables["Time"]
        assert t.long_name == b"Time" + b" " * 20 and t.units == b"seconds" + b" " * 9
        assert t.time_origin == b"01-JAN-1900 00:00:00"
        np.testing.assert_array_equal(t[:], [0.0, 210.0])
        assert f.variables["E"].dimensions == ("Time", "zt", "yt", "xt")
        assert f.variables["E"].units == b"m^2/s^2".ljust(16)
        assert f.variables["diss_ms"].long_name == b"energy change".ljust(24)
        assert f.variables["Nsqr"].long_name == b"buoyancy frequency".ljust(24)
        assert f.variables["bflux_le"].long_name == b"Energy flux across max l"
        assert f.variables["mflint"].units == b" " * 16
        assert f.f0 == case.f0 and f.k_fixed == 0.0 and f.m_fixed == 0.0
        # values: E_int(i,j,k) of record r holds 1e3*r + i + 10 j + 100 k
        E = f.variables["E"][:]
        assert E.shape == (2, 3, 4, 5)
        for r in range(2):
            for (i, j, k) in ((0, 0, 0), (4, 1, 2), (2, 3, 1)):
                assert E[r, k, j, i] == 1e3 * (r + 1) + i + 10 * j + 100 * k
        D = f.variables["diss_le"][:]
        assert D[1, 2, 3, 4] == 1e4 * 4 + 2 + 4 + 30 + 200
        assert f.variables["mflint"][1] == 2.5
        # diagnose's Diag order xs xe ys ye zs ze ks ...: bflux_ks is entry 6
        assert f.variables["bflux_ks"][0] == 106.0 and f.variables["bflux_ye"][1] == 203.0
        # background and grid: the interior of the halo'd arrays
        Nsqr = case.Nsqr.reshape([v + 4 for v in N[:3]], order="F")
        np.testing.assert_array_equal(f.variables["Nsqr"][:], Nsqr[2:-2, 2:-2, 2:-2].T)
        t_x, _, u_x, _ = cases.calc_grid(0, case.spacing[0])
        np.testing.assert_array_equal(f.variables["xt"][:], t_x[2:7])
        np.testing.assert_array_equal(f.variables["xu"][:], u_x[2:7])


def test_snap6d_schema_and_records(tmp_path):
    case = _case()
    m = FakeModel(N)
    s = snapshot.Snapshots(m, case, directory=str(tmp_path), write_6d=True)
    s.diagnose(itt=3, dt=30.0)
    with _open(str(tmp_path / "snap6D.cdf")) as f:
        assert f._dims == ["xt", "xu", "yt", "yu", "zt", "zu", "kt", "ku", "lt", "lu", "mt", "mu", "Time"]
        names = list(f.variables)
        assert names[:4] == ["Time", "xt", "xu", "dxt"] and names[-2:] == ["E", "meanfl"]
        assert names[19:27] == ["xdot", "ydot", "zdot", "kdot", "ldot", "mdot", "omega_i", "omega_e"]
        assert f.variables["dzt"].dimensions == ("zu",)
        assert f.variables["kt"].long_name == b"wavenumber".ljust(24)
        assert f.variables["E"].units == b"m^5/s^2".ljust(16)
        assert f.variables["E"].dimensions == ("Time", "mt", "lt", "kt", "zt", "yt", "xt")
        assert f.variables["xdot"].dimensions == ("mt", "lt", "kt", "zt", "yt", "xu")
        assert f.variables["ldot"].dimensions == ("mt", "lu", "kt", "zt", "yt", "xt")
        assert f.variables["mdot"].units == b"1/ms".ljust(16)
        np.testing.assert_array_equal(f.variables["Time"][:], [90.0])
        E = f.variables["E"][:]
        assert E.shape == (1,) + tuple(reversed(N))
        # the 6D E was downloaded after the diagnose call (calls = 1)
        assert E[0, 2, 1, 3, 2, 1, 4] == 2e7 + 1e6 + 4 + 10 + 200 + 3000 + 1e4 + 2e5
        # xdot(i) for i = 1..nx is the right face of cell i: face index i
        X = f.variables["xdot"][:]
        assert X.shape == (3, 2, 4, 3, 4, 5)
        assert X[0, 0, 0, 0, 0, 0] == 1e7 + 1
        assert X[2, 1, 3, 2, 3, 4] == 1e7 + 5 + 30 + 200 + 3000 + 1e4 + 2e5
        Z = f.variables["zdot"][:]
        assert Z[0, 0, 0, 0, 0, 0] == 3e7 + 100
        np.testing.assert_array_equal(f.variables["omega_i"][:].reshape(-1), m.download("omega_i") - 1e6)
        t_k, d_k, u_k, _ = cases.calc_grid(3, case.spacing[3])
        np.testing.assert_array_equal(f.variables["dkt"][:], d_k[2:6])
        np.testing.assert_array_equal(f.variables["ku"][:], u_k[2:6])


def test_record_slab_limit_and_names(tmp_path):
    f = snapshot.CdfFile(str(tmp_path / "x.cdf"))
    f.dim("Time", None)
    f.dim("a", 2**15)
    f.dim("b", 2**15)
    f.var("A", ["Time", "a", "b"])      # 8 GiB per record, not the last record variable
    f.var("B", ["Time"])
    with pytest.raises(ValueError, match="4 GiB"):
        f.create()
    # init_snap6D_cdf_many's name for a single block, format (a,i5,...) with blanks as '0'
    assert snapshot.per_pe_6d_name((24, 24, 24, 16, 16, 16)) == \
        "snap6D_i=00001:00024_j=00001:00024_k=00001:00024_l=00001:00016_n=00001:00016_m=00001:00016.cdf"
    with pytest.raises(ValueError):
        snapshot.char_att("x" * 25, 24)
