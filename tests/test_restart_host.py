"""Restart files (raytracer_3d_amd.restart) in the record layout of
for_src/restart.f90:1-93, without a device.

The reader is checked against bytes assembled here record by record from
the layout (7 + 6 + 6 int32, then three halo'd binary64 arrays, each record
framed by its 4-byte length), so a wrong order, a missing halo or a
transposed array fails on known values.
"""
import numpy as np
import pytest

from raytracer_3d_amd import restart

N = (4, 3, 5, 2, 3, 4)


def _coded(n, base):
    idx = np.indices(n)
    v = np.full(n, float(base))
    for ax in range(6):
        v += idx[ax] * 10.0 ** ax
    return v.reshape(-1, order="F")


def test_file_name_matches_the_reference_format():
    # write(fname,'(a,i9,a)') 'restart_pe=',my_pe,'.dta' with blanks made '0'
    assert restart.restart_file_name(0) == "restart_pe=000000000.dta"
    assert restart.restart_file_name(37) == "restart_pe=000000037.dta"


def test_layout_byte_for_byte_and_round_trip(tmp_path):
    E, dE, mf = _coded(N, 1e6), _coded(N, 2e6), _coded(N, 3e6)
    p = tmp_path / restart.restart_file_name(0)
    restart.write_restart(str(p), N, 17, E, dE, mf)
    data = p.read_bytes()
    H = [v + 4 for v in N]
    expect = b""
    for payload in (np.array(list(N) + [17], "<i4").tobytes(),
                     np.array([1, N[0], 1, N[1], 1, N[2]], "<i4").tobytes(),
                     np.array([1, N[3], 1, N[4], 1, N[5]], "<i4").tobytes()):
        expect += np.array([len(payload)], "<i4").tobytes() + payload + np.array([len(payload)], "<i4").tobytes()
    body = b""
    for a in (E, dE, mf):
        h = np.zeros(H, order="F")
        h[2:-2, 2:-2, 2:-2, 2:-2, 2:-2, 2:-2] = a.reshape(N, order="F")
        body += h.tobytes(order="F")
    expect += np.array([len(body)], "<i4").tobytes() + body + np.array([len(body)], "<i4").tobytes()
    assert data == expect
    assert (tmp_path / "ritt").read_text().split() == ["17"]
    r = restart.read_restart(str(p), N)
    assert r["n"] == N and r["itt"] == 17
    assert np.array_equal(r["E"], E) and np.array_equal(r["dE"], dE) and np.array_equal(r["meanfl"], mf)
    # E(i,j,k,l,n,m) at Fortran index (3,2,4,1,2,3) (0-based) holds its coded value
    flat = 3 + N[0] * (2 + N[1] * (4 + N[2] * (1 + N[3] * (2 + N[4] * 3))))
    assert r["E"][flat] == 1e6 + 3 + 20 + 400 + 1000 + 2e4 + 3e5


def test_reader_refuses_what_read_restart_refuses(tmp_path):
    p = str(tmp_path / "r.dta")
    restart.write_restart(p, N, 3, _coded(N, 0), ritt=False)
    with pytest.raises(ValueError, match="dimensions"):
        restart.read_restart(p, (4, 3, 5, 2, 3, 5))
    data = bytearray(open(p, "rb").read())
    data[-1] ^= 1                      # broken trailing length marker
    open(p, "wb").write(bytes(data))
    with pytest.raises(ValueError, match="framing"):
        restart.read_restart(p)
    # PE bounds of a decomposed run (a block of x) are not this device's grid
    q = str(tmp_path / "q.dta")
    with open(q, "wb") as f:
        for payload in (np.array(list(N) + [3], "<i4").tobytes(), np.array([1, 2, 1, 3, 1, 5], "<i4").tobytes(),
                        np.array([1, 2, 1, 3, 1, 4], "<i4").tobytes(), np.zeros(3 * int(np.prod([v + 4 for v in N]))).tobytes()):
            f.write(np.array([len(payload)], "<i4").tobytes() + payload + np.array([len(payload)], "<i4").tobytes())
    with pytest.raises(ValueError, match="PE boundaries"):
        restart.read_restart(q, N)


def _gfortran_subrecords(payload, limit):
    """gfortran's framing of one record, assembled here from its rule: the
    leading marker of every subrecord but the last is negative, the trailing
    marker of every subrecord but the first is negative."""
    parts = [payload[i:i + limit] for i in range(0, len(payload), limit)] or [b""]
    out = b""
    for k, p in enumerate(parts):
        lead = -len(p) if k < len(parts) - 1 else len(p)
        trail = -len(p) if k > 0 else len(p)
        out += np.array([lead], "<i4").tobytes() + p + np.array([trail], "<i4").tobytes()
    return out


@pytest.mark.parametrize("limit", [100, 8 * 37, 977])
def test_long_records_split_into_gfortran_subrecords(tmp_path, limit):
    """Records longer than the subrecord limit (2,147,483,639 bytes by
    default, which the field record passes from prod(n+4) > 89.5M) are split
    as gfortran splits them; a small limit checks the split byte for byte,
    including a limit that cuts a binary64 value and one that divides the
    record, and the reader joins the subrecords again."""
    E, dE, mf = _coded(N, 1e6), _coded(N, 2e6), _coded(N, 3e6)
    p = tmp_path / "r.dta"
    restart.write_restart(str(p), N, 5, E, dE, mf, ritt=False, max_subrecord=limit)
    H = [v + 4 for v in N]
    expect = b""
    for payload in (np.array(list(N) + [5], "<i4").tobytes(),
                     np.array([1, N[0], 1, N[1], 1, N[2]], "<i4").tobytes(),
                     np.array([1, N[3], 1, N[4], 1, N[5]], "<i4").tobytes()):
        expect += _gfortran_subrecords(payload, limit)
    body = b""
    for a in (E, dE, mf):
        h = np.zeros(H, order="F")
        h[2:-2, 2:-2, 2:-2, 2:-2, 2:-2, 2:-2] = a.reshape(N, order="F")
        body += h.tobytes(order="F")
    expect += _gfortran_subrecords(body, limit)
    data = p.read_bytes()
    assert data == expect
    assert np.frombuffer(data[:4], "<i4")[0] == 28          # short records stay whole
    r = restart.read_restart(str(p), N)
    assert r["itt"] == 5
    assert np.array_equal(r["E"], E) and np.array_equal(r["dE"], dE) and np.array_equal(r["meanfl"], mf)
    # a continuation flag that contradicts the framing is refused
    bad = bytearray(data)
    at = len(data) - len(_gfortran_subrecords(body, limit))
    bad[at:at + 4] = np.array([abs(int(np.frombuffer(data[at:at + 4], "<i4")[0]))], "<i4").tobytes()
    (tmp_path / "bad.dta").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="framing|records"):
        restart.read_restart(str(tmp_path / "bad.dta"))


def test_default_limit_is_gfortrans():
    assert restart.MAX_SUBRECORD == 2 ** 31 - 9
