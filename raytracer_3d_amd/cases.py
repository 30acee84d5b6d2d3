"""Synthetic integrations (the set_parameter / set_grid / set_initial_conditions
hooks of a reference config plug-in, e.g. for_config/gm1.f90:7-118) and the
binary case-file format shared by the Fortran driver, the reference harness
and the tests.

All randomness comes from one counter-based generator (splitmix64 keyed by
seed, member and linear cell index), evaluated on the host with numpy, so
every consumer of a case receives bitwise-identical inputs.

Arrays follow the reference's Fortran order (x fastest) and are held as flat
numpy float64 vectors:
  * ``E0``: interior, nx*ny*nz*nk*nl*nm;
  * ``Nsqr``, ``U``, ``V``: halo'd 3D, (nx+4)*(ny+4)*(nz+4) (onx = 2);
  * grids: halo'd 1D, n+4 per axis.
"""
from __future__ import annotations

import dataclasses
import math
import os
import struct

import numpy as np

ONX = 2
AXES = ("x", "y", "z", "k", "l", "m")
CASE_MAGIC = 1144149074  # b"RT6D" little-endian
CASE_VERSION = 1


# --------------------------------------------------------------------- PRNG

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, member: int, idx: np.ndarray, stream: int = 0) -> np.ndarray:
    """U[0,1) doubles keyed by (seed, member, stream, linear index)."""
    with np.errstate(over="ignore"):
        key = np.uint64((seed * 0x100000001B3 + member * 0x2545F4914F6CDD1D + stream * 0x27BB2EE687B0B0FD)
                        & 0xFFFFFFFFFFFFFFFF)
        h = splitmix64(np.asarray(idx, dtype=np.uint64) * _GOLD ^ key)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, member: int, idx: np.ndarray, stream: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller from two uniform streams."""
    u1 = uniform(seed, member, idx, 2 * stream + 1)
    u2 = uniform(seed, member, idx, 2 * stream + 2)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


# --------------------------------------------------------------------- grid


def calc_grid(axis: int, spacing: np.ndarray, fixed: float = 0.0):
    """One axis of calc_grid (for_src/calc_grid.f90:104-177, u_centered_grid
    :300-322).  Returns halo'd (t, dt, u, du), n+4 values each."""
    d_int = np.asarray(spacing, dtype=np.float64)
    n = d_int.size
    N = n + 2 * ONX
    D = np.zeros(N)
    D[2:n + 2] = d_int
    D[n + 2:] = D[n + 1]
    D[:2] = D[2]
    U = np.zeros(N)
    T = np.zeros(N)
    Du = np.zeros(N)
    for p in range(1, N):
        U[p] = U[p - 1] + D[p]
    T[0] = U[0] - D[0] * 0.5
    for p in range(1, N):
        T[p] = 2.0 * U[p - 1] - T[p - 1]
    for p in range(N - 1):
        Du[p] = T[p + 1] - T[p]
    Du[N - 1] = 2.0 * D[N - 1] - Du[max(N - 2, 0)]
    if axis <= 1:
        sh = Du[2]
    elif axis == 2:
        sh = U[n + 1]
    elif n > 1:
        sh = U[n // 2 + 1]
    else:
        return np.full(N, float(fixed)), D, np.full(N, float(fixed)), Du
    return T - sh, D, U - sh, Du


def check_grid(spacings) -> None:
    """Grid checks of calc_grid.f90:260-293: dz (T and U spacing) must not
    grow with the index, nor dm and dl over their first half."""
    def bad(a, lo, hi):
        for i in range(lo, hi + 1):
            if a[i + 1] - a[i] > 1e-12:
                return i
        return None

    for ax, half in ((2, False), (5, True), (4, True)):
        n = len(spacings[ax])
        _, dt, _, du = calc_grid(ax, spacings[ax])
        hi = n // 2 if half else n
        # Fortran i = 2..hi compares (i) with (i-1): halo slots i+1 and i
        for i in range(2, hi + 1):
            if du[i + 1] - du[i] > 1e-12 or dt[i + 1] - dt[i] > 1e-12:
                raise ValueError(f"grid check failed on axis {AXES[ax]} at index {i}: spacing grows")


# --------------------------------------------------------------------- case


@dataclasses.dataclass
class Case:
    name: str
    n: tuple
    spacing: list            # 6 arrays of interior T spacings
    Nsqr: np.ndarray         # halo'd 3D
    U: np.ndarray
    V: np.ndarray
    E0: np.ndarray           # interior 6D
    f0: float = 1e-4
    sign: float = 1.0
    fixed: tuple = (0.0, 0.0, 0.0)   # k_fixed, l_fixed, m_fixed
    reflect_up: bool = True
    reflect_down: bool = True
    dt: float | None = None
    nsteps: int = 1
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def cells(self) -> int:
        return int(np.prod(self.n))

    def grids(self):
        """halo'd (t, dt) per axis"""
        out = []
        for ax in range(6):
            fx = self.fixed[ax - 3] if ax >= 3 else 0.0
            t, d, _, _ = calc_grid(ax, self.spacing[ax], fx)
            out.append((t, d))
        return out

    def halo_positions(self, ax: int) -> np.ndarray:
        return self.grids()[ax][0]


def _halo3(n, fill_fn, grids):
    """Evaluate fill_fn(x, y, z) on the halo'd 3D grid (Fortran order, flat)."""
    x = grids[0][0]
    y = grids[1][0]
    z = grids[2][0]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    a = fill_fn(X, Y, Z)
    return np.asfortranarray(np.broadcast_to(a, X.shape)).ravel(order="F").astype(np.float64)


def _wavenumber_spacing(n: int, kmax: float) -> np.ndarray:
    """Uniform spacing covering [-kmax, kmax] with n cells."""
    return np.full(n, 2.0 * kmax / n)


def _interior_axes(case_n, grids):
    return [grids[a][0][2:2 + case_n[a]] for a in range(6)]


def packet_case(n=(32, 4, 32, 16, 4, 16), seed=1234, member=0, L=(100e3, 10e3, 4000.0),
                kmax=2 * math.pi / 1000.0, lmax=2 * math.pi / 1000.0, mmax=2 * math.pi / 50.0,
                N0=2e-3, f0=1e-4, x0=None, z0=-2000.0, sx=10e3, sz=500.0, k0=2 * math.pi / 10e3,
                m0=2 * math.pi / 400.0, peak=1e-3, noise=0.01, U0=0.0, Nscale=None, name=None) -> Case:
    """Gaussian wave packet in (x, z) and (k, m) with seeded 1% noise; constant
    N and no mean flow unless U0/Nscale are given (then N(z) = N0 exp(z/Nscale),
    U(z) = U0 tanh((z+500)/300))."""
    n = tuple(int(v) for v in n)
    sp = [np.full(n[0], L[0] / n[0]), np.full(n[1], L[1] / n[1]), np.full(n[2], L[2] / n[2]),
          _wavenumber_spacing(n[3], kmax), _wavenumber_spacing(n[4], lmax), _wavenumber_spacing(n[5], mmax)]
    c = Case(name or "packet", n, sp, None, None, None, None, f0=f0)
    g = c.grids()
    if Nscale:
        Nsq = _halo3(n, lambda X, Y, Z: (N0 * np.exp(Z / Nscale)) ** 2, g)
    else:
        Nsq = _halo3(n, lambda X, Y, Z: np.full_like(X, N0 * N0), g)
    if U0:
        Uf = _halo3(n, lambda X, Y, Z: U0 * np.tanh((Z + 500.0) / 300.0), g)
    else:
        Uf = np.zeros_like(Nsq)
    c.Nsqr, c.U, c.V = Nsq, Uf, np.zeros_like(Nsq)
    xs, ys, zs, ks, ls, ms = _interior_axes(n, g)
    xc = 0.5 * L[0] if x0 is None else x0
    gx = np.exp(-((xs - xc) ** 2) / (2 * sx * sx))
    gz = np.exp(-((zs - z0) ** 2) / (2 * sz * sz))
    gk = np.exp(-((ks - k0) ** 2) / (2 * (0.3 * k0) ** 2))
    gm = np.exp(-((ms - m0) ** 2) / (2 * (0.3 * m0) ** 2))
    E = peak * np.einsum("i,j,k,l,n,m->ijklnm", gx, np.ones(n[1]), gz, gk, np.ones(n[4]), gm)
    E = E.ravel(order="F")
    u = uniform(seed, member, np.arange(E.size, dtype=np.uint64))
    c.E0 = E * (1.0 + noise * (2.0 * u - 1.0))
    c.meta = dict(kind="packet", seed=seed, member=member, L=L, N0=N0, U0=U0)
    return c


def gm_case(n=(24, 24, 24, 16, 16, 16), seed=2024, L=(200e3, 200e3, 4000.0), kmax=2 * math.pi / 1000.0,
            lmax=2 * math.pi / 1000.0, mmax=2 * math.pi / 50.0, N0=5.2e-3, b=1300.0, jstar=3.0, E0=3e-3,
            U0=0.2, f0=1e-4, sigma=0.3, name=None) -> Case:
    """Garrett-Munk-shaped spectrum in (k,l,m) under N(z) = N0 exp(z/b) and
    U(z) = U0 tanh((z+500)/300), times a seeded lognormal (x,y) modulation.

    S(z,k,l,m) = A(m) B(omega) J N/N0 with A = (2/pi)/m*/(1+(m/m*)^2),
    m* = pi j* N/(b N0), B = (2/pi) f/(omega sqrt(omega^2-f^2)),
    J = (N^2-omega^2)^2/(m^2 omega (N^2-f^2)) (the shape of for_config/gm1.f90:96-109),
    normalised so that the domain mean of sum_klm E dk dl dm is E0."""
    n = tuple(int(v) for v in n)
    sp = [np.full(n[0], L[0] / n[0]), np.full(n[1], L[1] / n[1]), np.full(n[2], L[2] / n[2]),
          _wavenumber_spacing(n[3], kmax), _wavenumber_spacing(n[4], lmax), _wavenumber_spacing(n[5], mmax)]
    c = Case(name or "gm", n, sp, None, None, None, None, f0=f0)
    g = c.grids()
    c.Nsqr = _halo3(n, lambda X, Y, Z: (N0 * np.exp(Z / b)) ** 2, g)
    c.U = _halo3(n, lambda X, Y, Z: U0 * np.tanh((Z + 500.0) / 300.0), g)
    c.V = np.zeros_like(c.U)
    xs, ys, zs, ks, ls, ms = _interior_axes(n, g)
    Nz = N0 * np.exp(zs / b)
    Z, K, Lw, M = np.meshgrid(Nz, ks, ls, ms, indexing="ij")
    kh2 = K * K + Lw * Lw
    om = np.sqrt((Z * Z * kh2 + f0 * f0 * M * M) / (kh2 + M * M))
    mstar = math.pi * jstar * Z / (b * N0)
    A = (2.0 / math.pi) / mstar / (1.0 + (M / mstar) ** 2)
    Bw = (2.0 / math.pi) * f0 / (om * np.sqrt(np.maximum(om * om - f0 * f0, 1e-30)))
    J = (Z * Z - om * om) ** 2 / (M * M * om * (Z * Z - f0 * f0))
    S = A * Bw * J * Z / N0
    dk, dl, dm = sp[3][0], sp[4][0], sp[5][0]
    col = S.sum(axis=(1, 2, 3)) * dk * dl * dm      # per z
    S = S * (E0 / col.mean())
    gxy = normal(seed, 0, np.arange(n[0] * n[1], dtype=np.uint64)).reshape(n[0], n[1], order="F")
    mod = np.exp(sigma * gxy - 0.5 * sigma * sigma)
    E = np.einsum("ij,klnm->ijklnm", mod, S)
    c.E0 = E.ravel(order="F")
    c.meta = dict(kind="gm", seed=seed, L=L, N0=N0, b=b, jstar=jstar, E0=E0, U0=U0, sigma=sigma)
    return c


def random_case(n, seed=7, reflect_up=True, reflect_down=True, horizontal=True, name=None,
                fixed=(0.0, 0.0, 0.0)) -> Case:
    """Small test integration exercising every branch: non-uniform spacings
    (obeying calc_grid's checks), N^2, U, V varying in x, y and z (so kdot,
    ldot, lambda_x, lambda_y are non-zero when ``horizontal``), random E."""
    n = tuple(int(v) for v in n)
    rng = np.random.default_rng(seed)
    sp = []
    for ax in range(6):
        m = n[ax]
        # x, y, k non-uniform; z, l, m uniform (calc_grid.f90:260-293 asks
        # their T and U spacings not to grow, which a u-centred grid built from
        # a non-uniform T spacing generally violates)
        if ax in (0, 1):
            s = 1000.0 * (1.0 + 0.3 * rng.random(m))
        elif ax == 3:
            s = 2e-4 * (1.0 + 0.4 * rng.random(m))
        else:
            s = np.full(m, {2: 50.0, 4: 3e-4, 5: 4e-3}[ax])
        sp.append(s)
    c = Case(name or "random", n, sp, None, None, None, None, f0=1e-4, fixed=fixed,
             reflect_up=reflect_up, reflect_down=reflect_down)
    g = c.grids()
    hz = 1.0 if horizontal else 0.0
    c.Nsqr = _halo3(n, lambda X, Y, Z: (3e-3 * (1.0 + 0.2 * np.sin(X / 3000.0 * hz) + 0.1 * np.cos(Y / 2500.0 * hz))
                                         * np.exp(Z / 800.0)) ** 2, g)
    c.U = _halo3(n, lambda X, Y, Z: 0.15 * np.tanh((Z + 100.0) / 60.0) * (1.0 + 0.3 * hz * np.sin(X / 2000.0 + Y / 3000.0)), g)
    c.V = _halo3(n, lambda X, Y, Z: 0.05 * np.cos(Z / 70.0) * (1.0 + 0.2 * hz * np.cos(X / 1700.0 - Y / 2100.0)), g)
    E = 1e-3 * (0.5 + rng.random(c.cells))
    # a few flat stretches (R = 0 faces) and extrema exercise every limiter branch
    idx = np.arange(0, E.size - 1, 7)
    E[idx] = E[idx + 1]
    c.E0 = E
    check_grid(sp)
    c.meta = dict(kind="random", seed=seed, horizontal=horizontal)
    return c


def member_case(seed: int, member: int, fp32: bool = False) -> Case:
    """One ensemble member: grid sizes from {12,16,20,24} (x,y,z) and
    {8,12,16} (k,l,m); N0 ~ U[2e-3, 8e-3] with N(z) = N0 exp(z/1300),
    U0 ~ U[0, 0.3] tanh shear, Gaussian packet with seeded centres."""
    u = uniform(seed, member, np.arange(16, dtype=np.uint64), stream=7)
    nxyz = [12, 16, 20, 24]
    nklm = [8, 12, 16]
    n = (nxyz[int(u[0] * 4)], nxyz[int(u[1] * 4)], nxyz[int(u[2] * 4)],
         nklm[int(u[3] * 3)], nklm[int(u[4] * 3)], nklm[int(u[5] * 3)])
    N0 = 2e-3 + 6e-3 * u[6]
    U0 = 0.3 * u[7]
    L = (100e3, 100e3, 4000.0)
    x0 = L[0] * (0.25 + 0.5 * u[8])
    z0 = -L[2] * (0.25 + 0.5 * u[9])
    k0 = 2 * math.pi / (5e3 + 15e3 * u[10])
    m0 = 2 * math.pi / (200.0 + 600.0 * u[11])
    c = packet_case(n=n, seed=seed, member=member, L=L, N0=N0, U0=U0, Nscale=1300.0, x0=x0, z0=z0, k0=k0, m0=m0,
                    name=f"member{member}")
    c.meta.update(dict(kind="member", N0=N0, U0=U0, x0=x0, z0=z0, k0=k0, m0=m0, fp32=fp32))
    return c


def separable_E(factors, peak: float, noise: float, seed: int, member: int) -> np.ndarray:
    """E = peak*fx*fy*fz*fk*fl*fm*(1 + noise*(2u-1)), products left to right
    in binary64, u = uniform(seed, member, linear index): the host statement of
    rt6d_init_separable (Fortran order, x fastest)."""
    A = peak * np.asarray(factors[0], dtype=np.float64)
    for f in factors[1:]:
        A = A[..., None] * np.asarray(f, dtype=np.float64)
    E = A.ravel(order="F")
    u = uniform(seed, member, np.arange(E.size, dtype=np.uint64))
    return E * (1.0 + noise * (2.0 * u - 1.0))


def member_spec(seed: int, member: int, n=None) -> Case:
    """One ensemble or serving member without a host E: grid sizes from
    {12,16,20,24} (x,y,z) and {8,12,16} (k,l,m) unless ``n`` is given,
    N0 ~ U[2e-3, 8e-3] with N(z) = N0 exp(z/1300), U0 ~ U[0, 0.3] tanh shear,
    a Gaussian packet in (x,z,k,m) with seeded centres and 1% noise, held as
    separable factors (``meta["init"]``) for rt6d_init_separable."""
    u = uniform(seed, member, np.arange(16, dtype=np.uint64), stream=7)
    nxyz = [12, 16, 20, 24]
    nklm = [8, 12, 16]
    if n is None:
        n = (nxyz[int(u[0] * 4)], nxyz[int(u[1] * 4)], nxyz[int(u[2] * 4)],
             nklm[int(u[3] * 3)], nklm[int(u[4] * 3)], nklm[int(u[5] * 3)])
    n = tuple(int(v) for v in n)
    N0 = 2e-3 + 6e-3 * u[6]
    U0 = 0.3 * u[7]
    L = (100e3, 100e3, 4000.0)
    x0 = L[0] * (0.25 + 0.5 * u[8])
    z0 = -L[2] * (0.25 + 0.5 * u[9])
    k0 = 2 * math.pi / (5e3 + 15e3 * u[10])
    m0 = 2 * math.pi / (200.0 + 600.0 * u[11])
    kmax, lmax, mmax = 2 * math.pi / 1000.0, 2 * math.pi / 1000.0, 2 * math.pi / 50.0
    sp = [np.full(n[0], L[0] / n[0]), np.full(n[1], L[1] / n[1]), np.full(n[2], L[2] / n[2]),
          _wavenumber_spacing(n[3], kmax), _wavenumber_spacing(n[4], lmax), _wavenumber_spacing(n[5], mmax)]
    c = Case(f"member{member}", n, sp, None, None, None, None, f0=1e-4)
    g = c.grids()
    c.Nsqr = _halo3(n, lambda X, Y, Z: (N0 * np.exp(Z / 1300.0)) ** 2, g)
    c.U = _halo3(n, lambda X, Y, Z: U0 * np.tanh((Z + 500.0) / 300.0), g)
    c.V = np.zeros_like(c.Nsqr)
    xs, ys, zs, ks, ls, ms = _interior_axes(n, g)
    sx, sz = 10e3, 500.0
    factors = [np.exp(-((xs - x0) ** 2) / (2 * sx * sx)), np.ones(n[1]),
               np.exp(-((zs - z0) ** 2) / (2 * sz * sz)),
               np.exp(-((ks - k0) ** 2) / (2 * (0.3 * k0) ** 2)), np.ones(n[4]),
               np.exp(-((ms - m0) ** 2) / (2 * (0.3 * m0) ** 2))]
    c.meta = dict(kind="member_spec", seed=seed, member=member, N0=N0, U0=U0, x0=x0, z0=z0, k0=k0, m0=m0,
                  init=dict(factors=factors, peak=1e-3, noise=0.01, seed=seed, member=member))
    return c


# ---------------------------------------------------------------- case file


def write_case_file(case: Case, path: str, nsteps: int | None = None, dt: float | None = None) -> None:
    """Binary case file: int32[16] header, float64[8] parameters, the six
    interior spacings, Nsqr/U/V halo'd 3D, interior E (all little-endian)."""
    dt = case.dt if dt is None else dt
    if dt is None:
        raise ValueError("case has no dt")
    nst = case.nsteps if nsteps is None else nsteps
    hdr = [CASE_MAGIC, CASE_VERSION, *case.n, int(case.reflect_up), int(case.reflect_down), int(nst), 0, 0, 0, 0, 0]
    par = [dt, case.f0, case.sign, *case.fixed, 0.0, 0.0]
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(struct.pack("<16i", *hdr))
        f.write(struct.pack("<8d", *par))
        for s in case.spacing:
            f.write(np.ascontiguousarray(s, dtype="<f8").tobytes())
        for a in (case.Nsqr, case.U, case.V):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(case.E0, dtype="<f8").tobytes())
    os.replace(tmp, path)


def read_case_file(path: str) -> Case:
    with open(path, "rb") as f:
        hdr = struct.unpack("<16i", f.read(64))
        if hdr[0] != CASE_MAGIC:
            raise ValueError("not a case file")
        par = struct.unpack("<8d", f.read(64))
        n = tuple(hdr[2:8])
        sp = [np.frombuffer(f.read(8 * n[a]), dtype="<f8").copy() for a in range(6)]
        n3 = (n[0] + 4) * (n[1] + 4) * (n[2] + 4)
        bg = [np.frombuffer(f.read(8 * n3), dtype="<f8").copy() for _ in range(3)]
        E = np.frombuffer(f.read(8 * int(np.prod(n))), dtype="<f8").copy()
    return Case("file", n, sp, bg[0], bg[1], bg[2], E, f0=par[1], sign=par[2], fixed=tuple(par[3:6]),
                reflect_up=bool(hdr[8]), reflect_down=bool(hdr[9]), dt=par[0], nsteps=hdr[10])


REF_FIELDS = (
    ["omega_i", "omega_e"]
    + [f"{a}dot" for a in AXES]
    + ["lambda_x", "lambda_y", "lambda_z", "E", "dE", "meanfl"]
    + [f"flux_{a}" for a in AXES]
)


def face_count(n, d) -> int:
    return int(np.prod(n)) // n[d] * (n[d] + 1)


def read_ref_output(path: str, n) -> dict:
    """Read the stream written by oracle/ref_harness.f90."""
    n = tuple(n)
    cells = int(np.prod(n))
    out = {}
    with open(path, "rb") as f:
        for a in range(6):
            h = n[a] + 4
            t, dt, u, du = (np.frombuffer(f.read(8 * h), dtype="<f8").copy() for _ in range(4))
            out[f"grid_{AXES[a]}"] = np.stack([t, dt, u, du])
        for name in REF_FIELDS:
            if name.endswith("dot"):
                cnt = face_count(n, AXES.index(name[0]))
            elif name.startswith("flux_"):
                cnt = face_count(n, AXES.index(name[5]))
            else:
                cnt = cells
            out[name] = np.frombuffer(f.read(8 * cnt), dtype="<f8").copy()
    return out
