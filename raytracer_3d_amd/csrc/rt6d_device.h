// rt6d_device.h -- device-side building blocks of the 6D step (gfx950).
//
// Layout in HBM: every 6D field is stored without halos, Fortran order over
// (x,y,z,k,l,m), linear index lin = i + nx*(j + ny*(k + ...)) with 0-based
// coordinates.  Face fields of axis d (group velocities, fluxes) have n_d+1
// entries along d: face f lies between cells f and f+1 (1-based), face 0 and
// face n_d are the domain boundaries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt6d {

constexpr int kAxes = 6;

// Unsigned 32-bit division by a run-time constant: q = (umulhi(x, mul) + x) >> shift
// (round-up multiplier method, Granlund & Montgomery 1994).  Valid for x < 2^31.
struct FastDiv {
    uint32_t d, mul, shift;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.shift = l;
    f.mul = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t x, const FastDiv &f)
{
    return (__umulhi(x, f.mul) + x) >> f.shift;
}

// Geometry of one integration, passed by value to kernels.
struct Geo {
    int n[kAxes];        // interior extents
    int inner[kAxes];    // product of the extents of the axes below d
    FastDiv div[5];      // division by n[0..4]
    int act;             // bit d set when n[d] > 1 (the axis is advected)
    int up, down;        // z reflection switches
    uint32_t ncells;
    int uxy;             // bit 0 (1): every x (y) spacing the step reads is the same
};

// Decode lin into 1-based coordinates c[d] and quotients q[d] = lin / inner[d].
__device__ __forceinline__ void decode(const Geo &g, uint32_t lin, int c[kAxes], uint32_t q[kAxes + 1])
{
    q[0] = lin;
#pragma unroll
    for (int d = 0; d < 5; ++d) {
        q[d + 1] = fdiv(q[d], g.div[d]);
        c[d] = (int)(q[d] - q[d + 1] * (uint32_t)g.n[d]) + 1;
    }
    c[5] = (int)q[5] + 1;
    q[6] = 0;
}

// Offset (relative to the cell) of the neighbour at coordinate c+o along axis
// d, after the ghost treatment of integrate.f90:14-94: zero-gradient copies on
// every open side, and on z the m-mirrored copy of reflection_condition
// (parallel_mpi.f90:837-841,867-871) when that side reflects.
template <int D>
__device__ __forceinline__ int64_t nbr_off(const Geo &g, const int c[kAxes], int o)
{
    int cc = c[D] + o;
    const int n = g.n[D];
    if (D == 2) {
        if (cc < 1) {
            if (g.down) {
                int zr = 1 - cc;                    // E(1-i) = E(i, nm+1-m)
                int mr = g.n[5] + 1 - c[5];
                return (int64_t)(zr - c[2]) * g.inner[2] + (int64_t)(mr - c[5]) * g.inner[5];
            }
            cc = 1;
        } else if (cc > n) {
            if (g.up) {
                int zr = 2 * n + 1 - cc;            // E(nz+i) = E(nz-i+1, nm+1-m)
                int mr = g.n[5] + 1 - c[5];
                return (int64_t)(zr - c[2]) * g.inner[2] + (int64_t)(mr - c[5]) * g.inner[5];
            }
            cc = n;
        }
    } else {
        cc = cc < 1 ? 1 : (cc > n ? n : cc);
    }
    return (int64_t)(cc - c[D]) * g.inner[D];
}

// Superbee limiter of superbee_fluxes.f90:10.
template <typename T>
__device__ __forceinline__ T sb_limiter(T cr)
{
    T a = T(2) * cr;
    a = a < T(1) ? a : T(1);
    T b = cr < T(2) ? cr : T(2);
    T c = a > b ? a : b;
    return c > T(0) ? c : T(0);
}

// Face flux in the reference's operation order (superbee_fluxes.f90:19-30):
// IEEE divides, the single-precision 1.E20 of :26, no contraction.
__device__ __forceinline__ double sb_flux_exact(double v, double dt, double dxf,
                                                double Em1, double E0, double E1, double E2)
{
#pragma clang fp contract(off)
    const double big = (double)1.0e20f;
    double ucfl = fabs(v * dt / dxf);
    double Rjp = E2 - E1;
    double Rj = E1 - E0;
    double Rjm = E0 - Em1;
    double up = v > 0.0 ? Rjm : Rjp;
    double cr;
    if (Rj != 0.0) cr = up / Rj;
    else cr = up * big;
    cr = sb_limiter(cr);
    return v * (E1 + E0) * 0.5 - fabs(v) * ((1.0 - cr) + ucfl * cr) * Rj * 0.5;
}

// x ^ (y & 0x80000000) in one instruction: gfx950's three-input bitwise
// operation with the truth table of s0 ^ (s1 & s2), indexed by
// (s0 << 2) | (s1 << 1) | s2 (0xf0 ^ (0xcc & 0xaa) = 0x78)
__device__ __forceinline__ uint32_t xor_sign_of(uint32_t x, uint32_t y)
{
    return __builtin_amdgcn_bitop3_b32(x, y, 0x80000000u, 0x78);
}

// a with its sign flipped where b is negative (a * sign(b), exact)
__device__ __forceinline__ double flip_sign_by(double a, double b)
{
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const uint32_t hi = xor_sign_of((uint32_t)(ua >> 32), (uint32_t)(ub >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (ua & 0xffffffffull));
}
__device__ __forceinline__ float flip_sign_by(float a, float b)
{
    return __builtin_bit_cast(float, xor_sign_of(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b)));
}

// a with the sign of b, in T (the generic builtin is binary64: a binary32
// operand would be widened first)
__device__ __forceinline__ double copysign_t(double a, double b) { return __builtin_copysign(a, b); }
__device__ __forceinline__ float copysign_t(float a, float b) { return __builtin_copysignf(a, b); }

// a*b + c with one rounding, in T
template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c)
{
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// The same flux without a divide, in upwind form.  With R = E1-E0, Ru the
// upwind difference and r = Ru/R, superbee's phi(r)*R = sign(R) * max(0,
// min(2*min(|R|,su), max(|R|,su))) with su = Ru*sign(R): min/max/doubling
// only, so phi(r)*R is exact (one of 0, R, Ru, 2R, 2Ru) and 0 for R == 0 as
// in the reference (where the anti-diffusive term is multiplied by Rj = 0).
// The reference's v/2*(E1+E0) - |v|/2*R is v*E_up, E_up the upwind cell (E0
// for v > 0, else E1), here fma(vp, E0, vn*E1) = v*E_up rounded once, with
// vp = max(v,0) and vn = min(v,0), so that
//   F = v*E_up + w*phiR,  w = |v|/2*(1-c),  c = |v|*dt/dx.
// vp, vn, w and the upwind side depend on the face's velocity and spacing
// only (sb_coef): kernels that read them from a table and kernels that form
// them per face give the same bits.  On a boundary face with a zero-gradient
// ghost, R = 0 and F = v*E_up exactly.
template <typename T>
struct SbCoef {
    T vp, vn, w;
    bool pos;   // v > 0: the upwind difference is E0 - Em1
};

template <typename T>
__device__ __forceinline__ SbCoef<T> sb_coef(T v, T cfl_coef)
{
    const T a = fabs(v);
    SbCoef<T> k;
    k.vp = fmax(v, T(0));
    k.vn = fmin(v, T(0));
    k.w = (T(0.5) * a) * fma_t<T>(-a, cfl_coef, T(1));
    k.pos = v > T(0);
    return k;
}

template <typename T>
__device__ __forceinline__ T sb_flux_c(const SbCoef<T> &k, T Em1, T E0, T E1, T E2)
{
    const T R = E1 - E0;
    const T Ru = k.pos ? (E0 - Em1) : (E2 - E1);
    const T su = flip_sign_by(Ru, R);
    const T aR = fabs(R);
    const T mn = fmin(aR, su), mx = fmax(aR, su);
    T p;
    if constexpr (sizeof(T) == 4) p = __builtin_amdgcn_fmed3f(mn + mn, T(0), mx);   // mx >= 0
    else p = fmax(fmin(mn + mn, mx), T(0));
    const T phiR = copysign_t(p, R);
    return fma_t<T>(k.w, phiR, fma_t<T>(k.vp, E0, k.vn * E1));
}

// The same flux with the upwind side fixed (POS: v > 0): Ru without a select,
// and v*E_up as one product.  Where v*E_up is not zero it equals
// fma(vp, E0, vn*E1) of sb_flux_c exactly (the other product is a zero); a
// zero product may differ in its sign only, which the sums that follow do
// not see (E + dt*dE is the same).
template <typename T, bool POS>
__device__ __forceinline__ T sb_flux_up(const SbCoef<T> &k, T Em1, T E0, T E1, T E2)
{
    const T R = E1 - E0;
    const T Ru = POS ? (E0 - Em1) : (E2 - E1);
    const T su = flip_sign_by(Ru, R);
    const T aR = fabs(R);
    const T mn = fmin(aR, su), mx = fmax(aR, su);
    T p;
    if constexpr (sizeof(T) == 4) p = __builtin_amdgcn_fmed3f(mn + mn, T(0), mx);   // mx >= 0
    else p = fmax(fmin(mn + mn, mx), T(0));
    const T phiR = copysign_t(p, R);
    return fma_t<T>(k.w, phiR, POS ? k.vp * E0 : k.vn * E1);
}

// The upwind side of a group of faces that share the sign of v: fixed at
// compile time on each side of a wave-uniform branch (SignC<true/false>), or
// selected per face (SignSel).
template <bool B>
struct SignC {
    static constexpr bool sel = false, value = B;
};
struct SignSel {
    static constexpr bool sel = true, value = false;
};

template <typename T, typename S>
__device__ __forceinline__ T sb_flux_s(const SbCoef<T> &k, T Em1, T E0, T E1, T E2)
{
    if constexpr (S::sel) return sb_flux_c<T>(k, Em1, E0, E1, E2);
    else return sb_flux_up<T, S::value>(k, Em1, E0, E1, E2);
}

// f(S) for faces whose v has the sign of pos: with BR a branch on the sign
// and the upwind side fixed on each side, else the per-face select.
template <bool BR, typename F>
__device__ __forceinline__ void on_sign(bool pos, F &&f)
{
    if constexpr (BR) {
        if (pos) f(SignC<true>{});
        else f(SignC<false>{});
    } else {
        f(SignSel{});
    }
}

template <typename T>
__device__ __forceinline__ T sb_flux_fast(T v, T cfl_coef, T Em1, T E0, T E1, T E2)
{
    return sb_flux_c<T>(sb_coef<T>(v, cfl_coef), Em1, E0, E1, E2);
}

}  // namespace rt6d
