// rt6d_kernels.h -- kernel templates of the 6D step, instantiated in
// rt6d_exact.hip (compiled with -ffp-contract=off: reference operation order)
// and rt6d_fast.hip (reciprocals, division-free limiter, FMA contraction).
#pragma once
#include "rt6d_device.h"

namespace rt6d {

// Per-axis 1D coefficient block in device memory, n_d+1 entries each:
//   dx[f]   = dxt_d(f), f = 0..n_d (f = 0 is the ghost spacing)
//   rdx[f]  = 1 / dx[f]
//   cfl[f]  = dt / dx[f]  (recomputed when dt changes)
struct AxisCoef {
    const double *dx[kAxes];
    const double *rdx[kAxes];
    const double *cfl[kAxes];
    const double *cr[kAxes];   // (cfl[f], rdx[f]) interleaved: one wide scalar load per face
};

// Horizontally homogeneous coefficient tables over the 4D (z,k,l,m) space
// (q2 = lin / (nx*ny)):
//   vx3[cls*n4 + q2], vy3 likewise: cls 0 face 0, 1 interior faces, 2 face n
//   vz[(q2 + q3)]  faces along z: extent nz+1
//   vm[q2 + f*inner4m] faces along m: extent nm+1
//   lam[q2]
//   rec[kRec*P ..]  per (z,k,l,m) plane P = lin / (nx*ny): what one step of
//                   a cell of that plane reads, side by side, so that a wave
//                   fetches it in a few wide scalar loads: the fast-arithmetic
//                   face coefficients (vp, vn, w of sb_coef, for the step's dt)
//                   of the interior x and y faces (with the spacing of face
//                   1, which is every face's where the spacing is uniform),
//                   the z faces below and above and the m face above; the
//                   velocities of the x and y boundary faces; lam; bit
//                   flags of which coefficient sets have v > 0; and, in the
//                   record's precision, 1/dz of the plane's z cell, 1/dm of
//                   its m cell and the step's dt
constexpr int kRec = 24;
enum RecField {
    kRecXi = 0, kRecYi = 3, kRecZlo = 6, kRecZhi = 9, kRecMhi = 12,   // (vp, vn, w) each
    kRecVx0 = 15, kRecVxn = 16, kRecVy0 = 17, kRecVyn = 18, kRecLam = 19, kRecFlags = 20,
    kRecRz = 21, kRecRm = 22, kRecDt = 23
};
// flag bits of kRecFlags (in the low bits of that slot's bit pattern)
enum RecFlag { kPosXi = 1, kPosYi = 2, kPosZlo = 4, kPosZhi = 8, kPosMhi = 16 };
template <typename T>
struct HomTables {
    const T *vx3, *vy3, *vz, *vm, *lam;
    const T *rec;
    uint32_t n4, inner4m;
};

template <typename T>
struct StepPtrs {
    const T *Ein;
    T *Eout;
    const T *vel[kAxes];   // general path: face arrays of each axis
    const T *lam;          // general path: lambda summed over active x,y,z
    HomTables<T> hom;
    AxisCoef ax;
    double dt;
    // tendency outputs (dE, meanfl, face fluxes), used by the tendency kernel
    T *dE, *meanfl;
    T *flux[kAxes];
};

// Map a hardware block id to a logical block so that the 8 XCDs (hardware
// round-robins blocks over them) each work on one contiguous eighth of the
// cell range: neighbours along the strided axes then share an XCD's L2.
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nb_logical)
{
    const uint32_t per = (nb_logical + 7u) / 8u;
    return (b & 7u) * per + (b >> 3);
}

template <int D, typename T, bool HOMOG>
__device__ __forceinline__ void load_vel(const Geo &g, const StepPtrs<T> &p, uint32_t lin,
                                         const int c[kAxes], const uint32_t q[kAxes + 1], T &vL, T &vR)
{
    if (!HOMOG) {
        const uint32_t fl = lin + (uint32_t)g.inner[D] * q[D + 1];
        vL = p.vel[D][fl];
        vR = p.vel[D][fl + (uint32_t)g.inner[D]];
    } else {
        const uint32_t q2 = q[2];
        if (D == 0 || D == 1) {
            const T *t = D == 0 ? p.hom.vx3 : p.hom.vy3;
            const int cd = c[D], nd = g.n[D];
            vL = t[(cd == 1 ? 0u : 1u) * p.hom.n4 + q2];
            vR = t[(cd == nd ? 2u : 1u) * p.hom.n4 + q2];
        } else if (D == 2) {
            const uint32_t fl = q2 + q[3];
            vL = p.hom.vz[fl];
            vR = p.hom.vz[fl + 1];
        } else {  // D == 5; k and l have no HOMOG velocities (identically zero)
            vL = p.hom.vm[q2];
            vR = p.hom.vm[q2 + p.hom.inner4m];
        }
    }
}

// Contribution of axis D to the tendency of one cell: loads the four
// neighbours, computes the two face fluxes and subtracts their divergence
// from dE (integrate.f90:105-200 order).  FL/FR receive the face fluxes.
template <int D, typename T, bool EXACT, bool HOMOG>
__device__ __forceinline__ void axis_term(const Geo &g, const StepPtrs<T> &p, const T *__restrict__ E,
                                          uint32_t lin, const int c[kAxes], const uint32_t q[kAxes + 1],
                                          T e0, T &dE, T &FL, T &FR)
{
    const T em2 = E[(int64_t)lin + nbr_off<D>(g, c, -2)];
    const T em1 = E[(int64_t)lin + nbr_off<D>(g, c, -1)];
    const T ep1 = E[(int64_t)lin + nbr_off<D>(g, c, 1)];
    const T ep2 = E[(int64_t)lin + nbr_off<D>(g, c, 2)];
    T vL, vR;
    load_vel<D, T, HOMOG>(g, p, lin, c, q, vL, vR);
    const int cd = c[D];
    if constexpr (EXACT) {
        const double *dx = p.ax.dx[D];
        FL = sb_flux_exact(vL, p.dt, dx[cd - 1], em2, em1, e0, ep1);
        FR = sb_flux_exact(vR, p.dt, dx[cd], em1, e0, ep1, ep2);
        dE = dE - (FR - FL) / dx[cd];
    } else {
        const double *cf = p.ax.cfl[D];
        FL = sb_flux_fast<T>(vL, (T)cf[cd - 1], em2, em1, e0, ep1);
        FR = sb_flux_fast<T>(vR, (T)cf[cd], em1, e0, ep1, ep2);
        dE = dE - (FR - FL) * (T)p.ax.rdx[D][cd];
    }
}

// Whole tendency of one cell: dE (divergence minus meanfl) and meanfl.
// fl/fr receive the face fluxes of each axis (zero when the axis is skipped).
template <typename T, bool EXACT, bool HOMOG>
__device__ __forceinline__ void cell_tendency(const Geo &g, const StepPtrs<T> &p, const T *__restrict__ E,
                                              uint32_t lin, const int c[kAxes], const uint32_t q[kAxes + 1],
                                              T e0, T &dE, T &mf, T fl[kAxes], T fr[kAxes])
{
    dE = T(0);
#pragma unroll
    for (int d = 0; d < kAxes; ++d) { fl[d] = T(0); fr[d] = T(0); }
    if (g.act & 1) axis_term<0, T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, fl[0], fr[0]);
    if (g.act & 2) axis_term<1, T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, fl[1], fr[1]);
    if (g.act & 4) axis_term<2, T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, fl[2], fr[2]);
    if (!HOMOG) {
        if (g.act & 8) axis_term<3, T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, fl[3], fr[3]);
        if (g.act & 16) axis_term<4, T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, fl[4], fr[4]);
    }
    if (g.act & 32) axis_term<5, T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, fl[5], fr[5]);
    const T lam = HOMOG ? p.hom.lam[q[2]] : p.lam[lin];
    mf = e0 * lam;            // meanfl = E*lambda, integrate.f90:202-252
    dE = dE - mf;             // integrate.f90:261
}

// One forward-Euler step (integrate.f90:4-271) for the cells of one block.
template <typename T, bool EXACT, bool HOMOG>
__global__ void __launch_bounds__(1024) k_step(const Geo g, const StepPtrs<T> p, uint32_t nb_logical)
{
    const uint32_t b = xcd_logical_block(blockIdx.x, nb_logical);
    if (b >= nb_logical) return;
    const uint32_t lin = b * blockDim.x + threadIdx.x;
    if (lin >= g.ncells) return;
    int c[kAxes];
    uint32_t q[kAxes + 1];
    decode(g, lin, c, q);
    const T *__restrict__ E = p.Ein;
    const T e0 = E[lin];
    T dE, mf, fl[kAxes], fr[kAxes];
    cell_tendency<T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, mf, fl, fr);
    if constexpr (EXACT) {
        p.Eout[lin] = e0 + p.dt * dE;    // integrate.f90:262, separately rounded
    } else {
        p.Eout[lin] = e0 + (T)p.dt * dE;
    }
}

template <typename T, bool EXACT>
__device__ __forceinline__ T face_flux(T v, double dt, double dxf, double cfl, T Em1, T E0, T E1, T E2)
{
    if constexpr (EXACT) return sb_flux_exact(v, dt, dxf, Em1, E0, E1, E2);
    else return sb_flux_fast<T>(v, (T)cfl, Em1, E0, E1, E2);
}

template <typename T, bool EXACT>
__device__ __forceinline__ T div_term(T dE, T FR, T FL, double dxc, double rdxc)
{
    if constexpr (EXACT) return dE - (FR - FL) / dxc;
    else return dE - (FR - FL) * (T)rdxc;
}

// the fast-arithmetic div_term with the reciprocal spacing already in T
template <typename T>
__device__ __forceinline__ T div_term_r(T dE, T FR, T FL, T rdxc)
{
    return dE - (FR - FL) * rdxc;
}

#include "rt6d_column.h"
#include "rt6d_colx2.h"
#include "rt6d_colq.h"
#include "rt6d_ensemble.h"

// Tendency fields of one step (the by-products integrate leaves behind:
// dE, meanfl, flux_x..flux_m), computed from the state the step started from.
template <typename T, bool EXACT, bool HOMOG>
__global__ void __launch_bounds__(256) k_tendency(const Geo g, const StepPtrs<T> p)
{
    const uint32_t lin = blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= g.ncells) return;
    int c[kAxes];
    uint32_t q[kAxes + 1];
    decode(g, lin, c, q);
    const T *__restrict__ E = p.Ein;
    const T e0 = E[lin];
    T dE, mf, fl[kAxes], fr[kAxes];
    cell_tendency<T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, mf, fl, fr);
    p.dE[lin] = dE;
    p.meanfl[lin] = mf;
#pragma unroll
    for (int d = 0; d < kAxes; ++d) {
        if (!(g.act & (1 << d))) continue;
        const uint32_t f = lin + (uint32_t)g.inner[d] * q[d + 1];
        if (c[d] == 1) p.flux[d][f] = fl[d];
        p.flux[d][f + (uint32_t)g.inner[d]] = fr[d];
    }
}

// ---------------------------------------------------------------- diagnose

// Per-block partial sums of diagnose.f90:45-68 and :102-277.
// slot 0 eint, 1 mflint, 2 mfl_max, 3..14 bfl_xs..bfl_me
constexpr int kDiagSlots = 15;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Blocks of k_diag_cells: a grid-stride loop over the cells in at most
// kDiagParts blocks, so that the per-block partials take a bounded buffer.
constexpr uint32_t kDiagParts = 2048;

__host__ __device__ inline uint32_t diag_blocks(const Geo &g)
{
    const uint32_t nb = (g.ncells + 255u) / 256u;
    return nb < kDiagParts ? (nb ? nb : 1u) : kDiagParts;
}

// a + x into the compensated sum (sum, cmp) (Neumaier)
__device__ __forceinline__ void neumaier_add(double &sum, double &cmp, double x)
{
    const double t = sum + x;
    cmp += fabs(sum) >= fabs(x) ? (sum - t) + x : (x - t) + sum;
    sum = t;
}

template <typename T, bool EXACT, bool HOMOG>
__global__ void __launch_bounds__(256) k_diag_cells(const Geo g, const StepPtrs<T> p, const T *__restrict__ Ecur,
                                                    double *__restrict__ partial)
{
    __shared__ double red[4][kDiagSlots];
    double acc[kDiagSlots], cmp[kDiagSlots];
#pragma unroll
    for (int s = 0; s < kDiagSlots; ++s) acc[s] = cmp[s] = 0.0;
    for (uint32_t lin = blockIdx.x * blockDim.x + threadIdx.x; lin < g.ncells; lin += gridDim.x * blockDim.x) {
        double v[kDiagSlots];
#pragma unroll
        for (int s = 0; s < kDiagSlots; ++s) v[s] = 0.0;
        int c[kAxes];
        uint32_t q[kAxes + 1];
        decode(g, lin, c, q);
        const T *__restrict__ E = p.Ein;
        const T e0 = E[lin];
        T dE, mf, fl[kAxes], fr[kAxes];
        cell_tendency<T, EXACT, HOMOG>(g, p, E, lin, c, q, e0, dE, mf, fl, fr);
        double dxs[kAxes];
#pragma unroll
        for (int d = 0; d < kAxes; ++d) dxs[d] = p.ax.dx[d][c[d]];
        const double vol = dxs[0] * dxs[1] * dxs[2] * dxs[3] * dxs[4] * dxs[5];
        v[0] = (double)dE * vol;
        v[1] = (double)mf * vol;
        const double ec = (double)Ecur[lin];
        const double eps = (double)1e-12f;
        v[2] = fabs((double)mf) / (ec > eps ? ec : eps);
#pragma unroll
        for (int d = 0; d < kAxes; ++d) {
            // product of the other five spacings in axis order
            double w = 1.0;
            bool first = true;
#pragma unroll
            for (int e = 0; e < kAxes; ++e) {
                if (e == d) continue;
                w = first ? dxs[e] : w * dxs[e];
                first = false;
            }
            if (c[d] == 1) v[3 + 2 * d] = (double)fl[d] * w;
            if (c[d] == g.n[d]) v[4 + 2 * d] = (double)fr[d] * w;
        }
#pragma unroll
        for (int s = 0; s < kDiagSlots; ++s) {
            if (s == 2) acc[s] = fmax(acc[s], v[s]);
            else neumaier_add(acc[s], cmp[s], v[s]);
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < kDiagSlots; ++s) {
        double v = s == 2 ? wave_max(acc[s]) : wave_sum(acc[s] + cmp[s]);
        if (lane == 0) red[wv][s] = v;
    }
    __syncthreads();
    if (threadIdx.x < kDiagSlots) {
        const int s = threadIdx.x;
        const int nw = (blockDim.x + 63) >> 6;
        double v = red[0][s];
        for (int w = 1; w < nw; ++w) v = s == 2 ? fmax(v, red[w][s]) : v + red[w][s];
        partial[(size_t)blockIdx.x * kDiagSlots + s] = v;
    }
}

// Compensated (Neumaier) reduction of the per-block partials, one block.
static __global__ void __launch_bounds__(1024) k_diag_final(const double *__restrict__ partial, uint32_t nparts,
                                                     double *__restrict__ out)
{
    __shared__ double ssum[1024], scmp[1024];
    for (int s = 0; s < kDiagSlots; ++s) {
        double sum = 0.0, cmp = 0.0;
        for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x) {
            const double x = partial[(size_t)b * kDiagSlots + s];
            if (s == 2) { sum = fmax(sum, x); continue; }
            const double t = sum + x;
            cmp += fabs(sum) >= fabs(x) ? (sum - t) + x : (x - t) + sum;
            sum = t;
        }
        ssum[threadIdx.x] = sum;
        scmp[threadIdx.x] = cmp;
        __syncthreads();
        for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                const double a = ssum[threadIdx.x], x = ssum[threadIdx.x + w];
                if (s == 2) {
                    ssum[threadIdx.x] = fmax(a, x);
                } else {
                    const double t = a + x;
                    const double e = fabs(a) >= fabs(x) ? (a - t) + x : (x - t) + a;
                    ssum[threadIdx.x] = t;
                    scmp[threadIdx.x] += scmp[threadIdx.x + w] + e;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[s] = s == 2 ? ssum[0] : ssum[0] + scmp[0];
        __syncthreads();
    }
}

// Wavenumber-space integrals per (x,y,z) column (diagnose.f90:84-98) and the
// k/l/m boundary-flux columns diss_* (:186-277), summed sequentially in the
// reference's loop order by one thread per column.
template <typename T, bool EXACT, bool HOMOG>
__global__ void __launch_bounds__(256) k_diag_columns(const Geo g, const StepPtrs<T> p, const T *__restrict__ Ecur,
                                                      double *__restrict__ Eint, double *__restrict__ mfint,
                                                      double *__restrict__ diss)
{
    const uint32_t n3 = (uint32_t)g.n[0] * g.n[1] * g.n[2];
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n3) return;
    const int nk = g.n[3], nl = g.n[4], nm = g.n[5];
    const uint32_t s3 = (uint32_t)g.inner[3], s4 = (uint32_t)g.inner[4], s5 = (uint32_t)g.inner[5];
    const double *dkt = p.ax.dx[3], *dlt = p.ax.dx[4], *dmt = p.ax.dx[5];
    double ei = 0.0, mi = 0.0;
    for (int m = 1; m <= nm; ++m)
        for (int n = 1; n <= nl; ++n)
            for (int l = 1; l <= nk; ++l) {
                const uint32_t lin = col + s3 * (l - 1) + s4 * (n - 1) + s5 * (m - 1);
                ei = ei + (double)Ecur[lin] * dlt[n] * dmt[m] * dkt[l];
                int c[kAxes];
                uint32_t q[kAxes + 1];
                decode(g, lin, c, q);
                const T lam = HOMOG ? p.hom.lam[q[2]] : p.lam[lin];
                const double mf = (double)(p.Ein[lin] * lam);
                mi = mi + mf * dkt[l] * dlt[n] * dmt[m];
            }
    if (Eint) Eint[col] = ei;
    if (mfint) mfint[col] = mi;
    if (!diss) return;
    // boundary fluxes of k, l, m at faces 0 and n_d
    double dks = 0, dke = 0, dls = 0, dle = 0, dms = 0, dme = 0;
    const T *__restrict__ E = p.Ein;
    if (!HOMOG && (g.act & 8)) {
        for (int m = 1; m <= nm; ++m)
            for (int n = 1; n <= nl; ++n) {
                const uint32_t l0 = col + s4 * (n - 1) + s5 * (m - 1);
                const uint32_t l1 = l0 + s3 * (nk - 1);
                int c[kAxes];
                uint32_t q[kAxes + 1];
                T dE = T(0), fl, fr;
                decode(g, l0, c, q);
                axis_term<3, T, EXACT, HOMOG>(g, p, E, l0, c, q, E[l0], dE, fl, fr);
                dks = dks + (double)fl * dlt[n] * dmt[m];
                decode(g, l1, c, q);
                axis_term<3, T, EXACT, HOMOG>(g, p, E, l1, c, q, E[l1], dE, fl, fr);
                dke = dke + (double)fr * dlt[n] * dmt[m];
            }
    }
    if (!HOMOG && (g.act & 16)) {
        for (int m = 1; m <= nm; ++m)
            for (int l = 1; l <= nk; ++l) {
                const uint32_t l0 = col + s3 * (l - 1) + s5 * (m - 1);
                const uint32_t l1 = l0 + s4 * (nl - 1);
                int c[kAxes];
                uint32_t q[kAxes + 1];
                T dE = T(0), fl, fr;
                decode(g, l0, c, q);
                axis_term<4, T, EXACT, HOMOG>(g, p, E, l0, c, q, E[l0], dE, fl, fr);
                dls = dls + (double)fl * dkt[l] * dmt[m];
                decode(g, l1, c, q);
                axis_term<4, T, EXACT, HOMOG>(g, p, E, l1, c, q, E[l1], dE, fl, fr);
                dle = dle + (double)fr * dkt[l] * dmt[m];
            }
    }
    if (g.act & 32) {
        for (int n = 1; n <= nl; ++n)
            for (int l = 1; l <= nk; ++l) {
                const uint32_t l0 = col + s3 * (l - 1) + s4 * (n - 1);
                const uint32_t l1 = l0 + s5 * (nm - 1);
                int c[kAxes];
                uint32_t q[kAxes + 1];
                T dE = T(0), fl, fr;
                decode(g, l0, c, q);
                axis_term<5, T, EXACT, HOMOG>(g, p, E, l0, c, q, E[l0], dE, fl, fr);
                dms = dms + (double)fl * dkt[l] * dlt[n];
                decode(g, l1, c, q);
                axis_term<5, T, EXACT, HOMOG>(g, p, E, l1, c, q, E[l1], dE, fl, fr);
                dme = dme + (double)fr * dkt[l] * dlt[n];
            }
    }
    diss[0 * n3 + col] = dks;
    diss[1 * n3 + col] = dke;
    diss[2 * n3 + col] = dls;
    diss[3 * n3 + col] = dle;
    diss[4 * n3 + col] = dms;
    diss[5 * n3 + col] = dme;
}

// ---------------------------------------------------------------- launchers

struct LaunchCfg {
    int block;          // threads per block of the per-cell kernel
    int kernel;         // RT6D_OPT_KERNEL option (0 auto)
    int ncu;            // compute units of the device
    hipStream_t stream;
};

template <typename T, bool EXACT, int ZC, int YM = 0>
hipError_t launch_col(const Geo &g, const StepPtrs<T> &p, bool homog, hipStream_t stream)
{
    const uint32_t nxy = (uint32_t)g.n[0] * g.n[1];
    ColCfg cc;
    cc.nxyp = (int)(((nxy + 63u) / 64u) * 64u);
    cc.nzb = g.n[2] / ZC;
    cc.dxyp = make_fastdiv((uint32_t)cc.nxyp);
    cc.dzb = make_fastdiv((uint32_t)cc.nzb);
    const uint64_t nthreads = (uint64_t)cc.nxyp * cc.nzb * g.n[3] * g.n[4];
    const uint32_t nbl = (uint32_t)((nthreads + 255) / 256);
    const uint32_t nbh = ((nbl + 7u) / 8u) * 8u;
    if (homog) {
        if (is_full<true>(g))
            hipLaunchKernelGGL((k_step_col<T, EXACT, true, ZC, true, YM>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
        else
            hipLaunchKernelGGL((k_step_col<T, EXACT, true, ZC, false, YM>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
    } else {
        if (is_full<false>(g))
            hipLaunchKernelGGL((k_step_col<T, EXACT, false, ZC, true, YM>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
        else
            hipLaunchKernelGGL((k_step_col<T, EXACT, false, ZC, false, YM>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
    }
    return hipGetLastError();
}

// Byte-offset reach of the kernels that use buffer resources.  The pair
// kernels address E one m plane at a time (a resource at the plane's start),
// so only a plane has to stay below 2^31 bytes; the other column kernels
// address each array as a whole.
template <typename T>
inline bool whole_fits32(const Geo &g)
{
    return (uint64_t)g.ncells * 2u * sizeof(T) < (1ull << 32);
}
template <typename T>
inline bool mplane_fits32(const Geo &g)
{
    return (uint64_t)(uint32_t)g.inner[5] * 2u * sizeof(T) < (1ull << 32);
}

// Pair kernel (two x cells per thread): homogeneous path, nx even.
inline bool colx2_fits(const Geo &g, bool homog, int zc)
{
    return homog && g.n[0] >= 2 && g.n[0] % 2 == 0 && g.n[2] % zc == 0;
}

// The wave-local LDS variant needs the halo of two rows on each side of a
// wavefront's 64 pairs in one access: nx <= 32.
inline bool colx2wl_fits(const Geo &g, bool homog, int zc) { return colx2_fits(g, homog, zc) && g.n[0] <= 32; }

// As for the quad kernel: a grid that gives each SIMD 2.5 wavefronts of
// pairs.  On smaller grids the wave-local variant's halo loads and LDS
// exchange mostly cost more than they save (binary32, wave-local against
// plain pairs: 12^3 x 8^3 0.0182 against 0.0170 ms, 20^3 x 8^3 0.0236
// against 0.0225; 16^3 x 12^3, just below the threshold, 0.0354 against
// 0.0371; 24^3 x 8^3, above it, 0.0292 against 0.0305).
inline bool colx2_fills(const Geo &g, int zc, int ncu)
{
    const uint64_t waves_per_plane = ((uint64_t)(g.n[0] / 2) * g.n[1] + 63u) / 64u;
    const uint64_t waves = waves_per_plane * (uint64_t)(g.n[2] / zc) * g.n[3] * g.n[4];
    return 2 * waves >= 5ull * 4u * (uint64_t)(ncu > 0 ? ncu : 256);
}

// Whether the wave-local pair kernel with ZC z cells per thread hands the x
// faces between pairs from lane to lane (the XAL kernels): fast arithmetic
// with uniform x/y spacings on a full grid within 32-bit offsets, ZC <= 2,
// and wavefronts that start at row starts, which binary64 arranges by
// dealing whole rows where that takes no more wavefronts.
template <typename T>
inline bool colx2_rowal(const Geo &g, bool exact, int zc)
{
    return !exact && (g.uxy & 3) == 3 && is_full<true>(g) && zc <= 2 && whole_fits32<T>(g);
}
template <typename T>
inline bool colx2_hands_over(const Geo &g, bool exact, int zc)
{
    return colx2_rowal<T>(g, exact, zc) && colx2_config(g, zc, 2, sizeof(T) == 8).xal;
}

template <typename T, bool EXACT, int ZC, bool WL = false>
hipError_t launch_colx2(const Geo &g, const StepPtrs<T> &p, int ncu, hipStream_t stream)
{
    // the wave-local kernels in fast arithmetic take uniform x and y spacings
    // as scalars; where the wavefronts start at row starts they also hand
    // the x faces between pairs over from lane to lane.  Binary64 deals
    // whole rows to wavefronts where that takes no more of them (at 24^3 x
    // 16^3 0.2669 against 0.2725 ms); binary32 only where rows align anyway
    // (the 60-pair wavefronts of 24^3 measured 0.1485 against 0.1464 ms).
    const bool uxyw = WL && !EXACT && (g.uxy & 3) == 3 && is_full<true>(g);
    ColX2Cfg cc = colx2_config(g, ZC, 2, WL && sizeof(T) == 8 && colx2_rowal<T>(g, EXACT, ZC));
    const uint64_t nthreads = (uint64_t)cc.nxyp * cc.nzb * g.n[3] * g.n[4];
    // the binary32 sign-branch kernel gates its far outer z loads on the
    // face signs where the grid fills the device (32 wavefronts per CU)
    cc.fz = nthreads / 64u >= (uint64_t)ncu * 32u ? 1 : 0;
    const uint32_t nbl = (uint32_t)((nthreads + 255) / 256);
    const uint32_t nbh = ((nbl + 7u) / 8u) * 8u;
    // E beyond the reach of 32-bit offsets: addressed one m plane at a time
    if (!whole_fits32<T>(g)) {
        if (is_full<true>(g))
            hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, true, WL, false, true>), dim3(nbh), dim3(256), 0, stream, g,
                               p, cc, nbl);
        else
            hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, false, WL, false, true>), dim3(nbh), dim3(256), 0, stream, g,
                               p, cc, nbl);
        return hipGetLastError();
    }
    if constexpr (WL && !EXACT && ZC <= 2) {
        if (uxyw) {
            if constexpr (sizeof(T) == 8) {
                if (cc.xal)
                    hipLaunchKernelGGL((k_step_colx2_w4<T, EXACT, ZC, true, WL, true, false, true>), dim3(nbh),
                                       dim3(256), 0, stream, g, p, cc, nbl);
                else
                    hipLaunchKernelGGL((k_step_colx2_w4<T, EXACT, ZC, true, WL, true>), dim3(nbh), dim3(256), 0,
                                       stream, g, p, cc, nbl);
            } else {
                constexpr bool SB = ZC == 2;
                if (cc.xal)
                    hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, true, WL, true, false, true, SB>), dim3(nbh),
                                       dim3(256), 0, stream, g, p, cc, nbl);
                else
                    hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, true, WL, true, false, false, SB>), dim3(nbh),
                                       dim3(256), 0, stream, g, p, cc, nbl);
            }
            return hipGetLastError();
        }
    }
    if constexpr (WL && !EXACT && ZC > 2) {
        if (uxyw) {
            hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, true, WL, true>), dim3(nbh), dim3(256), 0, stream, g, p, cc,
                               nbl);
            return hipGetLastError();
        }
    }
    if (is_full<true>(g))
        hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, true, WL>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
    else
        hipLaunchKernelGGL((k_step_colx2<T, EXACT, ZC, false, WL>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
    return hipGetLastError();
}

// Quad kernel (four x cells per thread): binary32 homogeneous path, nx
// divisible by 4.
inline bool colq_fits(const Geo &g, bool homog, int zc)
{
    return homog && g.n[0] >= 4 && g.n[0] % 4 == 0 && g.n[2] % zc == 0;
}

// The quad kernel has half the pair kernel's threads.  On grids too small
// to give each SIMD (4 per compute unit) 2.5 wavefronts of quads it leaves
// the device underfilled and the pair kernel is faster (binary32, measured
// over the serving member sizes: 12^3 x 8^3 0.022 against 0.017 ms, 16^3 x
// 16^3 0.068 against 0.063 ms; 20^3 x 12^3 0.056 against 0.064 ms).
inline bool colq_fills(const Geo &g, int zc, int ncu)
{
    const uint64_t waves_per_plane = ((uint64_t)(g.n[0] / 4) * g.n[1] + 63u) / 64u;
    const uint64_t waves = waves_per_plane * (uint64_t)(g.n[2] / zc) * g.n[3] * g.n[4];
    return 2 * waves >= 5ull * 4u * (uint64_t)(ncu > 0 ? ncu : 256);
}

template <bool EXACT, int ZC>
hipError_t launch_colq(const Geo &g, const StepPtrs<float> &p, hipStream_t stream)
{
    const ColQCfg cc = colx2_config(g, ZC, 4, false);
    const uint64_t nthreads = (uint64_t)cc.nxyp * cc.nzb * g.n[3] * g.n[4];
    const uint32_t nbl = (uint32_t)((nthreads + 255) / 256);
    const uint32_t nbh = ((nbl + 7u) / 8u) * 8u;
    if (is_full<true>(g))
        hipLaunchKernelGGL((k_step_colq<EXACT, ZC, true>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
    else
        hipLaunchKernelGGL((k_step_colq<EXACT, ZC, false>), dim3(nbh), dim3(256), 0, stream, g, p, cc, nbl);
    return hipGetLastError();
}

// z cells per thread of the column kernel: two when nz is even (measured
// fastest at 24^3 x 16^3 in binary64 and binary32), else one
inline int col_zc(const Geo &g, int kernel)
{
    if (kernel == 4) return 1;
    if (kernel == 6) return g.n[2] % 4 == 0 ? 4 : (g.n[2] % 2 == 0 ? 2 : 1);
    return g.n[2] % 2 == 0 ? 2 : 1;
}

// One step.  kernel option: 0 auto, 1 per-cell kernel, 4/5/6 column kernel
// with 1/2/4 z cells per thread, 9/10/11 column kernel with LDS-staged y
// neighbours, 15-17 pair kernel, 18-20 binary32 quad kernel, 21-23 pair
// kernel with the wave-local exchange (1/2/4 z cells per thread in each
// family).  Options 2, 3, 7, 8, 12-14 and 24-26 named kernels that measured
// slower than these on every grid and were removed; they are rejected.
// The kernel option that the auto choice (option 0) resolves to.
template <typename T>
inline int auto_kernel(const Geo &g, bool homog, int ncu)
{
    const bool whole = whole_fits32<T>(g);
    if (!homog) {
        // general path: the y neighbours from LDS (measured 15% faster at
        // 24^3 x 16^3 in binary64; the homogeneous path is faster without)
        if (!whole) return 1;
        return g.n[2] % 2 == 0 ? 10 : 9;
    }
    if (!mplane_fits32<T>(g)) return 1;
    // the pair kernel with the wave-local neighbour exchange where nx <= 32
    // (24^3 x 16^3: binary64 0.300 against the plain pair kernel's 0.327 ms,
    // binary32 0.142 against 0.195 for the quad kernel; 32^3 x 16^3: 0.652
    // against 0.717 ms and 0.302 against 0.363 ms)
    if (colx2wl_fits(g, true, 1)) {
        if constexpr (sizeof(T) == 8) {
            // two z cells per thread where that fills the device, else one:
            // more wavefronts (12^3 x 8^3: 0.0159 against 0.0177 ms; 16^3 x
            // 12^3: 0.047 against 0.054 ms)
            return (g.n[2] % 2 == 0 && colx2_fills(g, 2, ncu)) ? 22 : 21;
        } else {
            // binary32 grids too small to fill the device with one z cell per
            // thread run the plain pair kernel (12^3 x 8^3: 0.0132 against
            // 0.0153 ms)
            if (colx2_fills(g, 1, ncu)) return g.n[2] % 2 == 0 ? 22 : 21;
        }
    }
    // binary32 rows too wide for the exchange: quads, two z cells per thread
    if constexpr (sizeof(T) == 4) {
        if (whole && g.n[0] > 32 && colq_fits(g, true, 2) && colq_fills(g, 2, ncu)) return 19;
    }
    // paired accesses, two z cells per thread where nz is even (binary64
    // 24^3 x 16^3: 10% faster than the column kernel)
    if (colx2_fits(g, true, 1)) return g.n[2] % 2 == 0 ? 16 : 15;
    if (!whole) return 1;
    return g.n[2] % 2 == 0 ? 5 : 4;
}

// Kernel options of kernels that were removed (slower on every grid measured).
inline bool kernel_option_removed(int k) { return k == 2 || k == 3 || k == 7 || k == 8 || (k >= 12 && k <= 14) || k >= 24; }

// z cells per thread of option k's family (1, 2 or 4 by its position in the
// family of three)
inline int family_zc(int k, int first) { return k == first ? 1 : (k == first + 1 ? 2 : 4); }

// The kernel a step with kernel option `opt` runs: the option itself where
// the grid fits that kernel, else the kernel the dispatch falls back to.
// The dispatch (launch_step_impl) and RT6D_Q_KERNEL both use it.
template <typename T>
inline int resolve_kernel(const Geo &g, bool homog, int opt, int ncu)
{
    const bool whole = whole_fits32<T>(g), mplane = mplane_fits32<T>(g);
    const int k = opt == 0 ? auto_kernel<T>(g, homog, ncu) : opt;
    if (k >= 9 && k <= 11 && whole) {
        const int zc = family_zc(k, 9);
        if (zc == 4 && g.n[2] % 4 == 0) return 11;
        if (zc >= 2 && g.n[2] % 2 == 0) return 10;
        return 9;
    }
    if (k >= 15 && k <= 17 && mplane && colx2_fits(g, homog, family_zc(k, 15))) return k;
    if (k >= 21 && k <= 23 && mplane && colx2wl_fits(g, homog, family_zc(k, 21))) return k;
    if constexpr (sizeof(T) == 4) {
        if (k >= 18 && k <= 20 && whole && colq_fits(g, homog, family_zc(k, 18))) return k;
    }
    if (k != 1 && whole) {
        const int zc = col_zc(g, k);
        return zc == 4 ? 6 : (zc == 2 ? 5 : 4);
    }
    return 1;
}

template <typename T, bool EXACT>
hipError_t launch_step_impl(const Geo &g, const StepPtrs<T> &p, bool homog, const LaunchCfg &lc)
{
    const int k = resolve_kernel<T>(g, homog, lc.kernel, lc.ncu);
    switch (k) {
    case 4: return launch_col<T, EXACT, 1>(g, p, homog, lc.stream);
    case 5: return launch_col<T, EXACT, 2>(g, p, homog, lc.stream);
    case 6: return launch_col<T, EXACT, 4>(g, p, homog, lc.stream);
    case 9: return launch_col<T, EXACT, 1, 1>(g, p, homog, lc.stream);
    case 10: return launch_col<T, EXACT, 2, 1>(g, p, homog, lc.stream);
    case 11: return launch_col<T, EXACT, 4, 1>(g, p, homog, lc.stream);
    case 15: return launch_colx2<T, EXACT, 1>(g, p, lc.ncu, lc.stream);
    case 16: return launch_colx2<T, EXACT, 2>(g, p, lc.ncu, lc.stream);
    case 17: return launch_colx2<T, EXACT, 4>(g, p, lc.ncu, lc.stream);
    case 21: return launch_colx2<T, EXACT, 1, true>(g, p, lc.ncu, lc.stream);
    case 22: return launch_colx2<T, EXACT, 2, true>(g, p, lc.ncu, lc.stream);
    case 23: return launch_colx2<T, EXACT, 4, true>(g, p, lc.ncu, lc.stream);
    default: break;
    }
    if constexpr (sizeof(T) == 4) {
        switch (k) {
        case 18: return launch_colq<EXACT, 1>(g, p, lc.stream);
        case 19: return launch_colq<EXACT, 2>(g, p, lc.stream);
        case 20: return launch_colq<EXACT, 4>(g, p, lc.stream);
        default: break;
        }
    }
    const uint32_t nbl = (uint32_t)((g.ncells + lc.block - 1) / lc.block);
    const uint32_t nbh = ((nbl + 7u) / 8u) * 8u;
    if (homog)
        hipLaunchKernelGGL((k_step<T, EXACT, true>), dim3(nbh), dim3(lc.block), 0, lc.stream, g, p, nbl);
    else
        hipLaunchKernelGGL((k_step<T, EXACT, false>), dim3(nbh), dim3(lc.block), 0, lc.stream, g, p, nbl);
    return hipGetLastError();
}

// Column-kernel configuration of an integration (two z cells per thread
// when nz is even).  Returns false when an array would exceed the 32-bit
// byte offsets of the column kernel.
template <typename T>
inline bool col_config(const Geo &g, ColCfg &cc, int &zc)
{
    const uint32_t nxy = (uint32_t)g.n[0] * g.n[1];
    zc = g.n[2] % 2 == 0 ? 2 : 1;
    cc.nxyp = (int)(((nxy + 63u) / 64u) * 64u);
    cc.nzb = g.n[2] / zc;
    cc.dxyp = make_fastdiv((uint32_t)cc.nxyp);
    cc.dzb = make_fastdiv((uint32_t)cc.nzb);
    return (uint64_t)g.ncells * 2u * sizeof(T) < (1ull << 32);
}

template <typename T>
inline uint32_t col_blocks(const Geo &g, const ColCfg &cc)
{
    const uint64_t nthreads = (uint64_t)cc.nxyp * cc.nzb * g.n[3] * g.n[4];
    return (uint32_t)((nthreads + 255) / 256);
}

// Ensemble step (rt6d_fast.hip): one launch per variant over its stretch of
// the task list (tasks + off[v], cnt[v] tasks); persist_blocks = 0 runs one
// block per task, else that many persistent blocks share the tasks through
// the work counters ctr (kEnsCounters entries: a set per step parity, each
// step clearing the other parity's set; zero both before the first step).
template <typename T>
hipError_t launch_ensemble(const EnsMember<T> *mem, const EnsTask *tasks, const uint32_t off[kEnsVariants],
                           const uint32_t cnt[kEnsVariants], uint32_t parity, uint32_t persist_blocks, uint32_t *ctr,
                           hipStream_t s);

// Superbee branch flips between a binary64 and a binary32 integration (rt6d_fast.hip).
hipError_t launch_branch_flips(const Geo &g, const StepPtrs<double> &a, bool h64, const StepPtrs<float> &b, bool h32,
                               unsigned long long *counts, hipStream_t s);

// Implemented in rt6d_exact.hip (EXACT, double only) and rt6d_fast.hip.
hipError_t launch_step_exact(const Geo &g, const StepPtrs<double> &p, bool homog, const LaunchCfg &lc);
hipError_t launch_tendency_exact(const Geo &g, const StepPtrs<double> &p, bool homog, hipStream_t s);
hipError_t launch_diag_exact(const Geo &g, const StepPtrs<double> &p, bool homog, const double *Ecur,
                             double *partial, uint32_t nparts, double *out, double *Eint, double *mfint,
                             double *diss, hipStream_t s);
template <typename T>
hipError_t launch_step_fast(const Geo &g, const StepPtrs<T> &p, bool homog, const LaunchCfg &lc);
template <typename T>
hipError_t launch_tendency_fast(const Geo &g, const StepPtrs<T> &p, bool homog, hipStream_t s);
template <typename T>
hipError_t launch_diag_fast(const Geo &g, const StepPtrs<T> &p, bool homog, const T *Ecur, double *partial,
                            uint32_t nparts, double *out, double *Eint, double *mfint, double *diss,
                            hipStream_t s);

// ------------------------------------------------ calc_parameter (exact TU)

struct ParamIn {
    // halo'd 1D grids (n+4), index f+1 for Fortran index f in [-1, n+2]
    const double *t[kAxes];
    const double *dt[kAxes];
    const double *Nsqr, *U, *V;   // halo'd 3D
    double f0, sign;
    int n[kAxes];
    int up, down;
};

// Face velocity of axis D on its face array; lambda sum; omegas.  Outputs in
// double; the API converts to the step precision.
hipError_t launch_param_velocity(int D, const ParamIn &pi, double *out, hipStream_t s);
hipError_t launch_param_lambda(const ParamIn &pi, const double *const vel[kAxes], double *lam_sum,
                               double *lam_xyz /* may be null: 3 arrays */, hipStream_t s);
hipError_t launch_param_omega(const ParamIn &pi, double *omega_i, double *omega_e, hipStream_t s);
// Homogeneous tables from the general fields and an exact check that every
// cell equals its table entry (and that kdot, ldot vanish).  *ok_dev is an
// int on the device set to 1 on any mismatch.
template <typename T>
hipError_t launch_build_homog(const Geo &g, const T *const vel[kAxes], const T *lam, T *vx3, T *vy3, T *vz,
                              T *vm, T *hlam, int *bad_dev, hipStream_t s);
// The plane records (HomTables::rec) from the homogeneous tables and the
// step coefficients ax (dt/dx for the step's dt).
template <typename T>
hipError_t launch_build_records(const Geo &g, const T *vx3, const T *vy3, const T *vz, const T *vm, const T *hlam,
                                const AxisCoef &ax, double dt, T *rec, hipStream_t s);
// max over faces 1..n_d of |v_d| into vmax_bits[6] (as ordered 64-bit ints)
hipError_t launch_vmax(const Geo &g, const double *const vel[kAxes], unsigned long long *vmax_bits,
                       hipStream_t s);
// max over cells of sum_d max(|vL|,|vR|)/dx_d(c) into *bits
hipError_t launch_cfl_max(const Geo &g, const double *const vel[kAxes], const AxisCoef &ax,
                          unsigned long long *bits, hipStream_t s);
hipError_t launch_convert_d2f(const double *in, float *out, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_bottom_forcing(const Geo &g, int kind, T *E, const double *F, const T *zdot, bool table, double dt,
                                 double dz1, hipStream_t s);
template <typename T>
hipError_t launch_init_separable(const Geo &g, const double *f, double peak, double noise, uint64_t key, T *E,
                                 hipStream_t s);
hipError_t launch_copy(const void *in, void *out, size_t bytes, hipStream_t s);
hipError_t launch_convert_f2d(const float *in, double *out, size_t n, hipStream_t s);

}  // namespace rt6d
