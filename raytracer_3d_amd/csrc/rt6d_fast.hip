// rt6d_fast.hip -- throughput variants of the step, tendency and diagnose
// kernels: multiplications by precomputed reciprocals (1/dxt, dt/dxt), the
// division-free superbee form and FMA contraction.  Instantiated for binary64
// (the plain step) and binary32 (the FP32 form: E, fluxes and their sums in
// binary32; the dt and grid coefficients arrive in binary64).
#include "rt6d_kernels.h"

namespace rt6d {

static inline uint32_t nblocks(uint64_t n, int b) { return (uint32_t)((n + b - 1) / b); }

template <typename T>
hipError_t launch_step_fast(const Geo &g, const StepPtrs<T> &p, bool homog, const LaunchCfg &lc)
{
    return launch_step_impl<T, false>(g, p, homog, lc);
}

template <typename T>
hipError_t launch_tendency_fast(const Geo &g, const StepPtrs<T> &p, bool homog, hipStream_t s)
{
    const uint32_t nb = nblocks(g.ncells, 256);
    if (homog) hipLaunchKernelGGL((k_tendency<T, false, true>), dim3(nb), dim3(256), 0, s, g, p);
    else hipLaunchKernelGGL((k_tendency<T, false, false>), dim3(nb), dim3(256), 0, s, g, p);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_diag_fast(const Geo &g, const StepPtrs<T> &p, bool homog, const T *Ecur, double *partial,
                            uint32_t nparts, double *out, double *Eint, double *mfint, double *diss, hipStream_t s)
{
    const uint32_t nb = diag_blocks(g);
    if (nb > nparts) return hipErrorInvalidValue;
    if (homog)
        hipLaunchKernelGGL((k_diag_cells<T, false, true>), dim3(nb), dim3(256), 0, s, g, p, Ecur, partial);
    else
        hipLaunchKernelGGL((k_diag_cells<T, false, false>), dim3(nb), dim3(256), 0, s, g, p, Ecur, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(1024), 0, s, partial, nb, out);
    if (Eint || mfint || diss) {
        const uint32_t n3 = (uint32_t)g.n[0] * g.n[1] * g.n[2];
        if (homog)
            hipLaunchKernelGGL((k_diag_columns<T, false, true>), dim3(nblocks(n3, 256)), dim3(256), 0, s, g, p, Ecur,
                               Eint, mfint, diss);
        else
            hipLaunchKernelGGL((k_diag_columns<T, false, false>), dim3(nblocks(n3, 256)), dim3(256), 0, s, g, p,
                               Ecur, Eint, mfint, diss);
    }
    return hipGetLastError();
}

template <typename T, bool HOMOG, int ZC, bool FULL, bool X2 = false, bool Q = false, bool W = false,
          bool UXY = false, bool XAL = false>
static void ens_launch_variant(const EnsMember<T> *mem, const EnsTask *tasks, uint32_t n, uint32_t parity,
                               uint32_t persist_blocks, uint32_t *ctr, hipStream_t s)
{
    if (persist_blocks) {
        uint32_t nb = persist_blocks < n ? persist_blocks : n;
        nb = ((nb + 7u) / 8u) * 8u;
        hipLaunchKernelGGL((k_ens_step<T, HOMOG, ZC, FULL, X2, Q, W, UXY, XAL, true>), dim3(nb), dim3(256), 0, s, mem, tasks,
                           n, parity, ctr);
    } else {
        const uint32_t nb = ((n + 7u) / 8u) * 8u;
        hipLaunchKernelGGL((k_ens_step<T, HOMOG, ZC, FULL, X2, Q, W, UXY, XAL, false>), dim3(nb), dim3(256), 0, s, mem,
                           tasks, n, parity, ctr);
    }
}

template <typename T>
hipError_t launch_ensemble(const EnsMember<T> *mem, const EnsTask *tasks, const uint32_t off[kEnsVariants],
                           const uint32_t cnt[kEnsVariants], uint32_t parity, uint32_t persist_blocks, uint32_t *ctr,
                           hipStream_t s)
{
    for (int v = 0; v < kEnsVariants; ++v) {
        if (!cnt[v]) continue;
        const EnsTask *tv = tasks + off[v];
        uint32_t *cv = ctr + 8 * v;
        switch (v) {
        case 0: ens_launch_variant<T, false, 1, false>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 1: ens_launch_variant<T, true, 1, false>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 2: ens_launch_variant<T, false, 2, false>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 3: ens_launch_variant<T, true, 2, false>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 4: ens_launch_variant<T, false, 1, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 5: ens_launch_variant<T, true, 1, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 6: ens_launch_variant<T, false, 2, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 7: ens_launch_variant<T, true, 2, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        // pair kernel: homogeneous members only (bit 0 set)
        case 9: ens_launch_variant<T, true, 1, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 11: ens_launch_variant<T, true, 2, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 13: ens_launch_variant<T, true, 1, true, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 15: ens_launch_variant<T, true, 2, true, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        // quad kernel: binary32 homogeneous members with two z cells per thread (bits 0 and 1 set)
        case 19:
            if constexpr (sizeof(T) == 4)
                ens_launch_variant<T, true, 2, false, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s);
            break;
        case 23:
            if constexpr (sizeof(T) == 4)
                ens_launch_variant<T, true, 2, true, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s);
            break;
        // wave-local pair kernel: homogeneous members, one or two z cells per
        // thread, with uniform x and y spacings as scalars (bit 6)
        case 33: ens_launch_variant<T, true, 1, false, false, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 35: ens_launch_variant<T, true, 2, false, false, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 37: ens_launch_variant<T, true, 1, true, false, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 39: ens_launch_variant<T, true, 2, true, false, false, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s); break;
        case 101:
            ens_launch_variant<T, true, 1, true, false, false, true, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s);
            break;
        case 103:
            ens_launch_variant<T, true, 2, true, false, false, true, true>(mem, tv, cnt[v], parity, persist_blocks, cv, s);
            break;
        // the same in binary64 with whole rows per wavefront and the x faces passed between lanes (bit 7)
        case 229:
            if constexpr (sizeof(T) == 8)
                ens_launch_variant<T, true, 1, true, false, false, true, true, true>(mem, tv, cnt[v], parity,
                                                                                    persist_blocks, cv, s);
            break;
        case 231:
            if constexpr (sizeof(T) == 8)
                ens_launch_variant<T, true, 2, true, false, false, true, true, true>(mem, tv, cnt[v], parity,
                                                                                    persist_blocks, cv, s);
            break;
        default: break;   // not produced by the host
        }
    }
    return hipGetLastError();
}
template hipError_t launch_ensemble<double>(const EnsMember<double> *, const EnsTask *, const uint32_t *,
                                            const uint32_t *, uint32_t, uint32_t, uint32_t *, hipStream_t);
template hipError_t launch_ensemble<float>(const EnsMember<float> *, const EnsTask *, const uint32_t *,
                                           const uint32_t *, uint32_t, uint32_t, uint32_t *, hipStream_t);

// Superbee branch of a face, from |R| and the upwind difference carried to
// the sign of R (r = su/|R|): 0 phi = 0 (r <= 0), 1 phi = 2r, 2 phi = 1,
// 3 phi = r, 4 phi = 2.  Ties sit where neighbouring branches agree.
template <typename T>
__device__ __forceinline__ int sb_branch(T v, T Em1, T E0, T E1, T E2)
{
    const T R = E1 - E0;
    const T Ru = v > T(0) ? (E0 - Em1) : (E2 - E1);
    const T a = fabs(R), s = R < T(0) ? -Ru : Ru;
    if (!(s > T(0))) return 0;
    if (s + s <= a) return 1;
    if (s <= a) return 2;
    if (s <= a + a) return 3;
    return 4;
}

template <int D, typename T, bool HOMOG>
__device__ __forceinline__ int upper_face_branch(const Geo &g, const StepPtrs<T> &p, uint32_t lin, const int c[kAxes],
                                                 const uint32_t q[kAxes + 1])
{
    const T *E = p.Ein;
    const T e0 = E[lin];
    const T em1 = E[(int64_t)lin + nbr_off<D>(g, c, -1)];
    const T ep1 = E[(int64_t)lin + nbr_off<D>(g, c, 1)];
    const T ep2 = E[(int64_t)lin + nbr_off<D>(g, c, 2)];
    T vL, vR;
    load_vel<D, T, HOMOG>(g, p, lin, c, q, vL, vR);
    return sb_branch<T>(vR, em1, e0, ep1, ep2);
}

// Faces whose superbee branch differs between a binary64 and a binary32
// integration of the same problem (each in its own arithmetic, each on its
// own state).  Counts per axis: [0..5] faces compared, [6..11] flips; the
// upper face of every cell (the lower boundary face has R = 0 in both).
template <bool H64, bool H32>
__global__ void __launch_bounds__(256) k_branch_flips(const Geo g, const StepPtrs<double> a, const StepPtrs<float> b,
                                                      unsigned long long *counts)
{
    __shared__ unsigned long long red[12];
    if (threadIdx.x < 12) red[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long nf[kAxes] = {0, 0, 0, 0, 0, 0}, fl[kAxes] = {0, 0, 0, 0, 0, 0};
    for (uint32_t lin = blockIdx.x * blockDim.x + threadIdx.x; lin < g.ncells; lin += gridDim.x * blockDim.x) {
        int c[kAxes];
        uint32_t q[kAxes + 1];
        decode(g, lin, c, q);
#define RT6D_FLIP_AXIS(D)                                                                              \
        if (g.act & (1 << D)) {                                                                         \
            const bool skip64 = H64 && (D == 3 || D == 4), skip32 = H32 && (D == 3 || D == 4);          \
            const int b64 = skip64 ? 0 : upper_face_branch<D, double, H64>(g, a, lin, c, q);            \
            const int b32 = skip32 ? 0 : upper_face_branch<D, float, H32>(g, b, lin, c, q);             \
            nf[D] += 1;                                                                                 \
            fl[D] += b64 != b32 ? 1 : 0;                                                                \
        }
        RT6D_FLIP_AXIS(0) RT6D_FLIP_AXIS(1) RT6D_FLIP_AXIS(2) RT6D_FLIP_AXIS(3) RT6D_FLIP_AXIS(4) RT6D_FLIP_AXIS(5)
#undef RT6D_FLIP_AXIS
    }
#pragma unroll
    for (int d = 0; d < kAxes; ++d) {
        atomicAdd(&red[d], nf[d]);
        atomicAdd(&red[6 + d], fl[d]);
    }
    __syncthreads();
    if (threadIdx.x < 12) atomicAdd(&counts[threadIdx.x], red[threadIdx.x]);
}

hipError_t launch_branch_flips(const Geo &g, const StepPtrs<double> &a, bool h64, const StepPtrs<float> &b, bool h32,
                               unsigned long long *counts, hipStream_t s)
{
    uint32_t nb = (uint32_t)((g.ncells + 255) / 256);
    nb = nb > 2048 ? 2048 : (nb ? nb : 1);
    if (h64 && h32) hipLaunchKernelGGL((k_branch_flips<true, true>), dim3(nb), dim3(256), 0, s, g, a, b, counts);
    else if (h64) hipLaunchKernelGGL((k_branch_flips<true, false>), dim3(nb), dim3(256), 0, s, g, a, b, counts);
    else if (h32) hipLaunchKernelGGL((k_branch_flips<false, true>), dim3(nb), dim3(256), 0, s, g, a, b, counts);
    else hipLaunchKernelGGL((k_branch_flips<false, false>), dim3(nb), dim3(256), 0, s, g, a, b, counts);
    return hipGetLastError();
}

template hipError_t launch_step_fast<double>(const Geo &, const StepPtrs<double> &, bool, const LaunchCfg &);
template hipError_t launch_step_fast<float>(const Geo &, const StepPtrs<float> &, bool, const LaunchCfg &);
template hipError_t launch_tendency_fast<double>(const Geo &, const StepPtrs<double> &, bool, hipStream_t);
template hipError_t launch_tendency_fast<float>(const Geo &, const StepPtrs<float> &, bool, hipStream_t);
template hipError_t launch_diag_fast<double>(const Geo &, const StepPtrs<double> &, bool, const double *, double *,
                                             uint32_t, double *, double *, double *, double *, hipStream_t);
template hipError_t launch_diag_fast<float>(const Geo &, const StepPtrs<float> &, bool, const float *, double *,
                                            uint32_t, double *, double *, double *, double *, hipStream_t);

}  // namespace rt6d
