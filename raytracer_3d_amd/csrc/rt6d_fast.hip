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
    const uint32_t nb = nblocks(g.ncells, 256);
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

template <typename T>
hipError_t launch_ensemble(const EnsMember<T> *mem, const EnsTask *tasks, uint32_t ntasks, uint32_t parity,
                           uint32_t nchunks, hipStream_t s)
{
    if (ntasks == 0) return hipSuccess;
    const uint32_t nbh = ((nchunks + 7u) / 8u) * 8u;
    hipLaunchKernelGGL(k_ens_step<T>, dim3(nbh), dim3(256), 0, s, mem, tasks, ntasks, parity, nchunks);
    return hipGetLastError();
}
template hipError_t launch_ensemble<double>(const EnsMember<double> *, const EnsTask *, uint32_t, uint32_t, uint32_t,
                                            hipStream_t);
template hipError_t launch_ensemble<float>(const EnsMember<float> *, const EnsTask *, uint32_t, uint32_t, uint32_t,
                                           hipStream_t);

template hipError_t launch_step_fast<double>(const Geo &, const StepPtrs<double> &, bool, const LaunchCfg &);
template hipError_t launch_step_fast<float>(const Geo &, const StepPtrs<float> &, bool, const LaunchCfg &);
template hipError_t launch_tendency_fast<double>(const Geo &, const StepPtrs<double> &, bool, hipStream_t);
template hipError_t launch_tendency_fast<float>(const Geo &, const StepPtrs<float> &, bool, hipStream_t);
template hipError_t launch_diag_fast<double>(const Geo &, const StepPtrs<double> &, bool, const double *, double *,
                                             uint32_t, double *, double *, double *, double *, hipStream_t);
template hipError_t launch_diag_fast<float>(const Geo &, const StepPtrs<float> &, bool, const float *, double *,
                                            uint32_t, double *, double *, double *, double *, hipStream_t);

}  // namespace rt6d
