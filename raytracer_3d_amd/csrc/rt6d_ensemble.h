// rt6d_ensemble.h -- one launch per kernel variant steps many independent
// integrations (included by rt6d_kernels.h after rt6d_column.h, inside
// namespace rt6d).
//
// Each member is described by its geometry, its two ping-pong orders of step
// pointers and its column-kernel configuration.  The work list holds one
// (member, logical block) task per 256-thread block of every member, grouped
// by kernel variant (homogeneous or general coefficients, one or two z cells
// per thread, all advected axes active or not), so that every launch is
// compiled for exactly one variant and keeps that variant's register budget.
// Two ways to run a variant's tasks:
//  * dispatch: one hardware block per task, in list order; block ids are
//    XCD-aware, so each XCD works on a contiguous stretch of the list (the
//    neighbouring blocks of one member share that XCD's L2);
//  * persistent: a grid of blocks_per_cu blocks per compute unit; the blocks
//    of each XCD take the tasks of that XCD's eighth of the list from a work
//    counter until it is exhausted (every block reaches the end).  The work
//    counters come in two sets, one per step parity: a step takes its tasks
//    from its parity's set and its first block clears the other set, which
//    the previous step used and the next step will use (steps of one
//    ensemble run in order on its stream).  The host clears both sets when
//    the work list is rebuilt, so a step is one launch per variant with no
//    clearing kernel or memset between launches.
#pragma once

template <typename T>
struct EnsMember {
    Geo g;
    StepPtrs<T> p[2];   // p[0]: E[a] -> E[b]; p[1]: E[b] -> E[a]
    ColCfg cc;
    ColX2Cfg cx;        // pair- or quad-kernel configuration (variants with bit 3 or 4)
    uint32_t variant;   // bit 0: homogeneous tables; bit 1: two z cells per thread; bit 2: all axes
                        // active; bit 3: two x cells per thread (homogeneous, nx even);
                        // bit 4: four x cells per thread (binary32, homogeneous, nx
                        // divisible by 4, two z cells per thread); bit 5: two x cells
                        // per thread with the wave-local neighbour exchange (homogeneous,
                        // nx even and <= 32); bit 6: with bit 5 and all axes active,
                        // uniform x and y spacings taken as scalars; bit 7: with bit 6
                        // (binary64), whole rows per wavefront and the x faces between
                        // pairs passed from lane to lane
    uint32_t pad;
};

struct EnsTask {
    uint32_t member, block;
};

constexpr int kEnsVariants = 256;
constexpr int kEnsCounterSet = kEnsVariants * 8;   // one work counter per variant and XCD
constexpr int kEnsCounters = 2 * kEnsCounterSet;   // a set per step parity

template <typename T, bool HOMOG, int ZC, bool FULL, bool X2, bool Q, bool W, bool UXY, bool XAL>
__device__ __forceinline__ void ens_task(const EnsMember<T> &M, uint32_t parity, uint32_t block)
{
    if constexpr (W) {
        colx2_body<T, false, ZC, FULL, true, UXY, false, XAL>(M.g, M.p[parity & 1u], M.cx, block);
    } else if constexpr (Q) {
        if constexpr (sizeof(T) == 4) colq_body<false, ZC, FULL>(M.g, M.p[parity & 1u], M.cx, block);
    } else if constexpr (X2) {
        colx2_body<T, false, ZC, FULL>(M.g, M.p[parity & 1u], M.cx, block);
    } else {
        col_body<T, false, HOMOG, ZC, FULL>(M.g, M.p[parity & 1u], M.cc, block);
    }
}

// Held to 128 VGPRs (four wavefronts per SIMD) where the binary64
// uniform-spacing wave-local body renames its m window, as the single-run
// kernel is (launched one block per task: the persistent form would spill);
// one wavefront per SIMD is the floor a 256-thread block sets anyway.
template <typename T, bool HOMOG, int ZC, bool FULL, bool X2, bool Q, bool W, bool UXY, bool XAL, bool PERSIST>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && W && UXY && ZC <= 2 && !PERSIST ? 4 : 1)))
k_ens_step(const EnsMember<T> *__restrict__ mem, const EnsTask *__restrict__ tasks, uint32_t ntasks,
           uint32_t parity, uint32_t *ctr)
{
    if constexpr (!PERSIST) {
        const uint32_t t = xcd_logical_block(blockIdx.x, ntasks);
        if (t >= ntasks) return;   // uniform over the block
        const EnsTask tk = tasks[t];
        ens_task<T, HOMOG, ZC, FULL, X2, Q, W, UXY, XAL>(mem[tk.member], parity, tk.block);
    } else {
        __shared__ uint32_t s_task;
        // ctr: this variant's counters in the set of parity 0
        uint32_t *const mine = ctr + (parity & 1u) * kEnsCounterSet;
        if (blockIdx.x == 0 && threadIdx.x < 8) ctr[((parity & 1u) ^ 1u) * kEnsCounterSet + threadIdx.x] = 0u;
        const uint32_t xcd = blockIdx.x & 7u;
        const uint32_t per = (ntasks + 7u) / 8u;
        const uint32_t t0 = xcd * per;
        const uint32_t t1 = t0 + per < ntasks ? t0 + per : ntasks;
        for (;;) {
            if (threadIdx.x == 0) s_task = t0 + atomicAdd(&mine[xcd], 1u);
            __syncthreads();
            const uint32_t t = s_task;
            __syncthreads();   // every thread has read s_task before thread 0 writes it again
            if (t >= t1) break;
            const EnsTask tk = tasks[t];
            ens_task<T, HOMOG, ZC, FULL, X2, Q, W, UXY, XAL>(mem[tk.member], parity, tk.block);
        }
    }
}
