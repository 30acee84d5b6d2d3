// rt6d_ensemble.h -- one launch steps many independent integrations
// (included by rt6d_kernels.h after rt6d_column.h, inside namespace rt6d).
//
// Each member is described by its geometry, its two ping-pong orders of step
// pointers and its column-kernel configuration.  The work list holds one
// (member, logical block) task per 256-thread block of every member.  A
// persistent grid of `nchunks` blocks splits the list into contiguous chunks
// (chunk ids are XCD-aware, so one XCD works on neighbouring planes of the
// same members) and runs the column-kernel body of each task in turn.
#pragma once

template <typename T>
struct EnsMember {
    Geo g;
    StepPtrs<T> p[2];   // p[0]: E[a] -> E[b]; p[1]: E[b] -> E[a]
    ColCfg cc;
    uint32_t variant;   // bit 0: horizontally homogeneous tables; bit 1: two z cells per thread;
                        // bit 2: every advected axis active
    uint32_t pad;
};

struct EnsTask {
    uint32_t member, block;
};

template <typename T>
__global__ void __launch_bounds__(256) k_ens_step(const EnsMember<T> *__restrict__ mem,
                                                  const EnsTask *__restrict__ tasks, uint32_t ntasks,
                                                  uint32_t parity, uint32_t nchunks)
{
    const uint32_t c = xcd_logical_block(blockIdx.x, nchunks);
    if (c >= nchunks) return;
    const uint32_t t0 = (uint32_t)(((uint64_t)ntasks * c) / nchunks);
    const uint32_t t1 = (uint32_t)(((uint64_t)ntasks * (c + 1)) / nchunks);
    for (uint32_t t = t0; t < t1; ++t) {
        const EnsTask tk = tasks[t];
        const EnsMember<T> &M = mem[tk.member];
        const StepPtrs<T> &P = M.p[parity & 1u];
        switch (M.variant) {
        case 0: col_body<T, false, false, 1, false>(M.g, P, M.cc, tk.block); break;
        case 1: col_body<T, false, true, 1, false>(M.g, P, M.cc, tk.block); break;
        case 2: col_body<T, false, false, 2, false>(M.g, P, M.cc, tk.block); break;
        case 3: col_body<T, false, true, 2, false>(M.g, P, M.cc, tk.block); break;
        case 4: col_body<T, false, false, 1, true>(M.g, P, M.cc, tk.block); break;
        case 5: col_body<T, false, true, 1, true>(M.g, P, M.cc, tk.block); break;
        case 6: col_body<T, false, false, 2, true>(M.g, P, M.cc, tk.block); break;
        default: col_body<T, false, true, 2, true>(M.g, P, M.cc, tk.block); break;
        }
    }
}
