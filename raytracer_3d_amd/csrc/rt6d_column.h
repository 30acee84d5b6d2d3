// rt6d_column.h -- the column step kernel (included by rt6d_kernels.h,
// inside namespace rt6d).
//
// Every thread owns one (x,y) column of ZC consecutive z cells at one (k,l)
// and marches it through m = 1..nm with no LDS and no barriers:
//  * E along m sits in a register window per cell (E(m-1..m+1), and E(m+2)
//    loaded at the start of step m for its m face; E(m+3) one step ahead
//    with four z cells per thread): each E value is read from HBM once and
//    the m-face flux of step m is carried as the left face of step m+1.
//    Loading two steps ahead rather than three keeps less of E waiting in L2
//    for the neighbours' loads (binary64 pair kernel at 24^3 x 16^3: 717
//    instead of 877 MB fetched per step, at the same speed);
//  * the ZC cells share their ZC-1 inner z faces; the four outer z
//    neighbours are read one plane ahead;
//  * x and y neighbours are plain loads (L1/L2 hits: the neighbouring lanes
//    read the same lines), or (YM != 0) the y neighbours come from the
//    block's planes staged in LDS; both faces are computed per cell;
//  * the (x,y) index is padded to a multiple of 64 so that a wavefront never
//    spans two (z,k,l) planes.  Every access is then a raw buffer access with
//    a wave-uniform byte offset (scalar register) and a per-thread 32-bit
//    byte offset that is fixed for the whole march, and the horizontally
//    homogeneous coefficient tables are read through the scalar cache.
//    Arrays must stay below 4 GiB (checked by the launcher).
// Per face the arithmetic is that of k_step, so results are identical.
#pragma once

struct ColCfg {
    FastDiv dxyp;  // division by the padded (x,y) plane size
    FastDiv dzb;   // division by the number of z blocks
    int nxyp;      // nx*ny rounded up to a multiple of 64
    int nzb;       // nz / ZC
};

// Raw buffer access: address = resource base + scalar (wave-uniform) byte
// offset + per-lane 32-bit byte offset; no VALU address arithmetic, and
// accesses beyond the resource's byte size are dropped by the hardware.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mk_rsrc(const void *p, uint64_t bytes)
{
    // 0x00020000: raw (untyped, stride 0) buffer descriptor word 3 for gfx9 parts
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                             (int)(bytes > 0xffffffffull ? 0xffffffffu : (uint32_t)bytes),
                                             0x00020000);
}

// Cache policy of the step kernels' stores of the new E: non-temporal (the
// values are not read again in this step), which keeps L2 for the inputs
// (measured 1-2% faster in binary64 at 24^3 x 16^3).
constexpr int kStoreNT = 2;

template <typename T>
__device__ __forceinline__ T bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double bld<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ float bld<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <typename T>
__device__ __forceinline__ void bst(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ void bst<double>(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v2i = int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)voff, (int)soff, kStoreNT);
}
template <>
__device__ __forceinline__ void bst<float>(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, kStoreNT);
}

// Wave-uniform read of a table that no kernel writes while it runs: through
// the constant address space, so that a uniform index becomes a scalar load.
template <typename T>
__device__ __forceinline__ T sld(const T *p, uint32_t i)
{
    return ((const __attribute__((address_space(4))) T *)p)[i];
}

// Axis mask of the common case in which every axis that the path advects is
// active (x, y, z, m on the homogeneous path; all six on the general path).
// With FULL the kernel is compiled for that mask, without branches on it.
template <bool HOMOG>
constexpr unsigned full_mask() { return HOMOG ? 1u | 2u | 4u | 32u : 63u; }

template <bool HOMOG>
inline bool is_full(const Geo &g) { return (g.act & full_mask<HOMOG>()) == full_mask<HOMOG>() && (HOMOG || g.act == 63); }

// The work of logical block b (256 threads) of one integration.
// YM selects where the y neighbours come from: 0 memory; 1 the block's copy
// of its 256 (x,y) positions in LDS (one barrier per m step, two buffers
// alternating with m) where the neighbour lies inside the block, else
// memory.  With YM != 0 lanes beyond the grid stay to the end (for the
// barriers) on clamped in-bounds coordinates and store nothing.
template <typename T, bool EXACT, bool HOMOG, int ZC, bool FULL = false, int YM = 0>
__device__ __forceinline__ void col_body(const Geo &g, const StepPtrs<T> &p, const ColCfg &cc, uint32_t b)
{
    const uint32_t t = b * blockDim.x + threadIdx.x;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2], nk = g.n[3], nl = g.n[4], nm = g.n[5];
    const int nxy = nx * ny;
    // the plane part of the thread index is uniform over the wavefront
    const uint32_t pl = __builtin_amdgcn_readfirstlane(fdiv(t, cc.dxyp));
    int xy = (int)(t - pl * (uint32_t)cc.nxyp);
    const uint32_t kl = fdiv(pl, cc.dzb);
    int zb = (int)(pl - kl * (uint32_t)cc.nzb);
    int bl = (int)fdiv(kl, g.div[3]);
    int bk = (int)kl - bl * nk;
    bool live = true;
    if constexpr (YM != 0) {
        if (bl >= nl) { bl = nl - 1; bk = nk - 1; zb = cc.nzb - 1; live = false; }
        if (xy >= nxy) { xy = nxy - 1; live = false; }
    } else {
        if (bl >= nl) return;   // whole wavefronts beyond the grid
        if (xy >= nxy) return;  // padding lanes (no barriers below)
    }
    const int y0 = (int)fdiv((uint32_t)xy, g.div[0]);
    const int x0 = xy - y0 * nx;
    const int x = x0 + 1, y = y0 + 1;
    const int z0 = zb * ZC + 1, k = bk + 1, l = bl + 1;
    constexpr uint32_t S = sizeof(T);

    const size_t s5 = (size_t)g.inner[5];
    const uint32_t nzkl = (uint32_t)nz * nk * nl;
    const uint32_t Qkl = (uint32_t)bk + (uint32_t)nk * (uint32_t)bl;
    const uint32_t P0 = (uint32_t)(z0 - 1) + (uint32_t)nz * Qkl;
    const uint64_t cellsB = (uint64_t)g.ncells * S;
    const __amdgpu_buffer_rsrc_t rE = mk_rsrc(p.Ein, cellsB);
    const __amdgpu_buffer_rsrc_t rO = mk_rsrc(p.Eout, cellsB);
    const uint32_t planeB = (uint32_t)nxy * S;          // bytes per (x,y) plane
    const uint32_t mB = (uint32_t)(s5 * S);             // bytes per m step
    const uint32_t base0 = P0 * planeB;                 // plane (z0,k,l,m=1)
    const bool ax = FULL || (g.act & 1), ay = FULL || (g.act & 2), az = FULL || (g.act & 4);
    const bool ak = !HOMOG && (FULL || (g.act & 8)), al = !HOMOG && (FULL || (g.act & 16));
    const bool am = FULL || (g.act & 32);
    const double dtv = p.dt;

    // per-thread byte offsets inside an (x,y) plane, fixed for the march
    const uint32_t oc = (uint32_t)xy * S;
    uint32_t ox[4], oy[4];
    int ly[4];      // YM != 0: block-local index of the y neighbour (own index when outside the block)
    bool yin[4];
    {
        const int xs[4] = {x - 2 < 1 ? 1 : x - 2, x - 1 < 1 ? 1 : x - 1, x + 1 > nx ? nx : x + 1,
                           x + 2 > nx ? nx : x + 2};
        const int ys[4] = {y - 2 < 1 ? 1 : y - 2, y - 1 < 1 ? 1 : y - 1, y + 1 > ny ? ny : y + 1,
                           y + 2 > ny ? ny : y + 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ox[q] = (uint32_t)((xs[q] - 1) + nx * y0) * S;
            oy[q] = (uint32_t)(x0 + nx * (ys[q] - 1)) * S;
            const int li = (int)threadIdx.x + (ys[q] - y) * nx;
            yin[q] = li >= 0 && li < (int)blockDim.x;
            ly[q] = yin[q] ? li : (int)threadIdx.x;
        }
    }
    double cxl = 0, cxr = 0, dxl = 1, dxr = 1, rxx = 0, cyl = 0, cyr = 0, dyl = 1, dyr = 1, ryy = 0;
    if (ax) { cxl = p.ax.cfl[0][x - 1]; cxr = p.ax.cfl[0][x]; dxl = p.ax.dx[0][x - 1]; dxr = p.ax.dx[0][x]; rxx = p.ax.rdx[0][x]; }
    if (ay) { cyl = p.ax.cfl[1][y - 1]; cyr = p.ax.cfl[1][y]; dyl = p.ax.dx[1][y - 1]; dyr = p.ax.dx[1][y]; ryy = p.ax.rdx[1][y]; }
    double ckl = 0, ckr = 0, rkk = 0, dkl = 1, dkr = 1, cll = 0, clr = 0, rll = 0, dll = 1, dlr = 1;
    if (ak) { ckl = p.ax.cfl[3][k - 1]; ckr = p.ax.cfl[3][k]; rkk = p.ax.rdx[3][k]; dkl = p.ax.dx[3][k - 1]; dkr = p.ax.dx[3][k]; }
    if (al) { cll = p.ax.cfl[4][l - 1]; clr = p.ax.cfl[4][l]; rll = p.ax.rdx[4][l]; dll = p.ax.dx[4][l - 1]; dlr = p.ax.dx[4][l]; }

    // outer z neighbours z0-2, z0-1, z0+ZC, z0+ZC+1: plane shift and mirror flag
    int zsh[4];
    bool zmir[4];
    {
        const int zz0[4] = {z0 - 2, z0 - 1, z0 + ZC, z0 + ZC + 1};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int zz = zz0[q];
            bool mir = false;
            if (zz < 1) { if (g.down) { zz = 1 - zz; mir = true; } else zz = 1; }
            else if (zz > nz) { if (g.up) { zz = 2 * nz + 1 - zz; mir = true; } else zz = nz; }
            zsh[q] = zz - z0;
            zmir[q] = mir;
        }
    }
    int kpl[4] = {0, 0, 0, 0}, lpl[4] = {0, 0, 0, 0};   // plane shifts along k and l
    if (!HOMOG) {
        const int os[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int kk = k + os[q];
            kk = kk < 1 ? 1 : (kk > nk ? nk : kk);
            kpl[q] = (kk - k) * nz;
            int ll = l + os[q];
            ll = ll < 1 ? 1 : (ll > nl ? nl : ll);
            lpl[q] = (ll - l) * nz * nk;
        }
    }

    // byte offset of plane (z0+j, k, l, m)
    auto poff = [&](int j, int mm) -> uint32_t { return base0 + (uint32_t)j * planeB + (uint32_t)(mm - 1) * mB; };
    auto ldm = [&](int j, int mm) -> T {
        mm = mm < 1 ? 1 : (mm > nm ? nm : mm);
        return bld<T>(rE, oc, poff(j, mm));
    };
    auto ldz = [&](int q, int mm) -> T {
        const int mu = zmir[q] ? nm + 1 - mm : mm;
        return bld<T>(rE, oc, base0 + (uint32_t)(zsh[q] * (int)planeB) + (uint32_t)(mu - 1) * mB);
    };
    // own-plane look-ahead: E(m+2) loaded at the start of step m, or with
    // four z cells per thread (whose first m face comes early in the step)
    // E(m+3) one step ahead
    constexpr int LA = ZC >= 4 ? 3 : 2;
    T wm1[ZC], w0[ZC], wp1[ZC], wp2[ZC], FLm[ZC];
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
        w0[j] = ldm(j, 1);
        wp1[j] = ldm(j, 2);
        if constexpr (LA == 3) wp2[j] = ldm(j, 3);
        wm1[j] = w0[j];
        FLm[j] = T(0);
    }
    T zq[4] = {T(0), T(0), T(0), T(0)};
    if (az) {
#pragma unroll
        for (int q = 0; q < 4; ++q) zq[q] = ldz(q, 1);
    }
    // general-path coefficient fields (face arrays and lambda)
    __amdgpu_buffer_rsrc_t rV[kAxes], rL;
    if (!HOMOG) {
#pragma unroll
        for (int d = 0; d < kAxes; ++d)
            rV[d] = mk_rsrc(p.vel[d], (uint64_t)g.ncells / g.n[d] * (g.n[d] + 1) * S);
        rL = mk_rsrc(p.lam, cellsB);
    }
    const uint32_t fxo = (uint32_t)(x0 + (nx + 1) * y0) * S;        // x faces: left face x-1
    const uint32_t fyo = (uint32_t)(x0 + nx * y0) * S;              // y faces: left face y-1
    const uint32_t fxpB = (uint32_t)((nx + 1) * ny) * S, fypB = (uint32_t)(nx * (ny + 1)) * S;
    if (am) {
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const T vm0 = HOMOG ? sld(p.hom.vm, P0 + j) : bld<T>(rV[5], oc, poff(j, 1));
            // m face 0: v*E(1), the reference's formula over the zero-gradient
            // ghosts in exact arithmetic (v*(E+E)*0.5 differs from v*E where
            // the product is subnormal)
            if constexpr (EXACT) FLm[j] = sb_flux_exact(vm0, p.dt, sld(p.ax.dx[5], 0), w0[j], w0[j], w0[j], wp1[j]);
            else FLm[j] = vm0 * w0[j];
        }
    }

    for (int m = 1; m <= nm; ++m) {
        const uint32_t q3 = Qkl + (uint32_t)(m - 1) * (uint32_t)nk * nl;
        const uint32_t q20 = P0 + (uint32_t)(m - 1) * nzkl;
        T znx[4], wnx[ZC];
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            if constexpr (LA == 3) wnx[j] = ldm(j, m + 3);
            else wp2[j] = ldm(j, m + 2);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) znx[q] = (az && m < nm) ? ldz(q, m + 1) : T(0);

        T *sb = nullptr;
        constexpr int NT = 256;
        if constexpr (YM == 1) {
            __shared__ T sEy[2][ZC][NT];
            sb = &sEy[m & 1][0][0];
        }
        if constexpr (YM != 0) {
#pragma unroll
            for (int j = 0; j < ZC; ++j) sb[j * NT + threadIdx.x] = w0[j];
            __syncthreads();
        }
        T Fz[ZC + 1];
        if (az) {
            T zs[ZC + 4];
            zs[0] = zq[0];
            zs[1] = zq[1];
#pragma unroll
            for (int j = 0; j < ZC; ++j) zs[2 + j] = w0[j];
            zs[ZC + 2] = zq[2];
            zs[ZC + 3] = zq[3];
#pragma unroll
            for (int i = 0; i <= ZC; ++i) {
                const int fz = z0 - 1 + i;
                T v;
                if (HOMOG) v = sld(p.hom.vz, (uint32_t)fz + (uint32_t)(nz + 1) * q3);
                else v = bld<T>(rV[2], oc, ((uint32_t)fz + (uint32_t)(nz + 1) * q3) * planeB);
                Fz[i] = face_flux<T, EXACT>(v, dtv, sld(p.ax.dx[2], fz), sld(p.ax.cfl[2], fz), zs[i], zs[i + 1], zs[i + 2],
                                            zs[i + 3]);
            }
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const uint32_t Pm = P0 + (uint32_t)j + (uint32_t)(m - 1) * nzkl;   // plane (z0+j,k,l,m)
            const uint32_t pb = Pm * planeB;
            const uint32_t q2 = q20 + (uint32_t)j;
            const T e0 = w0[j];
            T dE = T(0);
            if (ax) {
                const T a0 = bld<T>(rE, ox[0], pb), a1 = bld<T>(rE, ox[1], pb);
                const T a3 = bld<T>(rE, ox[2], pb), a4 = bld<T>(rE, ox[3], pb);
                T vL, vR;
                if (HOMOG) {
                    const T v0 = sld(p.hom.vx3, q2), vi = sld(p.hom.vx3, p.hom.n4 + q2), vb = sld(p.hom.vx3, 2 * p.hom.n4 + q2);
                    vL = x == 1 ? v0 : vi;
                    vR = x == nx ? vb : vi;
                } else {
                    vL = bld<T>(rV[0], fxo, Pm * fxpB);
                    vR = bld<T>(rV[0], fxo + S, Pm * fxpB);
                }
                const T FL = face_flux<T, EXACT>(vL, dtv, dxl, cxl, a0, a1, e0, a3);
                const T FR = face_flux<T, EXACT>(vR, dtv, dxr, cxr, a1, e0, a3, a4);
                dE = div_term<T, EXACT>(dE, FR, FL, dxr, rxx);
            }
            if (ay) {
                T a0, a1, a3, a4;
                if constexpr (YM != 0) {
                    const T *sj = sb + j * NT;
                    a0 = sj[ly[0]];
                    a1 = sj[ly[1]];
                    a3 = sj[ly[2]];
                    a4 = sj[ly[3]];
                    if (!(yin[0] && yin[1] && yin[2] && yin[3])) {   // rows next to the block's ends
                        if (!yin[0]) a0 = bld<T>(rE, oy[0], pb);
                        if (!yin[1]) a1 = bld<T>(rE, oy[1], pb);
                        if (!yin[2]) a3 = bld<T>(rE, oy[2], pb);
                        if (!yin[3]) a4 = bld<T>(rE, oy[3], pb);
                    }
                } else {
                    a0 = bld<T>(rE, oy[0], pb);
                    a1 = bld<T>(rE, oy[1], pb);
                    a3 = bld<T>(rE, oy[2], pb);
                    a4 = bld<T>(rE, oy[3], pb);
                }
                T vL, vR;
                if (HOMOG) {
                    const T v0 = sld(p.hom.vy3, q2), vi = sld(p.hom.vy3, p.hom.n4 + q2), vb = sld(p.hom.vy3, 2 * p.hom.n4 + q2);
                    vL = y == 1 ? v0 : vi;
                    vR = y == ny ? vb : vi;
                } else {
                    vL = bld<T>(rV[1], fyo, Pm * fypB);
                    vR = bld<T>(rV[1], fyo + (uint32_t)nx * S, Pm * fypB);
                }
                const T FL = face_flux<T, EXACT>(vL, dtv, dyl, cyl, a0, a1, e0, a3);
                const T FR = face_flux<T, EXACT>(vR, dtv, dyr, cyr, a1, e0, a3, a4);
                dE = div_term<T, EXACT>(dE, FR, FL, dyr, ryy);
            }
            if (az) {
                const int zc = z0 + j;
                dE = div_term<T, EXACT>(dE, Fz[j + 1], Fz[j], sld(p.ax.dx[2], zc), sld(p.ax.rdx[2], zc));
            }
            if (ak) {
                const T km2 = bld<T>(rE, oc, pb + (uint32_t)(kpl[0] * (int)planeB));
                const T km1 = bld<T>(rE, oc, pb + (uint32_t)(kpl[1] * (int)planeB));
                const T kp1 = bld<T>(rE, oc, pb + (uint32_t)(kpl[2] * (int)planeB));
                const T kp2 = bld<T>(rE, oc, pb + (uint32_t)(kpl[3] * (int)planeB));
                // k faces: plane (z, f, l, m) of extent nk+1 along k
                const uint32_t fpl = (uint32_t)(z0 - 1 + j) + (uint32_t)nz * ((uint32_t)bk + (uint32_t)(nk + 1) * ((uint32_t)bl + (uint32_t)nl * (m - 1)));
                const T vL = bld<T>(rV[3], oc, fpl * planeB), vR = bld<T>(rV[3], oc, (fpl + nz) * planeB);
                const T FL = face_flux<T, EXACT>(vL, dtv, dkl, ckl, km2, km1, e0, kp1);
                const T FR = face_flux<T, EXACT>(vR, dtv, dkr, ckr, km1, e0, kp1, kp2);
                dE = div_term<T, EXACT>(dE, FR, FL, dkr, rkk);
            }
            if (al) {
                const T lm2 = bld<T>(rE, oc, pb + (uint32_t)(lpl[0] * (int)planeB));
                const T lm1 = bld<T>(rE, oc, pb + (uint32_t)(lpl[1] * (int)planeB));
                const T lp1 = bld<T>(rE, oc, pb + (uint32_t)(lpl[2] * (int)planeB));
                const T lp2 = bld<T>(rE, oc, pb + (uint32_t)(lpl[3] * (int)planeB));
                // l faces: plane (z, k, f, m) of extent nl+1 along l
                const uint32_t fpl = (uint32_t)(z0 - 1 + j) + (uint32_t)nz * ((uint32_t)bk + (uint32_t)nk * ((uint32_t)bl + (uint32_t)(nl + 1) * (m - 1)));
                const T vL = bld<T>(rV[4], oc, fpl * planeB), vR = bld<T>(rV[4], oc, (fpl + nz * nk) * planeB);
                const T FL = face_flux<T, EXACT>(vL, dtv, dll, cll, lm2, lm1, e0, lp1);
                const T FR = face_flux<T, EXACT>(vR, dtv, dlr, clr, lm1, e0, lp1, lp2);
                dE = div_term<T, EXACT>(dE, FR, FL, dlr, rll);
            }
            if (am) {
                const T vR = HOMOG ? sld(p.hom.vm, q2 + p.hom.inner4m) : bld<T>(rV[5], oc, pb + mB);
                const double dm = sld(p.ax.dx[5], m);
                const T FR = face_flux<T, EXACT>(vR, dtv, dm, sld(p.ax.cfl[5], m), wm1[j], e0, wp1[j], wp2[j]);
                dE = div_term<T, EXACT>(dE, FR, FLm[j], dm, sld(p.ax.rdx[5], m));
                FLm[j] = FR;
            }
            const T lam = HOMOG ? sld(p.hom.lam, q2) : bld<T>(rL, oc, pb);
            const T mf = e0 * lam;
            dE = dE - mf;
            T enew;
            if constexpr (EXACT) enew = e0 + p.dt * dE;
            else enew = e0 + (T)p.dt * dE;
            if (YM == 0 || live) bst<T>(enew, rO, oc, pb);
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            wm1[j] = w0[j];
            w0[j] = wp1[j];
            wp1[j] = wp2[j];
            if constexpr (LA == 3) wp2[j] = wnx[j];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) zq[q] = znx[q];
    }
}

template <typename T, bool EXACT, bool HOMOG, int ZC, bool FULL, int YM = 0>
__global__ void __launch_bounds__(256) k_step_col(const Geo g, const StepPtrs<T> p, const ColCfg cc, uint32_t nbl)
{
    const uint32_t b = xcd_logical_block(blockIdx.x, nbl);
    if (b >= nbl) return;   // uniform over the block
    col_body<T, EXACT, HOMOG, ZC, FULL, YM>(g, p, cc, b);
}

