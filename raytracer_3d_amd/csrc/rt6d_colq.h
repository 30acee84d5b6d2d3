// rt6d_colq.h -- the column step kernel with four x cells per thread, for
// binary32 (included by rt6d_kernels.h after rt6d_colx2.h, inside namespace
// rt6d).
//
// The pair kernel's march with each thread owning the quad of x cells
// (x .. x+3), x0 = x-1 a multiple of 4, so that the own cells, the four
// neighbouring rows and the four outer z planes each move as one 16-byte
// access; the x stencil reads the pairs (x-2, x-1) and (x+4, x+5) as 8-byte
// accesses.  Five x faces serve four cells.  The vector-memory pipeline (the
// address and data units of each CU) bounds the binary32 pair kernel, and its
// cost per access grows more slowly than the bytes moved, so moving 16 bytes
// per access halves the accesses per cell for nearly the same bytes.
// Horizontally homogeneous path, nx divisible by 4 (the quads then stay
// aligned with the rows and with 16 bytes); each cell's face arguments and
// the order in which it sums its axes are those of the column kernel, so
// results are bit for bit the same.
#pragma once

struct Quad {
    float c[4];
};

__device__ __forceinline__ Quad bld4(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v4f = float __attribute__((ext_vector_type(4)));
    const v4f f = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
    return Quad{{f.x, f.y, f.z, f.w}};
}

__device__ __forceinline__ void bst4(const Quad &v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v4i = int __attribute__((ext_vector_type(4)));
    using v4f = float __attribute__((ext_vector_type(4)));
    v4f f;
    f.x = v.c[0];
    f.y = v.c[1];
    f.z = v.c[2];
    f.w = v.c[3];
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4i, f), r, (int)voff, (int)soff, kStoreNT);
}

// Column configuration of the quad kernel: the pair kernel's fields, counted
// in quads (nxyp: (nx/4)*ny rounded up to a multiple of 64; dxh: division by
// nx/4).
using ColQCfg = ColX2Cfg;

template <bool EXACT, int ZC, bool FULL>
__device__ __forceinline__ void colq_body(const Geo &g, const StepPtrs<float> &p, const ColQCfg &cc, uint32_t b)
{
    using T = float;
    const uint32_t t = b * blockDim.x + threadIdx.x;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2], nk = g.n[3], nl = g.n[4], nm = g.n[5];
    const int nxy = nx * ny, nxq = nx >> 2;
    const uint32_t pl = __builtin_amdgcn_readfirstlane(fdiv(t, cc.dxyp));
    const int xyq = (int)(t - pl * (uint32_t)cc.nxyp);
    const uint32_t kl = fdiv(pl, cc.dzb);
    const int zb = (int)(pl - kl * (uint32_t)cc.nzb);
    const int bl = (int)fdiv(kl, g.div[3]);
    const int bk = (int)kl - bl * nk;
    if (bl >= nl) return;          // whole wavefronts beyond the grid
    const int nquads = nxq * ny;
    if (xyq >= nquads) return;   // padding lanes (no barriers below)
    const int y0 = (int)fdiv((uint32_t)xyq, cc.dxh);
    const int x0 = 4 * (xyq - y0 * nxq);   // 0-based first cell of the quad
    const int x = x0 + 1, y = y0 + 1;      // 1-based: the quad is (x .. x+3)
    const int z0 = zb * ZC + 1, k = bk + 1, l = bl + 1;
    constexpr uint32_t S = sizeof(T);
    (void)k;
    (void)l;

    const uint32_t nzkl = (uint32_t)nz * nk * nl;
    const uint32_t Qkl = (uint32_t)bk + (uint32_t)nk * (uint32_t)bl;
    const uint32_t P0 = (uint32_t)(z0 - 1) + (uint32_t)nz * Qkl;
    const uint64_t cellsB = (uint64_t)g.ncells * S;
    const __amdgpu_buffer_rsrc_t rE = mk_rsrc(p.Ein, cellsB);
    const __amdgpu_buffer_rsrc_t rO = mk_rsrc(p.Eout, cellsB);
    const uint32_t planeB = (uint32_t)nxy * S;
    const uint32_t mB = (uint32_t)g.inner[5] * S;
    const uint32_t base0 = P0 * planeB;
    const bool ax = FULL || (g.act & 1), ay = FULL || (g.act & 2), az = FULL || (g.act & 4);
    const bool am = FULL || (g.act & 32);
    const double dtv = p.dt;

    const int xy0 = x0 + nx * y0;
    const uint32_t oc = (uint32_t)xy0 * S;
    // x stencil pairs: (x-2, x-1) and (x+4, x+5); at the row ends the
    // zero-gradient ghosts repeat the quad's own edge cell
    const bool xfirst = x0 == 0, xlast = x0 + 4 >= nx;
    const uint32_t oxl = xfirst ? oc : oc - 2 * S, oxr = xlast ? oc : oc + 4 * S;
    uint32_t oy[4];
    {
        const int ys[4] = {y - 2 < 1 ? 1 : y - 2, y - 1 < 1 ? 1 : y - 1, y + 1 > ny ? ny : y + 1,
                           y + 2 > ny ? ny : y + 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) oy[q] = (uint32_t)(x0 + nx * (ys[q] - 1)) * S;
    }
    // x faces x-1 .. x+3 and cells x .. x+3; y faces y-1, y of row y
    double cxf[5], dxf[5], rx[4];
#pragma unroll
    for (int i = 0; i < 5; ++i) { cxf[i] = 0; dxf[i] = 1; }
#pragma unroll
    for (int i = 0; i < 4; ++i) rx[i] = 0;
    if (ax) {
#pragma unroll
        for (int i = 0; i < 5; ++i) { cxf[i] = p.ax.cfl[0][x - 1 + i]; dxf[i] = p.ax.dx[0][x - 1 + i]; }
#pragma unroll
        for (int i = 0; i < 4; ++i) rx[i] = p.ax.rdx[0][x + i];
    }
    double cyl = 0, cyr = 0, dyl = 1, dyr = 1, ryy = 0;
    if (ay) { cyl = p.ax.cfl[1][y - 1]; cyr = p.ax.cfl[1][y]; dyl = p.ax.dx[1][y - 1]; dyr = p.ax.dx[1][y]; ryy = p.ax.rdx[1][y]; }

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
    auto poff = [&](int j, int mm) -> uint32_t { return base0 + (uint32_t)j * planeB + (uint32_t)(mm - 1) * mB; };
    auto ldm = [&](int j, int mm) -> Quad {
        mm = mm < 1 ? 1 : (mm > nm ? nm : mm);
        return bld4(rE, oc, poff(j, mm));
    };
    auto ldz = [&](int q, int mm) -> Quad {
        const int mu = zmir[q] ? nm + 1 - mm : mm;
        return bld4(rE, oc, base0 + (uint32_t)(zsh[q] * (int)planeB) + (uint32_t)(mu - 1) * mB);
    };
    constexpr int LA = ZC >= 4 ? 3 : 2;   // own-plane look-ahead, as in the column kernel
    Quad wm1[ZC], w0[ZC], wp1[ZC], wp2[ZC], FLm[ZC];
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
        w0[j] = ldm(j, 1);
        wp1[j] = ldm(j, 2);
        if constexpr (LA == 3) wp2[j] = ldm(j, 3);
        wm1[j] = w0[j];
        FLm[j] = Quad{{T(0), T(0), T(0), T(0)}};
    }
    Quad zq[4];
    if (am) {
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const T vm0 = sld(p.hom.vm, P0 + j);
#pragma unroll
            for (int i = 0; i < 4; ++i) FLm[j].c[i] = vm0 * w0[j].c[i];   // m face 0: v*E(1)
        }
    }

    for (int m = 1; m <= nm; ++m) {
        const uint32_t q3 = Qkl + (uint32_t)(m - 1) * (uint32_t)nk * nl;
        const uint32_t q20 = P0 + (uint32_t)(m - 1) * nzkl;
        Quad wnx[ZC];
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            if constexpr (LA == 3) wnx[j] = ldm(j, m + 3);
            else wp2[j] = ldm(j, m + 2);   // E(m+2), for this step's m face
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) zq[q] = az ? ldz(q, m) : Quad{{T(0), T(0), T(0), T(0)}};

        T Fz[4][ZC + 1];
        if (az) {
#pragma unroll
            for (int i = 0; i <= ZC; ++i) {
                const int fz = z0 - 1 + i;
                const T v = sld(p.hom.vz, (uint32_t)fz + (uint32_t)(nz + 1) * q3);
                const double dzf = sld(p.ax.dx[2], fz), czf = sld(p.ax.cfl[2], fz);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    T zs[ZC + 4];
                    zs[0] = zq[0].c[c];
                    zs[1] = zq[1].c[c];
#pragma unroll
                    for (int j = 0; j < ZC; ++j) zs[2 + j] = w0[j].c[c];
                    zs[ZC + 2] = zq[2].c[c];
                    zs[ZC + 3] = zq[3].c[c];
                    Fz[c][i] = face_flux<T, EXACT>(v, dtv, dzf, czf, zs[i], zs[i + 1], zs[i + 2], zs[i + 3]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const uint32_t Pm = P0 + (uint32_t)j + (uint32_t)(m - 1) * nzkl;
            // raised wave priority while this cell quad's neighbour loads
            // issue, as in the pair kernel
            __builtin_amdgcn_s_setprio(1);
            const uint32_t pb = Pm * planeB;
            const uint32_t q2 = q20 + (uint32_t)j;
            const Quad e = w0[j];
            T dE[4] = {T(0), T(0), T(0), T(0)};
            if (ax) {
                Pair<T> L = bld2<T>(rE, oxl, pb), R = bld2<T>(rE, oxr, pb);
                if (xfirst) L = Pair<T>{e.c[0], e.c[0]};
                if (xlast) R = Pair<T>{e.c[3], e.c[3]};
                const T v0 = sld(p.hom.vx3, q2), vi = sld(p.hom.vx3, p.hom.n4 + q2), vb = sld(p.hom.vx3, 2 * p.hom.n4 + q2);
                const T vL = x == 1 ? v0 : vi, vR = x + 3 == nx ? vb : vi;
                // the stencil of x faces x-1 .. x+3: cells x-2 .. x+5
                const T s[8] = {L.a, L.b, e.c[0], e.c[1], e.c[2], e.c[3], R.a, R.b};
                T F[5];
#pragma unroll
                for (int i = 0; i < 5; ++i)
                    F[i] = face_flux<T, EXACT>(i == 0 ? vL : (i == 4 ? vR : vi), dtv, dxf[i], cxf[i], s[i], s[i + 1],
                                               s[i + 2], s[i + 3]);
#pragma unroll
                for (int c = 0; c < 4; ++c) dE[c] = div_term<T, EXACT>(dE[c], F[c + 1], F[c], dxf[c + 1], rx[c]);
            }
            if (ay) {
                const Quad a0 = bld4(rE, oy[0], pb), a1 = bld4(rE, oy[1], pb);
                const Quad a3 = bld4(rE, oy[2], pb), a4 = bld4(rE, oy[3], pb);
                __builtin_amdgcn_s_setprio(0);
                const T v0 = sld(p.hom.vy3, q2), vi = sld(p.hom.vy3, p.hom.n4 + q2), vb = sld(p.hom.vy3, 2 * p.hom.n4 + q2);
                const T vL = y == 1 ? v0 : vi, vR = y == ny ? vb : vi;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T FL = face_flux<T, EXACT>(vL, dtv, dyl, cyl, a0.c[c], a1.c[c], e.c[c], a3.c[c]);
                    const T FR = face_flux<T, EXACT>(vR, dtv, dyr, cyr, a1.c[c], e.c[c], a3.c[c], a4.c[c]);
                    dE[c] = div_term<T, EXACT>(dE[c], FR, FL, dyr, ryy);
                }
            }
            if (!ay) __builtin_amdgcn_s_setprio(0);
            if (az) {
                const int zc = z0 + j;
                const double dzc = sld(p.ax.dx[2], zc), rzc = sld(p.ax.rdx[2], zc);
#pragma unroll
                for (int c = 0; c < 4; ++c) dE[c] = div_term<T, EXACT>(dE[c], Fz[c][j + 1], Fz[c][j], dzc, rzc);
            }
            if (am) {
                const T vR = sld(p.hom.vm, q2 + p.hom.inner4m);
                const double dm = sld(p.ax.dx[5], m), cm = sld(p.ax.cfl[5], m), rm = sld(p.ax.rdx[5], m);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T FR = face_flux<T, EXACT>(vR, dtv, dm, cm, wm1[j].c[c], e.c[c], wp1[j].c[c], wp2[j].c[c]);
                    dE[c] = div_term<T, EXACT>(dE[c], FR, FLm[j].c[c], dm, rm);
                    FLm[j].c[c] = FR;
                }
            }
            const T lam = sld(p.hom.lam, q2);
            Quad en;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T d = dE[c] - e.c[c] * lam;
                en.c[c] = e.c[c] + (T)p.dt * d;
            }
            bst4(en, rO, oc, pb);
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            wm1[j] = w0[j];
            w0[j] = wp1[j];
            wp1[j] = wp2[j];
            if constexpr (LA == 3) wp2[j] = wnx[j];
        }
    }
}

template <bool EXACT, int ZC, bool FULL>
__global__ void __launch_bounds__(256) k_step_colq(const Geo g, const StepPtrs<float> p, const ColQCfg cc, uint32_t nbl)
{
    const uint32_t b = xcd_logical_block(blockIdx.x, nbl);
    if (b >= nbl) return;
    colq_body<EXACT, ZC, FULL>(g, p, cc, b);
}
