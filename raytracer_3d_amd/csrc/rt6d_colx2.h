// rt6d_colx2.h -- the column step kernel with two x cells per thread
// (included by rt6d_kernels.h after rt6d_column.h, inside namespace rt6d).
//
// The column kernel's march (register window along m, carried m-face flux,
// ZC z cells per thread sharing their inner z faces; the outer z planes are
// read in the step that uses them, which keeps the binary64 register count
// at four wavefronts per SIMD) with each thread owning
// the pair of x cells (x, x+1), x odd, so that every access moves a pair: the
// own pair, the pairs (x-2, x-1) and (x+2, x+3) for the x stencil, the pairs
// of the four neighbouring rows for y and of the four outer z planes.  The
// pair shares its inner x face, so three x faces serve two cells.  Halving
// the vector-memory instructions per cell is the point: the column kernel is
// bounded by the vector-memory pipeline, not by HBM or the VALU.
// Horizontally homogeneous path only, nx even (the pairs then stay aligned
// with the rows); the face arithmetic and the order in which each cell sums
// its axes are those of the column kernel, so results are bit for bit the
// same.
#pragma once

template <typename T>
struct Pair {
    T a, b;
};

template <typename T>
__device__ __forceinline__ Pair<T> bld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ Pair<double> bld2<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v4i = int __attribute__((ext_vector_type(4)));
    const v4i v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    using v2d = double __attribute__((ext_vector_type(2)));
    const v2d d = __builtin_bit_cast(v2d, v);
    return Pair<double>{d.x, d.y};
}
template <>
__device__ __forceinline__ Pair<float> bld2<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v2i = int __attribute__((ext_vector_type(2)));
    const v2i v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    using v2f = float __attribute__((ext_vector_type(2)));
    const v2f f = __builtin_bit_cast(v2f, v);
    return Pair<float>{f.x, f.y};
}
template <typename T>
__device__ __forceinline__ void bst2(Pair<T> v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ void bst2<double>(Pair<double> v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v4i = int __attribute__((ext_vector_type(4)));
    using v2d = double __attribute__((ext_vector_type(2)));
    v2d d;
    d.x = v.a;
    d.y = v.b;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4i, d), r, (int)voff, (int)soff, kStoreNT);
}
template <>
__device__ __forceinline__ void bst2<float>(Pair<float> v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    using v2i = int __attribute__((ext_vector_type(2)));
    using v2f = float __attribute__((ext_vector_type(2)));
    v2f f;
    f.x = v.a;
    f.y = v.b;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, f), r, (int)voff, (int)soff, kStoreNT);
}

// Column configuration of the pair kernel: the padded plane counts pairs.
// A plane's pairs are dealt to wavefronts wn at a time: 64, or (where asked
// for and where that takes no more wavefronts) the whole rows that fit in 64
// lanes, so that every wavefront starts at a row start (xal; also where 64 is
// a multiple of nx/2).
struct ColX2Cfg {
    FastDiv dxyp;  // division by the padded (pair, y) plane size
    FastDiv dzb;   // division by the number of z blocks
    FastDiv dxh;   // division by nx/2
    int nxyp;      // 64 lanes per wavefront of the plane: 64 * ceil((nx/2)*ny / wn)
    int nzb;       // nz / ZC
    int wn;        // pairs per wavefront
    int xal;       // 1: wavefronts start at row starts (wn a multiple of nx/2)
    int fz;        // 1: the far outer z planes are loaded only where read (SB kernels)
};

// The pair (or quad) configuration of a grid: w x cells per thread, ZC z
// cells; row-aligned wavefronts where asked for and free of extra wavefronts.
inline ColX2Cfg colx2_config(const Geo &g, int zc, int w, bool rowal)
{
    const uint32_t nxh = (uint32_t)(g.n[0] / w), npairs = nxh * (uint32_t)g.n[1];
    uint32_t wn = 64;
    if (rowal && nxh <= 64) {
        const uint32_t wr = (64u / nxh) * nxh;
        if ((npairs + wr - 1) / wr <= (npairs + 63u) / 64u) wn = wr;
    }
    ColX2Cfg cc;
    cc.wn = (int)wn;
    cc.xal = wn % nxh == 0 ? 1 : 0;
    cc.fz = 0;
    cc.nxyp = (int)(((npairs + wn - 1) / wn) * 64u);
    cc.nzb = g.n[2] / zc;
    cc.dxyp = make_fastdiv((uint32_t)cc.nxyp);
    cc.dzb = make_fastdiv((uint32_t)cc.nzb);
    cc.dxh = make_fastdiv(nxh);
    return cc;
}

// v of the lane below (lane i receives lane i-1's value; lane 0 receives 0,
// which callers replace).  The plain DPP move, with no old value to merge,
// needs no copy of the destination first.
__device__ __forceinline__ float lane_shr1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138 /* wave_shr:1 */, 0xf,
                                                              0xf, true));
}
__device__ __forceinline__ double lane_shr1(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x138, 0xf, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Wave-local neighbour exchange (WL): each wavefront of the block keeps, per
// z cell, 128 pair slots in LDS: H = nx pairs (two rows) before its wn pairs,
// its own pairs, and H after (nx <= 32).  Every step a wave writes its own
// pairs and loads the halo pairs once (one pair access by lanes 0..2H-1),
// then reads the x and y neighbour pairs from the slots.  Only the wave's own
// lanes touch its slots, and the LDS executes a wave's accesses in order, so
// no barrier is needed; wave_barrier keeps the compiler from moving the
// reads above the writes.
constexpr int kWLSlots = 128;

// UXY: the x and y spacings are uniform (every face and cell of both axes
// has the same dx), so that their coefficients are wave-uniform scalars
// rather than per-lane values (fast arithmetic only).
// XAL (with WL, fast arithmetic): the wavefronts start at row starts
// (ColX2Cfg::xal), so the x face between two pairs is computed once, by the
// lane on its left, and passed to the lane on its right (lane_shr1), and the
// row-end lanes read their zero-gradient ghosts out of their own slot.
// SB (with WL and UXY): each face group branches on the sign of its velocity,
// which is uniform over the wavefront (on_sign), instead of selecting the
// upwind side per face.
template <typename T, bool EXACT, int ZC, bool FULL, bool WL = false, bool UXY = false, bool MP = false,
          bool XAL = false, bool SB = false>
__device__ __forceinline__ void colx2_body(const Geo &g, const StepPtrs<T> &p, const ColX2Cfg &cc, uint32_t b)
{
    static_assert(!(UXY && EXACT), "the exact arithmetic divides by the per-lane spacings");
    static_assert(!XAL || (WL && !EXACT), "the x-face hand-over needs row-aligned wave-local lanes");
    static_assert(!SB || (WL && UXY && !EXACT), "the sign branch takes the face groups' coefficient sets");
    const uint32_t t = b * blockDim.x + threadIdx.x;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2], nk = g.n[3], nl = g.n[4], nm = g.n[5];
    const int nxy = nx * ny, nxh = nx >> 1;
    const uint32_t pl = __builtin_amdgcn_readfirstlane(fdiv(t, cc.dxyp));
    const int lane = (int)(threadIdx.x & 63u);   // nxyp and the block are whole wavefronts
    // pairs per wavefront: fewer than 64 only where whole rows are dealt (XAL)
    const int wn = XAL ? cc.wn : 64;
    // the pair and the wavefront's first pair
    int xyp, pw;
    if constexpr (XAL) {
        pw = (int)__builtin_amdgcn_readfirstlane((t - pl * (uint32_t)cc.nxyp) >> 6) * wn;
        xyp = pw + lane;
    } else {
        xyp = (int)(t - pl * (uint32_t)cc.nxyp);
        pw = xyp & ~63;
    }
    const uint32_t kl = fdiv(pl, cc.dzb);
    const int zb = (int)(pl - kl * (uint32_t)cc.nzb);
    const int bl = (int)fdiv(kl, g.div[3]);
    const int bk = (int)kl - bl * nk;
    if (bl >= nl) return;          // whole wavefronts beyond the grid
    const int npairs = nxh * ny;
    if constexpr (!WL) {
        if (xyp >= npairs) return;   // padding lanes (no barriers below)
    }
    // WL: padding lanes stay to load halo pairs; they compute on the
    // wavefront's last pair and store nothing
    const bool live = lane < wn && xyp < npairs;
    // the wavefront's last pair (with 64 pairs per wavefront only the plane's
    // last wavefront has padding lanes)
    const int plast = XAL ? (pw + wn < npairs ? pw + wn : npairs) - 1 : npairs - 1;
    const int xyc = live ? xyp : plast;
    const int y0 = (int)fdiv((uint32_t)xyc, cc.dxh);
    const int x0 = 2 * (xyc - y0 * nxh);   // 0-based first cell of the pair
    const int x = x0 + 1, y = y0 + 1;      // 1-based: the pair is (x, x+1)
    const int z0 = zb * ZC + 1, k = bk + 1, l = bl + 1;
    constexpr uint32_t S = sizeof(T);

    const uint32_t nzkl = (uint32_t)nz * nk * nl;
    const uint32_t Qkl = (uint32_t)bk + (uint32_t)nk * (uint32_t)bl;
    const uint32_t P0 = (uint32_t)(z0 - 1) + (uint32_t)nz * Qkl;
    const uint64_t cellsB = (uint64_t)g.ncells * S;
    const __amdgpu_buffer_rsrc_t rE = mk_rsrc(p.Ein, cellsB);
    const __amdgpu_buffer_rsrc_t rO = mk_rsrc(p.Eout, cellsB);
    const uint32_t planeB = (uint32_t)nxy * S;
    const uint32_t mB = (uint32_t)g.inner[5] * S;
    const uint32_t base0 = P0 * planeB;
    // MP: E addressed one m plane at a time, through a buffer resource at the
    // start of plane mm and 32-bit byte offsets inside it, for arrays beyond
    // the reach of 32-bit offsets; else one resource per array
    auto rin = [&](int mm) {
        if constexpr (MP) return mk_rsrc(p.Ein + (size_t)(uint32_t)(mm - 1) * (uint32_t)g.inner[5], mB);
        else return rE;
    };
    auto rout = [&](int mm) {
        if constexpr (MP) return mk_rsrc(p.Eout + (size_t)(uint32_t)(mm - 1) * (uint32_t)g.inner[5], mB);
        else return rO;
    };
    auto moff = [&](int mm) -> uint32_t {
        if constexpr (MP) return 0u;
        else return (uint32_t)(mm - 1) * mB;
    };
    const bool ax = FULL || (g.act & 1), ay = FULL || (g.act & 2), az = FULL || (g.act & 4);
    const bool am = FULL || (g.act & 32);
    const double dtv = p.dt;

    const int xy0 = x0 + nx * y0;
    const uint32_t oc = (uint32_t)xy0 * S;
    // the wavefront holds a pair of row 1 or of row ny
    const bool ybw = __ballot(y == 1 || y == ny) != 0;
    // x stencil pairs: (x-2, x-1) and (x+2, x+3); at the row ends the
    // zero-gradient ghosts repeat the pair's own edge cell
    const bool xfirst = x0 == 0, xlast = x0 + 2 >= nx;
    const uint32_t oxl = xfirst ? oc : oc - 2 * S, oxr = xlast ? oc : oc + 2 * S;
    uint32_t oy[4];
    {
        const int ys[4] = {y - 2 < 1 ? 1 : y - 2, y - 1 < 1 ? 1 : y - 1, y + 1 > ny ? ny : y + 1,
                           y + 2 > ny ? ny : y + 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) oy[q] = (uint32_t)(x0 + nx * (ys[q] - 1)) * S;
    }
    // x faces x-1, x, x+1 and cells x, x+1; y faces y-1, y of row y
    double cx0 = 0, cx1 = 0, cx2 = 0, dx0 = 1, dx1 = 1, dx2 = 1, rxa = 0, rxb = 0;
    double cyl = 0, cyr = 0, dyl = 1, dyr = 1, ryy = 0;
    if constexpr (UXY) {
        const double cx = sld(p.ax.cr[0], 2), rx = sld(p.ax.cr[0], 3);
        const double cy = sld(p.ax.cr[1], 2), ry = sld(p.ax.cr[1], 3);
        cx0 = cx1 = cx2 = cx;
        rxa = rxb = rx;
        cyl = cyr = cy;
        ryy = ry;
    } else {
        if (ax) {
            cx0 = p.ax.cfl[0][x - 1]; cx1 = p.ax.cfl[0][x]; cx2 = p.ax.cfl[0][x + 1];
            dx0 = p.ax.dx[0][x - 1]; dx1 = p.ax.dx[0][x]; dx2 = p.ax.dx[0][x + 1];
            rxa = p.ax.rdx[0][x]; rxb = p.ax.rdx[0][x + 1];
        }
        if (ay) { cyl = p.ax.cfl[1][y - 1]; cyr = p.ax.cfl[1][y]; dyl = p.ax.dx[1][y - 1]; dyr = p.ax.dx[1][y]; ryy = p.ax.rdx[1][y]; }
    }

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
    // WL slots: the halo pair this lane loads, and the slots of the x and y
    // neighbour pairs
    using V2 = T __attribute__((ext_vector_type(2)));
    const int H = nx;                              // pairs in two rows
    const int wv = (int)(threadIdx.x >> 6);
    int hslot = 0, hp = -1;
    if (lane < H) { hp = pw - H + lane; hslot = lane; }
    else if (lane < 2 * H) { hp = pw + wn + (lane - H); hslot = wn + lane; }
    const bool hval = WL && hp >= 0 && hp < npairs;
    uint32_t ohal = 0;
    if (hval) {
        const int yh = hp / nxh;
        ohal = (uint32_t)(2 * (hp - yh * nxh) + nx * yh) * S;
    }
    int sy[4] = {0, 0, 0, 0};
    if constexpr (WL) {
        const int ys[4] = {y - 2 < 1 ? 1 : y - 2, y - 1 < 1 ? 1 : y - 1, y + 1 > ny ? ny : y + 1,
                           y + 2 > ny ? ny : y + 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) sy[q] = (ys[q] - 1) * nxh + (x0 >> 1) - pw + H;
    }
    const int sown = xyc - pw + H;
    // XAL x neighbours, in elements of T over the slots: the pair left of
    // the own one, and the pair right of it; at a row start the left read
    // starts one element later, so that its second element is the own first
    // cell (the zero-gradient ghost that face x reads), and at a row end the
    // right read one element earlier, so that its first element is the own
    // second cell (the elements beyond are read but not used: the boundary
    // faces are v*E of the edge cell)
    const int lel = 2 * sown - (xfirst ? 1 : 2), rel = 2 * sown + (xlast ? 1 : 2);
    auto poff = [&](int j, int mm) -> uint32_t { return base0 + (uint32_t)j * planeB + moff(mm); };
    auto ldm = [&](int j, int mm) -> Pair<T> {
        mm = mm < 1 ? 1 : (mm > nm ? nm : mm);
        return bld2<T>(rin(mm), oc, poff(j, mm));
    };
    auto ldz = [&](int q, int mm) -> Pair<T> {
        const int mu = zmir[q] ? nm + 1 - mm : mm;
        return bld2<T>(rin(mu), oc, base0 + (uint32_t)(zsh[q] * (int)planeB) + moff(mu));
    };
    constexpr int LA = ZC >= 4 ? 3 : 2;   // own-plane look-ahead, as in the column kernel
    // the m window E(m-1), E(m), E(m+1), E(m+2) of each z cell: four slots
    // that the march renames rather than copies (RN: the uniform-spacing
    // kernels and binary32; otherwise, and with four z cells per thread,
    // where a fifth slot holds E(m+3), the slots are rotated by copies, which
    // needs fewer registers)
    Pair<T> win[5][ZC], FLm[ZC];
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
        win[1][j] = ldm(j, 1);
        win[2][j] = ldm(j, 2);
        if constexpr (LA == 3) win[3][j] = ldm(j, 3);
        win[0][j] = win[1][j];
        FLm[j] = Pair<T>{T(0), T(0)};
    }
    Pair<T>(&w0i)[ZC] = win[1];
    Pair<T>(&wp1i)[ZC] = win[2];
    Pair<T> zq[4];
    // lanes without a halo pair never store theirs: the XAL kernels leave it
    // unset (no zeroing per step); the others zero it, which measured fewer
    // registers there
    auto ldhal = [&](int j, int mm) -> Pair<T> {
        Pair<T> h;
        if constexpr (!XAL) h = Pair<T>{T(0), T(0)};
        if (hval) h = bld2<T>(rin(mm), ohal, poff(j, mm));
        return h;
    };
    if (am) {
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const T vm0 = sld(p.hom.vm, P0 + j);
            // m face 0: v*E(1), the reference's formula over the zero-gradient
            // ghosts in exact arithmetic (v*(E+E)*0.5 differs from v*E where
            // the product is subnormal)
            if constexpr (EXACT) {
                const double d0 = sld(p.ax.dx[5], 0);
                FLm[j] = Pair<T>{sb_flux_exact(vm0, p.dt, d0, w0i[j].a, w0i[j].a, w0i[j].a, wp1i[j].a),
                                 sb_flux_exact(vm0, p.dt, d0, w0i[j].b, w0i[j].b, w0i[j].b, wp1i[j].b)};
            } else {
                FLm[j] = Pair<T>{vm0 * w0i[j].a, vm0 * w0i[j].b};
            }
        }
    }

    auto step = [&](int m, Pair<T>(&wm1)[ZC], Pair<T>(&w0)[ZC], Pair<T>(&wp1)[ZC], Pair<T>(&wp2)[ZC],
                    Pair<T>(&wnx)[ZC]) {
        const uint32_t q20 = P0 + (uint32_t)(m - 1) * nzkl;
        // this step's coefficients: the plane records of the ZC cells, side by side
        const T *rp = p.hom.rec + (size_t)kRec * q20;
        auto rv = [&](int j, int f) -> T { return sld(rp, (uint32_t)(j * kRec + f)); };
        // a face coefficient set of cell j's record, and the velocity it came from
        auto rk = [&](int j, int at, uint32_t bit) -> SbCoef<T> {
            SbCoef<T> k;
            k.vp = rv(j, at);
            k.vn = rv(j, at + 1);
            k.w = rv(j, at + 2);
            const T fv = rv(j, kRecFlags);
            uint32_t fl;
            if constexpr (sizeof(T) == 8) fl = (uint32_t)__builtin_bit_cast(uint64_t, fv);
            else fl = __builtin_bit_cast(uint32_t, fv);
            k.pos = (fl & bit) != 0;
            return k;
        };
        auto rvel = [&](int j, int at) -> T { return rv(j, at) + rv(j, at + 1); };   // vp + vn = v exactly
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            if constexpr (LA == 3) wnx[j] = ldm(j, m + 3);
            else wp2[j] = ldm(j, m + 2);   // E(m+2), for this step's m face
        }
        bool farz = false;   // the far outer z planes gated by the sign flags
        if constexpr (SB) farz = cc.fz != 0;
        if (farz) {
            // E(z0-2) only for a lower face with v > 0, E(z0+ZC+1) only for an
            // upper face with v <= 0: the upwind sides that read them.  The
            // flags are this step's record flags, which the sign branches
            // read anyway; the loads wait for them, which pays where the
            // grid fills the device (binary32 at 24^3 x 16^3: 0.1336 instead
            // of 0.1394 ms, 32^3 0.2829 instead of 0.2908, E bitwise the
            // same) and costs on smaller grids, which keep all four loads
            // (ColX2Cfg::fz; 16^3 x 12^3 with the gate on: 0.0311 against
            // 0.0289 ms)
            auto flg = [&](int j) -> uint32_t {
                const T fv = rv(j, kRecFlags);
                if constexpr (sizeof(T) == 8) return (uint32_t)__builtin_bit_cast(uint64_t, fv);
                else return __builtin_bit_cast(uint32_t, fv);
            };
            const bool need0 = (flg(0) & kPosZlo) != 0, need3 = (flg(ZC - 1) & kPosZhi) == 0;
            zq[0] = Pair<T>{T(0), T(0)};
            zq[3] = Pair<T>{T(0), T(0)};
            zq[1] = az ? ldz(1, m) : Pair<T>{T(0), T(0)};
            zq[2] = az ? ldz(2, m) : Pair<T>{T(0), T(0)};
            if (az && need0) zq[0] = ldz(0, m);
            if (az && need3) zq[3] = ldz(3, m);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) zq[q] = az ? ldz(q, m) : Pair<T>{T(0), T(0)};
        }
        Pair<T> hq[ZC];
        if constexpr (WL) {
#pragma unroll
            for (int j = 0; j < ZC; ++j) hq[j] = ldhal(j, m);
        }

        T FzA[ZC + 1], FzB[ZC + 1];
        if (az) {
            T za[ZC + 4], zbv[ZC + 4];
            za[0] = zq[0].a; za[1] = zq[1].a; zbv[0] = zq[0].b; zbv[1] = zq[1].b;
#pragma unroll
            for (int j = 0; j < ZC; ++j) { za[2 + j] = w0[j].a; zbv[2 + j] = w0[j].b; }
            za[ZC + 2] = zq[2].a; za[ZC + 3] = zq[3].a; zbv[ZC + 2] = zq[2].b; zbv[ZC + 3] = zq[3].b;
#pragma unroll
            for (int i = 0; i <= ZC; ++i) {
                if constexpr (EXACT) {
                    const int fz = z0 - 1 + i;
                    const T v = i < ZC ? rvel(i, kRecZlo) : rvel(ZC - 1, kRecZhi);
                    const double dzf = sld(p.ax.dx[2], fz), czf = sld(p.ax.cfl[2], fz);
                    FzA[i] = face_flux<T, EXACT>(v, dtv, dzf, czf, za[i], za[i + 1], za[i + 2], za[i + 3]);
                    FzB[i] = face_flux<T, EXACT>(v, dtv, dzf, czf, zbv[i], zbv[i + 1], zbv[i + 2], zbv[i + 3]);
                } else {
                    const SbCoef<T> k = i < ZC ? rk(i, kRecZlo, kPosZlo) : rk(ZC - 1, kRecZhi, kPosZhi);
                    on_sign<SB>(k.pos, [&](auto sg) {
                        using S = decltype(sg);
                        FzA[i] = sb_flux_s<T, S>(k, za[i], za[i + 1], za[i + 2], za[i + 3]);
                        FzB[i] = sb_flux_s<T, S>(k, zbv[i], zbv[i + 1], zbv[i + 2], zbv[i + 3]);
                    });
                }
            }
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            // raised wave priority while this cell pair's neighbour loads issue,
            // so that they leave early and other waves' arithmetic covers them
            // (binary64 at 24^3 x 16^3: 0.396 instead of 0.417 ms)
            __builtin_amdgcn_s_setprio(1);
            const uint32_t pb = poff(j, m);   // cell pair (x, y, z0+j, k, l, m)
            const __amdgpu_buffer_rsrc_t rEm = rin(m);
            const Pair<T> e = w0[j];
            T dEa = T(0), dEb = T(0);
            [[maybe_unused]] V2 *sl = nullptr;
            if constexpr (WL) {
                __shared__ V2 sW[4][ZC][kWLSlots];
                sl = &sW[wv][j][0];
                __builtin_amdgcn_wave_barrier();   // the previous step's reads stay above
                V2 own;
                own.x = e.a;
                own.y = e.b;
                sl[sown] = own;
                if (hval) {
                    V2 h;
                    h.x = hq[j].a;
                    h.y = hq[j].b;
                    sl[hslot] = h;
                }
                __builtin_amdgcn_wave_barrier();   // the reads below stay below
            }
            auto slot = [&](int i) -> Pair<T> {
                const V2 v = sl[i];
                return Pair<T>{v.x, v.y};
            };
            if (ax) {
                Pair<T> L, R;
                if constexpr (XAL) {
                    const T *se = (const T *)sl;
                    L = Pair<T>{se[lel], se[lel + 1]};
                    R = Pair<T>{se[rel], se[rel + 1]};
                } else if constexpr (WL) {
                    L = slot(sown - 1);
                    R = slot(sown + 1);
                } else {
                    L = bld2<T>(rEm, oxl, pb);
                    R = bld2<T>(rEm, oxr, pb);
                }
                if constexpr (EXACT) {
                    if (xfirst) L = Pair<T>{e.a, e.a};
                    if (xlast) R = Pair<T>{e.b, e.b};
                } else if constexpr (!XAL) {
                    // the fast arithmetic replaces the boundary faces (F0 where
                    // xfirst, F2 where xlast) by v*E of the edge cell, so only
                    // the ghosts that face 1 reads are needed (the XAL reads
                    // above take them from the own pair)
                    if (xfirst) L.b = e.a;
                    if (xlast) R.a = e.b;
                }
                const T v0 = rv(j, kRecVx0), vb = rv(j, kRecVxn);
                T F0, F1, F2;
                if constexpr (EXACT) {
                    const T vi = rvel(j, kRecXi);
                    const T vL = x == 1 ? v0 : vi, vR = x + 1 == nx ? vb : vi;
                    F0 = face_flux<T, EXACT>(vL, dtv, dx0, cx0, L.a, L.b, e.a, e.b);
                    F1 = face_flux<T, EXACT>(vi, dtv, dx1, cx1, L.b, e.a, e.b, R.a);
                    F2 = face_flux<T, EXACT>(vR, dtv, dx2, cx2, e.a, e.b, R.a, R.b);
                } else {
                    // interior faces from the coefficient sets; the boundary
                    // faces 0 and nx, whose zero-gradient ghost makes R = 0,
                    // are v*E of the edge cell
                    SbCoef<T> k0, k1, k2;
                    if constexpr (UXY) {
                        k0 = k1 = k2 = rk(j, kRecXi, kPosXi);
                    } else {
                        const T vi = rvel(j, kRecXi);
                        k0 = sb_coef<T>(vi, (T)cx0);
                        k1 = sb_coef<T>(vi, (T)cx1);
                        k2 = sb_coef<T>(vi, (T)cx2);
                    }
                    // the three faces share the sign of vi
                    on_sign<SB>(k1.pos, [&](auto sg) {
                        using S = decltype(sg);
                        F1 = sb_flux_s<T, S>(k1, L.b, e.a, e.b, R.a);
                        if constexpr (!XAL) F0 = sb_flux_s<T, S>(k0, L.a, L.b, e.a, e.b);
                        F2 = sb_flux_s<T, S>(k2, e.a, e.b, R.a, R.b);
                    });
                    if constexpr (XAL) {
                        // face x-1 is face x+1 of the pair to the left, held by
                        // the lane below, the wavefronts starting at row starts
                        // (the same operands and coefficients, so the same bits)
                        if (xlast) F2 = vb * e.b;
                        F0 = lane_shr1(F2);   // lane 0 starts a row
                        if (xfirst) F0 = v0 * e.a;
                    } else {
                        if (xfirst) F0 = v0 * e.a;
                        if (xlast) F2 = vb * e.b;
                    }
                }
                dEa = div_term<T, EXACT>(dEa, F1, F0, dx1, rxa);
                dEb = div_term<T, EXACT>(dEb, F2, F1, dx2, rxb);
            }
            if (ay) {
                Pair<T> a0, a1, a3, a4;
                if constexpr (WL) {
                    a0 = slot(sy[0]);
                    a1 = slot(sy[1]);
                    a3 = slot(sy[2]);
                    a4 = slot(sy[3]);
                } else {
                    a0 = bld2<T>(rEm, oy[0], pb);
                    a1 = bld2<T>(rEm, oy[1], pb);
                    a3 = bld2<T>(rEm, oy[2], pb);
                    a4 = bld2<T>(rEm, oy[3], pb);
                }
                __builtin_amdgcn_s_setprio(0);
                const T v0 = rv(j, kRecVy0), vb = rv(j, kRecVyn);
                T FLa, FRa, FLb, FRb;
                if constexpr (EXACT) {
                    const T vi = rvel(j, kRecYi);
                    const T vL = y == 1 ? v0 : vi, vR = y == ny ? vb : vi;
                    FLa = face_flux<T, EXACT>(vL, dtv, dyl, cyl, a0.a, a1.a, e.a, a3.a);
                    FRa = face_flux<T, EXACT>(vR, dtv, dyr, cyr, a1.a, e.a, a3.a, a4.a);
                    FLb = face_flux<T, EXACT>(vL, dtv, dyl, cyl, a0.b, a1.b, e.b, a3.b);
                    FRb = face_flux<T, EXACT>(vR, dtv, dyr, cyr, a1.b, e.b, a3.b, a4.b);
                } else {
                    SbCoef<T> kl, kr;
                    if constexpr (UXY) {
                        kl = kr = rk(j, kRecYi, kPosYi);
                    } else {
                        const T vi = rvel(j, kRecYi);
                        kl = sb_coef<T>(vi, (T)cyl);
                        kr = sb_coef<T>(vi, (T)cyr);
                    }
                    on_sign<SB>(kl.pos, [&](auto sg) {   // both faces have the sign of vi
                        using S = decltype(sg);
                        FLa = sb_flux_s<T, S>(kl, a0.a, a1.a, e.a, a3.a);
                        FRa = sb_flux_s<T, S>(kr, a1.a, e.a, a3.a, a4.a);
                        FLb = sb_flux_s<T, S>(kl, a0.b, a1.b, e.b, a3.b);
                        FRb = sb_flux_s<T, S>(kr, a1.b, e.b, a3.b, a4.b);
                    });
                    // boundary rows: only wavefronts that hold row 1 or row ny
                    // (a wave-uniform branch)
                    if (ybw) {
                        if (y == 1) { FLa = v0 * e.a; FLb = v0 * e.b; }
                        if (y == ny) { FRa = vb * e.a; FRb = vb * e.b; }
                    }
                }
                dEa = div_term<T, EXACT>(dEa, FRa, FLa, dyr, ryy);
                dEb = div_term<T, EXACT>(dEb, FRb, FLb, dyr, ryy);
            }
            if (!ay) __builtin_amdgcn_s_setprio(0);
            if (az) {
                if constexpr (EXACT) {
                    const int zc = z0 + j;
                    const double dzc = sld(p.ax.dx[2], zc), rzc = sld(p.ax.rdx[2], zc);
                    dEa = div_term<T, EXACT>(dEa, FzA[j + 1], FzA[j], dzc, rzc);
                    dEb = div_term<T, EXACT>(dEb, FzB[j + 1], FzB[j], dzc, rzc);
                } else {
                    // 1/dz of the cell from its record, already in T
                    const T rz = rv(j, kRecRz);
                    dEa = div_term_r<T>(dEa, FzA[j + 1], FzA[j], rz);
                    dEb = div_term_r<T>(dEb, FzB[j + 1], FzB[j], rz);
                }
            }
            if (am) {
                T FRa, FRb;
                if constexpr (EXACT) {
                    const T vR = rvel(j, kRecMhi);
                    const double dm = sld(p.ax.dx[5], m), cm = sld(p.ax.cr[5], 2 * m);
                    FRa = face_flux<T, EXACT>(vR, dtv, dm, cm, wm1[j].a, e.a, wp1[j].a, wp2[j].a);
                    FRb = face_flux<T, EXACT>(vR, dtv, dm, cm, wm1[j].b, e.b, wp1[j].b, wp2[j].b);
                } else {
                    const SbCoef<T> k = rk(j, kRecMhi, kPosMhi);
                    on_sign<SB>(k.pos, [&](auto sg) {
                        using S = decltype(sg);
                        FRa = sb_flux_s<T, S>(k, wm1[j].a, e.a, wp1[j].a, wp2[j].a);
                        FRb = sb_flux_s<T, S>(k, wm1[j].b, e.b, wp1[j].b, wp2[j].b);
                    });
                }
                if constexpr (EXACT) {
                    const double dm = sld(p.ax.dx[5], m);
                    dEa = div_term<T, EXACT>(dEa, FRa, FLm[j].a, dm, 0.0);
                    dEb = div_term<T, EXACT>(dEb, FRb, FLm[j].b, dm, 0.0);
                } else {
                    const T rm = rv(j, kRecRm);
                    dEa = div_term_r<T>(dEa, FRa, FLm[j].a, rm);
                    dEb = div_term_r<T>(dEb, FRb, FLm[j].b, rm);
                }
                FLm[j] = Pair<T>{FRa, FRb};
            }
            const T lam = rv(j, kRecLam);
            dEa = dEa - e.a * lam;
            dEb = dEb - e.b * lam;
            Pair<T> en;
            if constexpr (EXACT) {
                en.a = e.a + p.dt * dEa;
                en.b = e.b + p.dt * dEb;
            } else {
                const T dtt = rv(j, kRecDt);   // the step's dt in T
                en.a = e.a + dtt * dEa;
                en.b = e.b + dtt * dEb;
            }
            if (!WL || live) bst2<T>(en, rout(m), oc, pb);
        }
    };
    constexpr bool RN = LA == 2 && (UXY || sizeof(T) == 4);
    if constexpr (RN) {
        // four steps per trip, the window's slots renamed step by step
        for (int m = 1; m <= nm; m += 4) {
            step(m, win[0], win[1], win[2], win[3], win[4]);
            if (m + 1 <= nm) step(m + 1, win[1], win[2], win[3], win[0], win[4]);
            if (m + 2 <= nm) step(m + 2, win[2], win[3], win[0], win[1], win[4]);
            if (m + 3 <= nm) step(m + 3, win[3], win[0], win[1], win[2], win[4]);
        }
    } else {
        for (int m = 1; m <= nm; ++m) {
            step(m, win[0], win[1], win[2], win[3], win[4]);
#pragma unroll
            for (int j = 0; j < ZC; ++j) {
                win[0][j] = win[1][j];
                win[1][j] = win[2][j];
                win[2][j] = win[3][j];
                if constexpr (LA == 3) win[3][j] = win[4][j];
            }
        }
    }
}

template <typename T, bool EXACT, int ZC, bool FULL, bool WL = false, bool UXY = false, bool MP = false,
          bool XAL = false, bool SB = false>
__global__ void __launch_bounds__(256) k_step_colx2(const Geo g, const StepPtrs<T> p, const ColX2Cfg cc, uint32_t nbl)
{
    const uint32_t b = xcd_logical_block(blockIdx.x, nbl);
    if (b >= nbl) return;
    colx2_body<T, EXACT, ZC, FULL, WL, UXY, MP, XAL, SB>(g, p, cc, b);
}

// The same kernel held to 128 VGPRs, four wavefronts per SIMD: the renamed
// binary64 window of the uniform-spacing kernel takes one register more than
// that otherwise (and then three wavefronts per SIMD), without spilling when
// held to it.
template <typename T, bool EXACT, int ZC, bool FULL, bool WL = false, bool UXY = false, bool MP = false,
          bool XAL = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_step_colx2_w4(const Geo g, const StepPtrs<T> p, const ColX2Cfg cc, uint32_t nbl)
{
    const uint32_t b = xcd_logical_block(blockIdx.x, nbl);
    if (b >= nbl) return;
    colx2_body<T, EXACT, ZC, FULL, WL, UXY, MP, XAL>(g, p, cc, b);
}
