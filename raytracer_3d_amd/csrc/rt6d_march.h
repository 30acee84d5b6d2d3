// rt6d_march.h -- the m-marching step kernel (included by rt6d_kernels.h).
//
// One block owns ZC consecutive (x,y) planes, z = z0 .. z0+ZC-1, of one (k,l)
// position (nx*ny <= 1024 threads, one (x,y) column per thread, ZC cells per
// thread) and marches them through m = 1..nm:
//  * E along m lives in a register window per cell (E(m-1..m+2), E(m+3) in
//    flight), so every E value is read from HBM once; the m-face flux of step
//    m is the left face of step m+1 and is carried, not recomputed;
//  * the current (x,y) planes are staged in LDS; each cell computes the flux
//    of its right x and y faces once, publishes it in LDS and takes its left
//    faces from its neighbours.  Faces 0 of the clamped axes equal v*E(1)
//    exactly (the ghost copy makes R = E(1)-E(0) vanish, superbee_fluxes.f90:
//    29-30), so they cost one product;
//  * along z the ZC cells share their ZC-1 inner faces; the four outer
//    neighbours are read from global memory one m-plane ahead (L2 hits: the
//    neighbouring blocks march the same planes concurrently);
//  * k and l (general path only) read their neighbours from global memory
//    and compute both faces per cell.
// Per face the operations are those of k_step, so the results are identical.
// Included inside namespace rt6d.
#pragma once

struct MarchCfg {
    FastDiv dzb;   // division by the number of z blocks
    int nzb;       // nz / ZC
};

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

template <typename T, bool EXACT, bool HOMOG, int ZC>
__global__ void __launch_bounds__(1024) k_step_march(const Geo g, const StepPtrs<T> p, const MarchCfg mc,
                                                     uint32_t nb_logical)
{
    extern __shared__ unsigned char smem_raw[];
    const int NT = blockDim.x;
    T *sE = reinterpret_cast<T *>(smem_raw);   // [ZC][NT]
    T *sFx = sE + ZC * NT;                       // [ZC][NT]
    T *sFy = sFx + ZC * NT;                      // [ZC][NT]

    const uint32_t b = xcd_logical_block(blockIdx.x, nb_logical);
    if (b >= nb_logical) return;  // uniform over the block
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2], nk = g.n[3], nl = g.n[4], nm = g.n[5];
    const int nxy = nx * ny;
    const uint32_t bkl = fdiv(b, mc.dzb);
    const int bzb = (int)(b - bkl * (uint32_t)mc.nzb);
    const int bl = (int)fdiv(bkl, g.div[3]);
    const int bk = (int)bkl - bl * nk;
    const int z0 = bzb * ZC + 1, k = bk + 1, l = bl + 1;

    const int tid = threadIdx.x;
    const bool live = tid < nxy;
    const int y0 = (int)fdiv((uint32_t)tid, g.div[0]);
    const int x0 = tid - y0 * nx;
    const int x = x0 + 1, y = y0 + 1;

    const uint32_t s5 = (uint32_t)g.inner[5];   // m stride
    const uint32_t nzkl = (uint32_t)nz * nk * nl;
    const uint32_t Qkl = (uint32_t)bk + (uint32_t)nk * (uint32_t)bl;
    const uint32_t P0 = (uint32_t)(z0 - 1) + (uint32_t)nz * Qkl;   // plane index of (z0,k,l), m = 1
    const uint32_t col0 = (uint32_t)nxy * P0 + (uint32_t)tid;      // cell (x,y,z0,k,l,1)
    const T *__restrict__ E = p.Ein;
    const bool ax = g.act & 1, ay = g.act & 2, az = g.act & 4;
    const bool ak = !HOMOG && (g.act & 8), al = !HOMOG && (g.act & 16), am = g.act & 32;
    const double dtv = p.dt;

    // per-thread x/y coefficients of the right faces (face x uses dxt(x))
    double cxr = 0, rxx = 0, dxxr = 1, cyr = 0, ryy = 0, dxyr = 1;
    if (ax && live) { cxr = p.ax.cfl[0][x]; rxx = p.ax.rdx[0][x]; dxxr = p.ax.dx[0][x]; }
    if (ay && live) { cyr = p.ax.cfl[1][y]; ryy = p.ax.rdx[1][y]; dxyr = p.ax.dx[1][y]; }
    double ckl = 0, ckr = 0, rkk = 0, dkl = 1, dkr = 1, cll = 0, clr = 0, rll = 0, dll = 1, dlr = 1;
    if (ak) { ckl = p.ax.cfl[3][k - 1]; ckr = p.ax.cfl[3][k]; rkk = p.ax.rdx[3][k]; dkl = p.ax.dx[3][k - 1]; dkr = p.ax.dx[3][k]; }
    if (al) { cll = p.ax.cfl[4][l - 1]; clr = p.ax.cfl[4][l]; rll = p.ax.rdx[4][l]; dll = p.ax.dx[4][l - 1]; dlr = p.ax.dx[4][l]; }

    // LDS neighbour indices along x and y (zero-gradient ghosts = clamped);
    // padding threads (tid >= nx*ny) read their own slot only
    const int ixm1 = live ? tid - (x > 1 ? 1 : 0) : tid;
    const int ixp1 = live ? tid + (x < nx ? 1 : 0) : tid;
    const int ixp2 = live ? tid + (x + 2 <= nx ? 2 : (x < nx ? 1 : 0)) : tid;
    const int iym1 = live ? tid - (y > 1 ? nx : 0) : tid;
    const int iyp1 = live ? tid + (y < ny ? nx : 0) : tid;
    const int iyp2 = live ? tid + (y + 2 <= ny ? 2 * nx : (y < ny ? nx : 0)) : tid;

    // outer z neighbours: z0-2, z0-1 (relative to cell z0) and z0+ZC, z0+ZC+1
    // (relative to cell z0+ZC-1); `mir` marks m-mirrored reflection ghosts
    int zoff[4];
    bool zmir[4];
    {
        const int zz0[4] = {z0 - 2, z0 - 1, z0 + ZC, z0 + ZC + 1};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int zz = zz0[q];
            bool mir = false;
            if (zz < 1) { if (g.down) { zz = 1 - zz; mir = true; } else zz = 1; }
            else if (zz > nz) { if (g.up) { zz = 2 * nz + 1 - zz; mir = true; } else zz = nz; }
            zoff[q] = (zz - z0) * nxy;     // relative to cell z0
            zmir[q] = mir;
        }
    }
    int koff[4] = {0, 0, 0, 0}, loff[4] = {0, 0, 0, 0};
    if (!HOMOG) {
        const int os[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int kk = k + os[q];
            kk = kk < 1 ? 1 : (kk > nk ? nk : kk);
            koff[q] = (kk - k) * g.inner[3];
            int ll = l + os[q];
            ll = ll < 1 ? 1 : (ll > nl ? nl : ll);
            loff[q] = (ll - l) * g.inner[4];
        }
    }

    // register windows along m (clamped ghosts, integrate.f90:84-94)
    auto ldm = [&](int j, int mm) -> T {
        mm = mm < 1 ? 1 : (mm > nm ? nm : mm);
        return E[(size_t)col0 + (size_t)j * nxy + (size_t)(mm - 1) * s5];
    };
    // z neighbours of plane m (outer ones; mirrored ghosts read plane nm+1-m)
    auto ldz = [&](int q, int mm) -> T {
        const int mu = zmir[q] ? nm + 1 - mm : mm;
        return E[(int64_t)col0 + zoff[q] + (int64_t)(mu - 1) * s5];
    };
    T wm1[ZC], w0[ZC], wp1[ZC], wp2[ZC], FLm[ZC];
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
        w0[j] = live ? ldm(j, 1) : T(0);
        wp1[j] = live ? ldm(j, 2) : T(0);
        wp2[j] = live ? ldm(j, 3) : T(0);
        wm1[j] = w0[j];
        FLm[j] = T(0);
    }
    T zq[4] = {T(0), T(0), T(0), T(0)};
    if (az && live) {
#pragma unroll
        for (int q = 0; q < 4; ++q) zq[q] = ldz(q, 1);
    }

    // general-path face-array bases (at m = 1) and per-m strides
    const uint32_t sfx = (uint32_t)(nx + 1) * ny * nzkl;
    const uint32_t sfy = (uint32_t)nx * (ny + 1) * nzkl;
    const uint32_t sfz = (uint32_t)nxy * (nz + 1) * nk * nl;
    const uint32_t sfk = (uint32_t)nxy * nz * (nk + 1) * nl;
    const uint32_t sfl = (uint32_t)nxy * nz * nk * (nl + 1);
    const uint32_t fz0 = (uint32_t)tid + (uint32_t)nxy * ((uint32_t)(z0 - 1) + (uint32_t)(nz + 1) * Qkl);
    const uint32_t fk0 = (uint32_t)tid + (uint32_t)nxy * (uint32_t)(z0 - 1) +
                         (uint32_t)nxy * nz * ((uint32_t)bk + (uint32_t)(nk + 1) * bl);
    const uint32_t fl0 = (uint32_t)tid + (uint32_t)nxy * ((uint32_t)(z0 - 1) + (uint32_t)nz * bk) +
                         (uint32_t)nxy * nz * nk * (uint32_t)bl;

    // m face 0 of every cell: closed form v*E(1)
    if (am && live) {
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const T vm0 = HOMOG ? p.hom.vm[P0 + j] : p.vel[5][col0 + (uint32_t)j * nxy];
            FLm[j] = vm0 * w0[j];
        }
    }

    for (int m = 1; m <= nm; ++m) {
        const uint32_t q3 = Qkl + (uint32_t)(m - 1) * (uint32_t)nk * nl;   // (k,l,m) index
        const uint32_t q20 = P0 + (uint32_t)(m - 1) * nzkl;                 // (z0,k,l,m) index
        // loads for the next plane, in flight during this one
        T wnx[ZC], znx[4];
#pragma unroll
        for (int j = 0; j < ZC; ++j) wnx[j] = live ? ldm(j, m + 3) : T(0);
#pragma unroll
        for (int q = 0; q < 4; ++q) znx[q] = (az && live && m < nm) ? ldz(q, m + 1) : T(0);

#pragma unroll
        for (int j = 0; j < ZC; ++j) sE[j * NT + tid] = w0[j];
        __syncthreads();

        // right x / y faces, published in LDS
        T vx0[ZC], vy0[ZC];
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const uint32_t q2 = q20 + (uint32_t)j;
            const uint32_t Pm = P0 + (uint32_t)j + (uint32_t)(m - 1) * nzkl;
            vx0[j] = T(0);
            vy0[j] = T(0);
            const T *sj = sE + j * NT;
            if (ax) {
                T vR;
                if (HOMOG) {
                    const T vi = p.hom.vx3[p.hom.n4 + q2], vb = p.hom.vx3[2 * p.hom.n4 + q2];
                    vR = (live && x == nx) ? vb : vi;
                    vx0[j] = p.hom.vx3[q2];
                } else {
                    const uint32_t fb = (uint32_t)(nx + 1) * ((uint32_t)y0 + (uint32_t)ny * Pm);
                    vR = live ? p.vel[0][fb + (uint32_t)x] : T(0);
                    vx0[j] = (live && x == 1) ? p.vel[0][fb] : T(0);
                }
                sFx[j * NT + tid] = face_flux<T, EXACT>(vR, dtv, dxxr, cxr, sj[ixm1], w0[j], sj[ixp1], sj[ixp2]);
            }
            if (ay) {
                T vR;
                if (HOMOG) {
                    const T vi = p.hom.vy3[p.hom.n4 + q2], vb = p.hom.vy3[2 * p.hom.n4 + q2];
                    vR = (live && y == ny) ? vb : vi;
                    vy0[j] = p.hom.vy3[q2];
                } else {
                    const uint32_t fb = (uint32_t)x0 + (uint32_t)nx * (uint32_t)(ny + 1) * Pm;
                    vR = live ? p.vel[1][fb + (uint32_t)nx * (uint32_t)y] : T(0);
                    vy0[j] = (live && y == 1) ? p.vel[1][fb] : T(0);
                }
                sFy[j * NT + tid] = face_flux<T, EXACT>(vR, dtv, dxyr, cyr, sj[iym1], w0[j], sj[iyp1], sj[iyp2]);
            }
        }
        __syncthreads();

        if (live) {
            // z faces 0..ZC of this block's column stack
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
                    const int fz = z0 - 1 + i;   // face index = its left cell
                    T v;
                    if (HOMOG) v = p.hom.vz[(uint32_t)fz + (uint32_t)(nz + 1) * q3];
                    else v = p.vel[2][fz0 + (uint32_t)i * nxy + (uint32_t)(m - 1) * sfz];
                    Fz[i] = face_flux<T, EXACT>(v, dtv, p.ax.dx[2][fz], p.ax.cfl[2][fz], zs[i], zs[i + 1], zs[i + 2],
                                                zs[i + 3]);
                }
            }
#pragma unroll
            for (int j = 0; j < ZC; ++j) {
                const uint32_t cell = col0 + (uint32_t)j * nxy + (uint32_t)(m - 1) * s5;
                const uint32_t q2 = q20 + (uint32_t)j;
                const T e0 = w0[j];
                T dE = T(0);
                if (ax) {
                    const T FR = sFx[j * NT + tid];
                    const T FL = x > 1 ? sFx[j * NT + tid - 1] : vx0[j] * e0;
                    dE = div_term<T, EXACT>(dE, FR, FL, dxxr, rxx);
                }
                if (ay) {
                    const T FR = sFy[j * NT + tid];
                    const T FL = y > 1 ? sFy[j * NT + tid - nx] : vy0[j] * e0;
                    dE = div_term<T, EXACT>(dE, FR, FL, dxyr, ryy);
                }
                if (az) {
                    const int zc = z0 + j;
                    dE = div_term<T, EXACT>(dE, Fz[j + 1], Fz[j], p.ax.dx[2][zc], p.ax.rdx[2][zc]);
                }
                if (ak) {
                    const T km2 = E[(int64_t)cell + koff[0]], km1 = E[(int64_t)cell + koff[1]];
                    const T kp1 = E[(int64_t)cell + koff[2]], kp2 = E[(int64_t)cell + koff[3]];
                    const uint32_t f = fk0 + (uint32_t)j * nxy + (uint32_t)(m - 1) * sfk;
                    const T vL = p.vel[3][f], vR = p.vel[3][f + (uint32_t)nxy * nz];
                    const T FL = face_flux<T, EXACT>(vL, dtv, dkl, ckl, km2, km1, e0, kp1);
                    const T FR = face_flux<T, EXACT>(vR, dtv, dkr, ckr, km1, e0, kp1, kp2);
                    dE = div_term<T, EXACT>(dE, FR, FL, dkr, rkk);
                }
                if (al) {
                    const T lm2 = E[(int64_t)cell + loff[0]], lm1 = E[(int64_t)cell + loff[1]];
                    const T lp1 = E[(int64_t)cell + loff[2]], lp2 = E[(int64_t)cell + loff[3]];
                    const uint32_t f = fl0 + (uint32_t)j * nxy + (uint32_t)(m - 1) * sfl;
                    const T vL = p.vel[4][f], vR = p.vel[4][f + (uint32_t)nxy * nz * nk];
                    const T FL = face_flux<T, EXACT>(vL, dtv, dll, cll, lm2, lm1, e0, lp1);
                    const T FR = face_flux<T, EXACT>(vR, dtv, dlr, clr, lm1, e0, lp1, lp2);
                    dE = div_term<T, EXACT>(dE, FR, FL, dlr, rll);
                }
                if (am) {
                    const T vR = HOMOG ? p.hom.vm[q2 + p.hom.inner4m] : p.vel[5][cell + s5];
                    const double dm = p.ax.dx[5][m];
                    const T FR = face_flux<T, EXACT>(vR, dtv, dm, p.ax.cfl[5][m], wm1[j], e0, wp1[j], wp2[j]);
                    dE = div_term<T, EXACT>(dE, FR, FLm[j], dm, p.ax.rdx[5][m]);
                    FLm[j] = FR;
                }
                const T lam = HOMOG ? p.hom.lam[q2] : p.lam[cell];
                const T mf = e0 * lam;
                dE = dE - mf;
                if constexpr (EXACT) p.Eout[cell] = e0 + p.dt * dE;
                else p.Eout[cell] = e0 + (T)p.dt * dE;
            }
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            wm1[j] = w0[j];
            w0[j] = wp1[j];
            wp1[j] = wp2[j];
            wp2[j] = wnx[j];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) zq[q] = znx[q];
    }
}
