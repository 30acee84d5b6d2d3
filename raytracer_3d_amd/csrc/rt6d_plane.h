// rt6d_plane.h -- the plane step kernel with face-flux exchange (included by
// rt6d_kernels.h after rt6d_column.h, inside namespace rt6d).
//
// One block owns ZC consecutive (x,y) planes, z = z0 .. z0+ZC-1, of one (k,l)
// position (one thread per (x,y) column, nx*ny <= 1024) and marches them
// through m = 1..nm like the column kernel (register window along m, the
// m-face flux carried, inner z faces shared by the ZC cells).  Along x and y
// every thread computes only the flux through the upper face of its cell and
// publishes it in LDS; after one barrier it takes the flux through its lower
// face from the neighbouring thread.  Each x and y face is thus computed once
// instead of twice; the lower boundary face (x = 1 or y = 1) has the closed
// form v*E(1) of the zero-gradient ghosts.  Two LDS buffers alternate with m
// (the buffer written at plane m was last read at plane m-2, before the
// barrier of plane m-1), so one barrier per plane suffices.
// Per face the arithmetic is that of k_step.
#pragma once

struct PlaneCfg {
    FastDiv dzb;   // division by the number of z blocks
    int nzb;       // nz / ZC
};

template <typename T, bool EXACT, bool HOMOG, int ZC>
__global__ void __launch_bounds__(1024) k_step_plane(const Geo g, const StepPtrs<T> p, const PlaneCfg pc,
                                                     uint32_t nb_logical)
{
    extern __shared__ unsigned char smem_raw[];
    const int NT = blockDim.x;
    T *sF = reinterpret_cast<T *>(smem_raw);   // [2 buffers][2 axes][ZC][NT]

    const uint32_t b = xcd_logical_block(blockIdx.x, nb_logical);
    if (b >= nb_logical) return;  // uniform over the block
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2], nk = g.n[3], nl = g.n[4], nm = g.n[5];
    const int nxy = nx * ny;
    const uint32_t bkl = fdiv(b, pc.dzb);
    const int bzb = (int)(b - bkl * (uint32_t)pc.nzb);
    const int bl = (int)fdiv(bkl, g.div[3]);
    const int bk = (int)bkl - bl * nk;
    const int z0 = bzb * ZC + 1, k = bk + 1, l = bl + 1;

    const int tid = threadIdx.x;
    const bool live = tid < nxy;
    const int xy = live ? tid : nxy - 1;   // padding threads shadow the last cell and never store
    const int y0 = (int)fdiv((uint32_t)xy, g.div[0]);
    const int x0 = xy - y0 * nx;
    const int x = x0 + 1, y = y0 + 1;
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
    const bool ax = g.act & 1, ay = g.act & 2, az = g.act & 4;
    const bool ak = !HOMOG && (g.act & 8), al = !HOMOG && (g.act & 16), am = g.act & 32;
    const double dtv = p.dt;
    const uint32_t oc = (uint32_t)xy * S;

    // byte offsets of x-1, x+1, x+2 and y-1, y+1, y+2 (clamped = zero-gradient ghosts)
    uint32_t ox[3], oy[3];
    {
        const int xs[3] = {x - 1 < 1 ? 1 : x - 1, x + 1 > nx ? nx : x + 1, x + 2 > nx ? nx : x + 2};
        const int ys[3] = {y - 1 < 1 ? 1 : y - 1, y + 1 > ny ? ny : y + 1, y + 2 > ny ? ny : y + 2};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            ox[q] = (uint32_t)((xs[q] - 1) + nx * y0) * S;
            oy[q] = (uint32_t)(x0 + nx * (ys[q] - 1)) * S;
        }
    }
    // upper-face coefficients of x and y; lower-face fluxes come from the neighbour
    double cxr = 0, dxr = 1, rxx = 0, cyr = 0, dyr = 1, ryy = 0;
    if (ax) { cxr = p.ax.cfl[0][x]; dxr = p.ax.dx[0][x]; rxx = p.ax.rdx[0][x]; }
    if (ay) { cyr = p.ax.cfl[1][y]; dyr = p.ax.dx[1][y]; ryy = p.ax.rdx[1][y]; }
    double ckl = 0, ckr = 0, rkk = 0, dkl = 1, dkr = 1, cll = 0, clr = 0, rll = 0, dll = 1, dlr = 1;
    if (ak) { ckl = sld(p.ax.cfl[3], k - 1); ckr = sld(p.ax.cfl[3], k); rkk = sld(p.ax.rdx[3], k); dkl = sld(p.ax.dx[3], k - 1); dkr = sld(p.ax.dx[3], k); }
    if (al) { cll = sld(p.ax.cfl[4], l - 1); clr = sld(p.ax.cfl[4], l); rll = sld(p.ax.rdx[4], l); dll = sld(p.ax.dx[4], l - 1); dlr = sld(p.ax.dx[4], l); }

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
    int kpl[4] = {0, 0, 0, 0}, lpl[4] = {0, 0, 0, 0};
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
    auto poff = [&](int j, int mm) -> uint32_t { return base0 + (uint32_t)j * planeB + (uint32_t)(mm - 1) * mB; };
    auto ldm = [&](int j, int mm) -> T {
        mm = mm < 1 ? 1 : (mm > nm ? nm : mm);
        return bld<T>(rE, oc, poff(j, mm));
    };
    auto ldz = [&](int q, int mm) -> T {
        const int mu = zmir[q] ? nm + 1 - mm : mm;
        return bld<T>(rE, oc, base0 + (uint32_t)(zsh[q] * (int)planeB) + (uint32_t)(mu - 1) * mB);
    };
    T wm1[ZC], w0[ZC], wp1[ZC], wp2[ZC], FLm[ZC];
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
        w0[j] = ldm(j, 1);
        wp1[j] = ldm(j, 2);
        wp2[j] = ldm(j, 3);
        wm1[j] = w0[j];
        FLm[j] = T(0);
    }
    T zq[4] = {T(0), T(0), T(0), T(0)};
    if (az) {
#pragma unroll
        for (int q = 0; q < 4; ++q) zq[q] = ldz(q, 1);
    }
    __amdgpu_buffer_rsrc_t rV[kAxes], rL;
    if (!HOMOG) {
#pragma unroll
        for (int d = 0; d < kAxes; ++d)
            rV[d] = mk_rsrc(p.vel[d], (uint64_t)g.ncells / g.n[d] * (g.n[d] + 1) * S);
        rL = mk_rsrc(p.lam, cellsB);
    }
    const uint32_t fxo = (uint32_t)(x0 + (nx + 1) * y0) * S;   // x faces: lower face x-1
    const uint32_t fyo = (uint32_t)(x0 + nx * y0) * S;         // y faces: lower face y-1
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
        T wnx[ZC], znx[4];
#pragma unroll
        for (int j = 0; j < ZC; ++j) wnx[j] = ldm(j, m + 3);
#pragma unroll
        for (int q = 0; q < 4; ++q) znx[q] = (az && m < nm) ? ldz(q, m + 1) : T(0);

        // upper x and y faces of the ZC cells, published for the neighbours
        T *sb = sF + (size_t)(m & 1) * 2 * ZC * NT;
        T vx0[ZC], vy0[ZC];
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const uint32_t Pm = P0 + (uint32_t)j + (uint32_t)(m - 1) * nzkl;
            const uint32_t pb = Pm * planeB;
            const uint32_t q2 = q20 + (uint32_t)j;
            const T e0 = w0[j];
            vx0[j] = T(0);
            vy0[j] = T(0);
            if (ax) {
                const T a1 = bld<T>(rE, ox[0], pb), a3 = bld<T>(rE, ox[1], pb), a4 = bld<T>(rE, ox[2], pb);
                T vR;
                if (HOMOG) {
                    const T vi = sld(p.hom.vx3, p.hom.n4 + q2), vb = sld(p.hom.vx3, 2 * p.hom.n4 + q2);
                    vx0[j] = sld(p.hom.vx3, q2);
                    vR = x == nx ? vb : vi;
                } else {
                    vx0[j] = bld<T>(rV[0], fxo, Pm * fxpB);
                    vR = bld<T>(rV[0], fxo + S, Pm * fxpB);
                }
                sb[j * NT + tid] = face_flux<T, EXACT>(vR, dtv, dxr, cxr, a1, e0, a3, a4);
            }
            if (ay) {
                const T a1 = bld<T>(rE, oy[0], pb), a3 = bld<T>(rE, oy[1], pb), a4 = bld<T>(rE, oy[2], pb);
                T vR;
                if (HOMOG) {
                    const T vi = sld(p.hom.vy3, p.hom.n4 + q2), vb = sld(p.hom.vy3, 2 * p.hom.n4 + q2);
                    vy0[j] = sld(p.hom.vy3, q2);
                    vR = y == ny ? vb : vi;
                } else {
                    vy0[j] = bld<T>(rV[1], fyo, Pm * fypB);
                    vR = bld<T>(rV[1], fyo + (uint32_t)nx * S, Pm * fypB);
                }
                sb[(ZC + j) * NT + tid] = face_flux<T, EXACT>(vR, dtv, dyr, cyr, a1, e0, a3, a4);
            }
        }
        __syncthreads();

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
                Fz[i] = face_flux<T, EXACT>(v, dtv, sld(p.ax.dx[2], fz), sld(p.ax.cfl[2], fz), zs[i], zs[i + 1],
                                            zs[i + 2], zs[i + 3]);
            }
        }
#pragma unroll
        for (int j = 0; j < ZC; ++j) {
            const uint32_t Pm = P0 + (uint32_t)j + (uint32_t)(m - 1) * nzkl;
            const uint32_t pb = Pm * planeB;
            const uint32_t q2 = q20 + (uint32_t)j;
            const T e0 = w0[j];
            T dE = T(0);
            if (ax) {
                const T FR = sb[j * NT + tid];
                const T FN = sb[j * NT + (x == 1 ? tid : tid - 1)];   // in bounds for every thread
                const T FL = x == 1 ? vx0[j] * e0 : FN;
                dE = div_term<T, EXACT>(dE, FR, FL, dxr, rxx);
            }
            if (ay) {
                const T FR = sb[(ZC + j) * NT + tid];
                const T FN = sb[(ZC + j) * NT + (y == 1 ? tid : tid - nx)];
                const T FL = y == 1 ? vy0[j] * e0 : FN;
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
            if (live) bst<T>(enew, rO, oc, pb);
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
