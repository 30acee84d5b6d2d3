/*
 * rt6d_oracle.c -- CPU restatement of the reference 6D wave-energy time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker against which the HIP
 * product path is compared.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls it, and nothing here is a fallback for the GPU path.
 *
 * It follows the reference Fortran loop by loop on the reference's own data
 * model: every 6D field is a halo'd array over (1-onx:n+onx) per axis, onx=2,
 * Fortran order (x fastest), zero-initialised as in
 * for_src/main_module.f90:131-238.  It is compiled with -ffp-contract=off so
 * that every floating-point operation is rounded as the Fortran statements
 * are written (left-to-right evaluation, no fused multiply-add).
 *
 * Function -> reference lines it restates:
 *   or_calc_grid        for_src/calc_grid.f90:104-177 and u_centered_grid :300-322
 *   or_calc_parameter   for_src/calc_parameter.f90:2-264
 *   or_fill_ghosts      for_src/integrate.f90:14-94 and the single-rank branches
 *                       of reflection_condition, for_src/parallel_mpi.f90:837-841,867-871
 *   or_superbee_fluxes  for_src/superbee_fluxes.f90:4-174
 *   or_integrate        for_src/integrate.f90:4-271
 *   or_diagnose         for_src/diagnose.f90:45-329 (the reductions; no netCDF)
 *
 * Pinning (what checks this file):
 *   - golden vectors written by the reference Fortran compiled in this
 *     container (oracle/build_ref.sh, tests/golden/make_golden.py): grid,
 *     calc_parameter outputs, fluxes, dE, meanfl and E after 1 and N steps;
 *   - closed forms and invariants in tests/test_oracle_invariants.py
 *     (constant E gives zero divergence, linear ramp gives Lax-Wendroff,
 *     extrema give upwind, telescoping energy budget, uniform-grid positions,
 *     dispersion relation).
 *   - or_diagnose: the reference diagnose.f90 cannot be built here (it needs
 *     netCDF, absent from the image).  Its sums are pinned against math.fsum
 *     of the reference-dumped dE/meanfl/flux fields and by the budget
 *     residual identity; the exact sequential rounding of the reference's
 *     sums is "parity unpinned".
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ONX 2

typedef struct or_state {
    int n[6];          /* nx ny nz nk nl nm */
    int H[6];          /* n + 2*onx */
    int reflect_up, reflect_down;
    double f0, sign_omega;
    double k_fixed, l_fixed, m_fixed;
    size_t nh6;        /* number of halo'd 6D points */
    size_t nh3;        /* number of halo'd 3D points */
    /* 1D grids, halo'd: index i+1 for Fortran index i in [-1, n+2] */
    double *t[6], *dt_[6], *u[6], *du[6];
    /* 3D background, halo'd */
    double *Nsqr, *U, *V;
    /* 6D fields, halo'd */
    double *E, *dE, *meanfl;
    double *omega_i, *omega_e;
    double *vel[6];    /* xdot ydot zdot kdot ldot mdot */
    double *flux[6];   /* flux_x .. flux_m */
    double *lam[3];    /* lambda_x lambda_y lambda_z */
} or_state;

/* Fortran index (i,j,k,l,n,m), each in [1-onx, n+onx], to linear offset. */
static inline size_t H6(const or_state *s, int i, int j, int k, int l, int n, int m)
{
    return (size_t)(i + 1) +
           (size_t)s->H[0] * ((size_t)(j + 1) +
           (size_t)s->H[1] * ((size_t)(k + 1) +
           (size_t)s->H[2] * ((size_t)(l + 1) +
           (size_t)s->H[3] * ((size_t)(n + 1) +
           (size_t)s->H[4] * (size_t)(m + 1)))));
}
static inline size_t H3(const or_state *s, int i, int j, int k)
{
    return (size_t)(i + 1) + (size_t)s->H[0] * ((size_t)(j + 1) + (size_t)s->H[1] * (size_t)(k + 1));
}
#define G1(a, i) ((a)[(i) + 1])

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

or_state *or_create(const int *n, int reflect_up, int reflect_down)
{
    or_state *s = (or_state *)calloc(1, sizeof(or_state));
    if (!s) return NULL;
    s->nh6 = 1;
    for (int d = 0; d < 6; ++d) {
        s->n[d] = n[d];
        s->H[d] = n[d] + 2 * ONX;
        s->nh6 *= (size_t)s->H[d];
        s->t[d] = zalloc(s->H[d]);
        s->dt_[d] = zalloc(s->H[d]);
        s->u[d] = zalloc(s->H[d]);
        s->du[d] = zalloc(s->H[d]);
    }
    s->nh3 = (size_t)s->H[0] * s->H[1] * s->H[2];
    s->reflect_up = reflect_up;
    s->reflect_down = reflect_down;
    s->f0 = 1e-4;
    s->sign_omega = 1.0;
    s->Nsqr = zalloc(s->nh3);
    s->U = zalloc(s->nh3);
    s->V = zalloc(s->nh3);
    s->E = zalloc(s->nh6);
    s->dE = zalloc(s->nh6);
    s->meanfl = zalloc(s->nh6);
    s->omega_i = zalloc(s->nh6);
    s->omega_e = zalloc(s->nh6);
    for (int d = 0; d < 6; ++d) {
        s->vel[d] = zalloc(s->nh6);
        s->flux[d] = zalloc(s->nh6);
    }
    for (int d = 0; d < 3; ++d) s->lam[d] = zalloc(s->nh6);
    return s;
}

void or_destroy(or_state *s)
{
    if (!s) return;
    for (int d = 0; d < 6; ++d) {
        free(s->t[d]); free(s->dt_[d]); free(s->u[d]); free(s->du[d]);
        free(s->vel[d]); free(s->flux[d]);
    }
    for (int d = 0; d < 3; ++d) free(s->lam[d]);
    free(s->Nsqr); free(s->U); free(s->V);
    free(s->E); free(s->dE); free(s->meanfl);
    free(s->omega_i); free(s->omega_e);
    free(s);
}

void or_set_params(or_state *s, double f0, double sign_omega,
                   double k_fixed, double l_fixed, double m_fixed)
{
    s->f0 = f0;
    s->sign_omega = sign_omega;
    s->k_fixed = k_fixed;
    s->l_fixed = l_fixed;
    s->m_fixed = m_fixed;
}

/* ------------------------------------------------------------------ grid */

/* u_centered_grid, calc_grid.f90:300-322, on a 1-based array of length N. */
static void u_centered_grid(const double *dyt, double *dyu, double *yt, double *yu, int N)
{
    /* local 1-based indexing: element p is [p-1] */
    yu[0] = 0.0;
    for (int i = 2; i <= N; ++i) yu[i - 1] = yu[i - 2] + dyt[i - 1];
    yt[0] = yu[0] - dyt[0] * 0.5;
    for (int i = 2; i <= N; ++i) yt[i - 1] = 2.0 * yu[i - 2] - yt[i - 2];
    for (int i = 1; i <= N - 1; ++i) dyu[i - 1] = yt[i] - yt[i - 1];
    int im = (N - 1 > 1) ? N - 1 : 1;
    dyu[N - 1] = 2.0 * dyt[N - 1] - dyu[im - 1];
}

/* Interior T-cell spacings d[0..n-1] for axis `axis`.  calc_grid.f90:104-177. */
void or_set_spacing(or_state *s, int axis, const double *d)
{
    int n = s->n[axis];
    double *a = s->dt_[axis];
    for (int i = 1; i <= n; ++i) G1(a, i) = d[i - 1];
}

void or_calc_grid(or_state *s)
{
    for (int d = 0; d < 6; ++d) {
        int n = s->n[d];
        double *dt = s->dt_[d], *du = s->du[d], *t = s->t[d], *u = s->u[d];
        /* fill boundary values, calc_grid.f90:104-111 */
        for (int i = 1; i <= ONX; ++i) {
            G1(dt, n + i) = G1(dt, n);
            G1(dt, 1 - i) = G1(dt, 1);
        }
        u_centered_grid(dt, du, t, u, n + 2 * ONX);
        int N = n + 2 * ONX;
        if (d == 0 || d == 1) {
            /* x and y: shift by dxu(1), calc_grid.f90:115,124 */
            double sh = G1(du, 1);
            for (int p = 0; p < N; ++p) { t[p] = t[p] - sh; }
            for (int p = 0; p < N; ++p) { u[p] = u[p] - sh; }
        } else if (d == 2) {
            /* z: zu(nz) becomes 0, calc_grid.f90:133 */
            double sh = G1(u, n);
            for (int p = 0; p < N; ++p) { t[p] = t[p] - sh; }
            for (int p = 0; p < N; ++p) { u[p] = u[p] - sh; }
        } else {
            /* k, l, m: u(n/2) becomes 0, or the fixed wavenumber, :142-172 */
            double fixed = d == 3 ? s->k_fixed : (d == 4 ? s->l_fixed : s->m_fixed);
            if (n > 1) {
                double sh = G1(u, n / 2);
                for (int p = 0; p < N; ++p) { t[p] = t[p] - sh; }
                for (int p = 0; p < N; ++p) { u[p] = u[p] - sh; }
            } else {
                for (int p = 0; p < N; ++p) { t[p] = fixed; u[p] = fixed; }
            }
        }
    }
}

void or_get_grid(const or_state *s, int axis, double *t, double *dt, double *u, double *du)
{
    size_t nb = (size_t)s->H[axis] * sizeof(double);
    if (t) memcpy(t, s->t[axis], nb);
    if (dt) memcpy(dt, s->dt_[axis], nb);
    if (u) memcpy(u, s->u[axis], nb);
    if (du) memcpy(du, s->du[axis], nb);
}

/* Overwrite the 1D grids (halo'd, length n+4 each) with host-given arrays. */
void or_set_grid(or_state *s, int axis, const double *t, const double *dt)
{
    size_t nb = (size_t)s->H[axis] * sizeof(double);
    memcpy(s->t[axis], t, nb);
    memcpy(s->dt_[axis], dt, nb);
}

/* ------------------------------------------------------------ background */

void or_set_background(or_state *s, const double *Nsqr, const double *U, const double *V)
{
    memcpy(s->Nsqr, Nsqr, s->nh3 * sizeof(double));
    memcpy(s->U, U, s->nh3 * sizeof(double));
    memcpy(s->V, V, s->nh3 * sizeof(double));
}

/* --------------------------------------------------------- calc_parameter */

#define LOOP6H(s)                                                   \
    for (int m = 1 - ONX; m <= s->n[5] + ONX; ++m)                  \
    for (int n = 1 - ONX; n <= s->n[4] + ONX; ++n)                  \
    for (int l = 1 - ONX; l <= s->n[3] + ONX; ++l)                  \
    for (int k = 1 - ONX; k <= s->n[2] + ONX; ++k)                  \
    for (int j = 1 - ONX; j <= s->n[1] + ONX; ++j)                  \
    for (int i = 1 - ONX; i <= s->n[0] + ONX; ++i)

#define LOOP6I(s)                                                   \
    for (int m = 1; m <= s->n[5]; ++m)                              \
    for (int n = 1; n <= s->n[4]; ++n)                              \
    for (int l = 1; l <= s->n[3]; ++l)                              \
    for (int k = 1; k <= s->n[2]; ++k)                              \
    for (int j = 1; j <= s->n[1]; ++j)                              \
    for (int i = 1; i <= s->n[0]; ++i)

/* Shift of one unit along axis d in the halo'd linear layout. */
static size_t stride6(const or_state *s, int d)
{
    size_t st = 1;
    for (int e = 0; e < d; ++e) st *= (size_t)s->H[e];
    return st;
}

/* Coordinate of linear halo offset p along axis d, as a Fortran index. */
static inline int coord6(const or_state *s, size_t p, int d)
{
    for (int e = 0; e < d; ++e) p /= (size_t)s->H[e];
    return (int)(p % (size_t)s->H[d]) - 1;
}

void or_calc_parameter(or_state *s)
{
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2], nk = s->n[3], nl = s->n[4], nm = s->n[5];
    const double f0 = s->f0;
    double *aloc = zalloc(s->nh6);
    const double *kt = s->t[3], *lt = s->t[4], *mt = s->t[5];

    /* omega_i, omega_e on all halo points, calc_parameter.f90:13-27 */
    LOOP6H(s) {
        double kh2 = G1(lt, n) * G1(lt, n) + G1(kt, l) * G1(kt, l);
        size_t p = H6(s, i, j, k, l, n, m);
        size_t q = H3(s, i, j, k);
        double mm = G1(mt, m) * G1(mt, m);
        s->omega_i[p] = s->sign_omega * sqrt((s->Nsqr[q] * kh2 + (f0 * f0) * mm) / (mm + kh2));
        s->omega_e[p] = s->omega_i[p] + G1(kt, l) * s->U[q] + G1(lt, n) * s->V[q];
    }

    const size_t sx = stride6(s, 0), sy = stride6(s, 1), sz = stride6(s, 2);
    const size_t sk = stride6(s, 3), sl = stride6(s, 4), sm = stride6(s, 5);
    const double *oe = s->omega_e;
    size_t N6 = s->nh6;

    /* ------------------------------------------------------------ x / k */
    if (nx > 1 && nk > 1) {
        /* x average (later overwritten by the k average, except its last k slice) :34-36 */
        for (size_t p = 0; p < N6; ++p) {
            int ci = coord6(s, p, 0);
            if (ci <= nx + ONX - 1) aloc[p] = 0.5 * (oe[p] + oe[p + sx]);
        }
        /* k average :38-40 */
        for (size_t p = 0; p < N6; ++p) {
            int cl = coord6(s, p, 3);
            if (cl <= nk + ONX - 1) aloc[p] = 0.5 * (oe[p] + oe[p + sk]);
        }
        /* xdot = d omega / dk :44-46 */
        for (size_t p = 0; p < N6; ++p) {
            int cl = coord6(s, p, 3);
            if (cl >= 1 - ONX + 1) s->vel[0][p] = (aloc[p] - aloc[p - sk]) / G1(s->dt_[3], cl);
        }
        /* clip at x boundaries :55-66 */
        for (size_t p = 0; p < N6; ++p) {
            int ci = coord6(s, p, 0);
            if (ci == nx) { double v = s->vel[0][p]; s->vel[0][p] = v > 0.0 ? v : 0.0; }
        }
        for (size_t p = 0; p < N6; ++p) {
            int ci = coord6(s, p, 0);
            if (ci > nx) s->vel[0][p] = s->vel[0][p - (size_t)(ci - nx) * sx];
        }
        for (size_t p = 0; p < N6; ++p) {
            int ci = coord6(s, p, 0);
            if (ci == 0) { double mv = -s->vel[0][p]; s->vel[0][p] = -(mv > 0.0 ? mv : 0.0); }
        }
        for (size_t p = 0; p < N6; ++p) {
            int ci = coord6(s, p, 0);
            if (ci < 0) s->vel[0][p] = s->vel[0][p + (size_t)(0 - ci) * sx];
        }
        /* kdot = -d omega / dx :69-71 */
        for (size_t p = 0; p < N6; ++p) {
            int ci = coord6(s, p, 0);
            if (ci >= 1 - ONX + 1) s->vel[3][p] = -(aloc[p] - aloc[p - sx]) / G1(s->dt_[0], ci);
        }
        /* clip at k boundaries :76-87 */
        for (size_t p = 0; p < N6; ++p) {
            int cl = coord6(s, p, 3);
            if (cl == nk) { double v = s->vel[3][p]; s->vel[3][p] = v > 0.0 ? v : 0.0; }
        }
        for (size_t p = 0; p < N6; ++p) {
            int cl = coord6(s, p, 3);
            if (cl > nk) s->vel[3][p] = s->vel[3][p - (size_t)(cl - nk) * sk];
        }
        for (size_t p = 0; p < N6; ++p) {
            int cl = coord6(s, p, 3);
            if (cl == 0) { double mv = -s->vel[3][p]; s->vel[3][p] = -(mv > 0.0 ? mv : 0.0); }
        }
        for (size_t p = 0; p < N6; ++p) {
            int cl = coord6(s, p, 3);
            if (cl < 0) s->vel[3][p] = s->vel[3][p + (size_t)(0 - cl) * sk];
        }
    } else {
        /* analytic group velocity on the interior, :91-104 */
        LOOP6I(s) {
            double kh2 = G1(lt, n) * G1(lt, n) + G1(kt, l) * G1(kt, l);
            size_t p = H6(s, i, j, k, l, n, m);
            size_t q = H3(s, i, j, k);
            double mm = G1(mt, m) * G1(mt, m);
            double den = mm + kh2;
            s->vel[0][p] = s->U[q] + (s->Nsqr[q] - f0 * f0) / s->omega_i[p] * mm / (den * den) * G1(kt, l);
        }
    }

    /* ------------------------------------------------------------ y / l */
    if (ny > 1 && nl > 1) {
        for (size_t p = 0; p < N6; ++p) {
            int cj = coord6(s, p, 1);
            if (cj <= ny + ONX - 1) aloc[p] = 0.5 * (oe[p] + oe[p + sy]);
        }
        for (size_t p = 0; p < N6; ++p) {
            int cn = coord6(s, p, 4);
            if (cn <= nl + ONX - 1) aloc[p] = 0.5 * (oe[p] + oe[p + sl]);
        }
        for (size_t p = 0; p < N6; ++p) {
            int cn = coord6(s, p, 4);
            if (cn >= 1 - ONX + 1) s->vel[1][p] = (aloc[p] - aloc[p - sl]) / G1(s->dt_[4], cn);
        }
        for (size_t p = 0; p < N6; ++p) {
            int cj = coord6(s, p, 1);
            if (cj == ny) { double v = s->vel[1][p]; s->vel[1][p] = v > 0.0 ? v : 0.0; }
        }
        for (size_t p = 0; p < N6; ++p) {
            int cj = coord6(s, p, 1);
            if (cj > ny) s->vel[1][p] = s->vel[1][p - (size_t)(cj - ny) * sy];
        }
        for (size_t p = 0; p < N6; ++p) {
            int cj = coord6(s, p, 1);
            if (cj == 0) { double mv = -s->vel[1][p]; s->vel[1][p] = -(mv > 0.0 ? mv : 0.0); }
        }
        for (size_t p = 0; p < N6; ++p) {
            int cj = coord6(s, p, 1);
            if (cj < 0) s->vel[1][p] = s->vel[1][p + (size_t)(0 - cj) * sy];
        }
        for (size_t p = 0; p < N6; ++p) {
            int cj = coord6(s, p, 1);
            if (cj >= 1 - ONX + 1) s->vel[4][p] = -(aloc[p] - aloc[p - sy]) / G1(s->dt_[1], cj);
        }
        for (size_t p = 0; p < N6; ++p) {
            int cn = coord6(s, p, 4);
            if (cn == nl) { double v = s->vel[4][p]; s->vel[4][p] = v > 0.0 ? v : 0.0; }
        }
        for (size_t p = 0; p < N6; ++p) {
            int cn = coord6(s, p, 4);
            if (cn > nl) s->vel[4][p] = s->vel[4][p - (size_t)(cn - nl) * sl];
        }
        for (size_t p = 0; p < N6; ++p) {
            int cn = coord6(s, p, 4);
            if (cn == 0) { double mv = -s->vel[4][p]; s->vel[4][p] = -(mv > 0.0 ? mv : 0.0); }
        }
        for (size_t p = 0; p < N6; ++p) {
            int cn = coord6(s, p, 4);
            if (cn < 0) s->vel[4][p] = s->vel[4][p + (size_t)(0 - cn) * sl];
        }
    } else {
        /* :169-182 */
        LOOP6I(s) {
            double kh2 = G1(lt, n) * G1(lt, n) + G1(kt, l) * G1(kt, l);
            size_t p = H6(s, i, j, k, l, n, m);
            size_t q = H3(s, i, j, k);
            double mm = G1(mt, m) * G1(mt, m);
            double den = mm + kh2;
            s->vel[1][p] = s->V[q] + (s->Nsqr[q] - f0 * f0) / s->omega_i[p] * mm / (den * den) * G1(lt, n);
        }
    }

    /* ------------------------------------------------------------ z / m (always FD) */
    for (size_t p = 0; p < N6; ++p) {
        int ck = coord6(s, p, 2);
        if (ck <= nz + ONX - 1) aloc[p] = 0.5 * (oe[p] + oe[p + sz]);
    }
    for (size_t p = 0; p < N6; ++p) {
        int cm = coord6(s, p, 5);
        if (cm <= nm + ONX - 1) aloc[p] = 0.5 * (oe[p] + oe[p + sm]);
    }
    /* zdot = d omega / dm :198-200 */
    for (size_t p = 0; p < N6; ++p) {
        int cm = coord6(s, p, 5);
        if (cm >= 1 - ONX + 1) s->vel[2][p] = (aloc[p] - aloc[p - sm]) / G1(s->dt_[5], cm);
    }
    /* :204-216 */
    if (!s->reflect_up) {
        for (size_t p = 0; p < N6; ++p) {
            int ck = coord6(s, p, 2);
            if (ck == nz) { double v = s->vel[2][p]; s->vel[2][p] = v > 0.0 ? v : 0.0; }
        }
    }
    for (size_t p = 0; p < N6; ++p) {
        int ck = coord6(s, p, 2);
        if (ck > nz) s->vel[2][p] = s->vel[2][p - (size_t)(ck - nz) * sz];
    }
    if (!s->reflect_down) {
        for (size_t p = 0; p < N6; ++p) {
            int ck = coord6(s, p, 2);
            if (ck == 0) { double mv = -s->vel[2][p]; s->vel[2][p] = -(mv > 0.0 ? mv : 0.0); }
        }
    }
    for (size_t p = 0; p < N6; ++p) {
        int ck = coord6(s, p, 2);
        if (ck < 0) s->vel[2][p] = s->vel[2][p + (size_t)(0 - ck) * sz];
    }
    /* mdot = -d omega / dz :219-221 */
    for (size_t p = 0; p < N6; ++p) {
        int ck = coord6(s, p, 2);
        if (ck >= 1 - ONX + 1) s->vel[5][p] = -(aloc[p] - aloc[p - sz]) / G1(s->dt_[2], ck);
    }
    /* :226-237 */
    for (size_t p = 0; p < N6; ++p) {
        int cm = coord6(s, p, 5);
        if (cm == nm) { double v = s->vel[5][p]; s->vel[5][p] = v > 0.0 ? v : 0.0; }
    }
    for (size_t p = 0; p < N6; ++p) {
        int cm = coord6(s, p, 5);
        if (cm > nm) s->vel[5][p] = s->vel[5][p - (size_t)(cm - nm) * sm];
    }
    for (size_t p = 0; p < N6; ++p) {
        int cm = coord6(s, p, 5);
        if (cm == 0) { double mv = -s->vel[5][p]; s->vel[5][p] = -(mv > 0.0 ? mv : 0.0); }
    }
    for (size_t p = 0; p < N6; ++p) {
        int cm = coord6(s, p, 5);
        if (cm < 0) s->vel[5][p] = s->vel[5][p + (size_t)(0 - cm) * sm];
    }

    /* growth rates for the mean-flow interaction, :240-260 */
    const double *dxt = s->dt_[0], *dyt = s->dt_[1], *dzt = s->dt_[2];
    LOOP6I(s) {
        size_t p = H6(s, i, j, k, l, n, m);
        size_t q = H3(s, i, j, k);
        double rw = 1.0 / s->omega_i[p];
        double xd = s->vel[0][p], yd = s->vel[1][p], zd = s->vel[2][p];
        double Uq = s->U[q], Vq = s->V[q];
        s->lam[0][p] = rw * G1(kt, l) *
            ((xd - Uq) * (s->U[H3(s, i + 1, j, k)] - s->U[H3(s, i - 1, j, k)]) / (2.0 * G1(dxt, i))
           + (yd - Vq) * (s->V[H3(s, i + 1, j, k)] - s->V[H3(s, i - 1, j, k)]) / (2.0 * G1(dxt, i)));
        s->lam[1][p] = rw * G1(lt, n) *
            ((xd - Uq) * (s->U[H3(s, i, j + 1, k)] - s->U[H3(s, i, j - 1, k)]) / (2.0 * G1(dyt, j))
           + (yd - Vq) * (s->V[H3(s, i, j + 1, k)] - s->V[H3(s, i, j - 1, k)]) / (2.0 * G1(dyt, j)));
        s->lam[2][p] = rw * zd *
            (G1(kt, l) * (s->U[H3(s, i, j, k + 1)] - s->U[H3(s, i, j, k - 1)]) / (2.0 * G1(dzt, k))
           + G1(lt, n) * (s->V[H3(s, i, j, k + 1)] - s->V[H3(s, i, j, k - 1)]) / (2.0 * G1(dzt, k)));
    }
    free(aloc);
}

/* ------------------------------------------------------------ ghost fill */

void or_fill_ghosts(or_state *s)
{
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2], nk = s->n[3], nl = s->n[4], nm = s->n[5];
    double *E = s->E;
    /* z reflection: m-mirrored copy, parallel_mpi.f90:837-841 and :867-871.
     * Only the interior of x, y, k, l is written; all interior m. */
    if (s->reflect_up) {
        for (int ii = 1; ii <= ONX; ++ii)
            for (int m = 1; m <= nm; ++m)
            for (int n = 1; n <= nl; ++n)
            for (int l = 1; l <= nk; ++l)
            for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                E[H6(s, i, j, nz + ii, l, n, m)] = E[H6(s, i, j, nz - ii + 1, l, n, nm + 1 - m)];
    }
    if (s->reflect_down) {
        for (int ii = 1; ii <= ONX; ++ii)
            for (int m = 1; m <= nm; ++m)
            for (int n = 1; n <= nl; ++n)
            for (int l = 1; l <= nk; ++l)
            for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                E[H6(s, i, j, 1 - ii, l, n, m)] = E[H6(s, i, j, ii, l, n, nm + 1 - m)];
    }
    /* zero-gradient copies, integrate.f90:24-94; each covers only the
     * interior of the other five axes */
#define ZG(AX, NAX, ENABLE)                                                          \
    if (ENABLE) {                                                                    \
        for (int ii = 1; ii <= ONX; ++ii)                                            \
            LOOP6I(s) {                                                              \
                int c[6] = {i, j, k, l, n, m};                                       \
                if (c[AX] != NAX) continue;                                          \
                int hi[6] = {i, j, k, l, n, m}; hi[AX] = NAX + ii;                   \
                E[H6(s, hi[0], hi[1], hi[2], hi[3], hi[4], hi[5])] = E[H6(s, i, j, k, l, n, m)]; \
            }                                                                        \
    }
#define ZGL(AX, ENABLE)                                                              \
    if (ENABLE) {                                                                    \
        for (int ii = 1; ii <= ONX; ++ii)                                            \
            LOOP6I(s) {                                                              \
                int c[6] = {i, j, k, l, n, m};                                       \
                if (c[AX] != 1) continue;                                            \
                int lo[6] = {i, j, k, l, n, m}; lo[AX] = 1 - ii;                     \
                E[H6(s, lo[0], lo[1], lo[2], lo[3], lo[4], lo[5])] = E[H6(s, i, j, k, l, n, m)]; \
            }                                                                        \
    }
    ZG(0, nx, nx > 1) ZGL(0, nx > 1)
    ZG(1, ny, ny > 1) ZGL(1, ny > 1)
    ZG(2, nz, !s->reflect_up) ZGL(2, !s->reflect_down)
    ZG(3, nk, nk > 1) ZGL(3, nk > 1)
    ZG(4, nl, nl > 1) ZGL(4, nl > 1)
    ZG(5, nm, nm > 1) ZGL(5, nm > 1)
#undef ZG
#undef ZGL
}

/* ------------------------------------------------------------ superbee */

static inline double limiter(double cr)
{
    /* max(0, max(min(1, 2 cr), min(2, cr))), superbee_fluxes.f90:10 */
    double a = 2.0 * cr;
    a = a < 1.0 ? a : 1.0;
    double b = cr < 2.0 ? cr : 2.0;
    double c = a > b ? a : b;
    return c > 0.0 ? c : 0.0;
}

/* single-precision literal 1.E20 of superbee_fluxes.f90:26 */
#define BIG_R4 ((double)1.0e20f)

void or_superbee_fluxes(or_state *s, double dt)
{
    const double *E = s->E;
    for (int d = 0; d < 6; ++d) {
        if (s->n[d] <= 1) continue;
        size_t st = stride6(s, d);
        const double *vel = s->vel[d];
        const double *dd = s->dt_[d];
        double *F = s->flux[d];
        int lo[6] = {1, 1, 1, 1, 1, 1};
        lo[d] = 0; /* faces xs_pe-1 .. xe_pe */
        for (int m = lo[5]; m <= s->n[5]; ++m)
        for (int n = lo[4]; n <= s->n[4]; ++n)
        for (int l = lo[3]; l <= s->n[3]; ++l)
        for (int k = lo[2]; k <= s->n[2]; ++k)
        for (int j = lo[1]; j <= s->n[1]; ++j)
        for (int i = lo[0]; i <= s->n[0]; ++i) {
            size_t p = H6(s, i, j, k, l, n, m);
            int c[6] = {i, j, k, l, n, m};
            double v = vel[p];
            /* uCFL = ABS(v*dt/dxt(i)): (v*dt)/dxt */
            double ucfl_arg = v * dt / G1(dd, c[d]);
            double Rjp = E[p + 2 * st] - E[p + st];
            double Rj = E[p + st] - E[p];
            double Rjm = E[p] - E[p - st];
            double cr;
            if (Rj != 0.0) cr = (v > 0.0 ? Rjm : Rjp) / Rj;
            else cr = (v > 0.0 ? Rjm : Rjp) * BIG_R4;
            cr = limiter(cr);
            F[p] = v * (E[p + st] + E[p]) * 0.5 - fabs(v) * ((1.0 - cr) + fabs(ucfl_arg) * cr) * Rj * 0.5;
        }
    }
}

/* ------------------------------------------------------------ integrate */

void or_integrate(or_state *s, double dt, int nsteps)
{
    for (int it = 0; it < nsteps; ++it) {
        or_fill_ghosts(s);
        or_superbee_fluxes(s, dt);
        /* dE = 0 on the interior, then flux divergence in x,y,z,k,l,m order, :105-200 */
        LOOP6I(s) { s->dE[H6(s, i, j, k, l, n, m)] = 0.0; }
        for (int d = 0; d < 6; ++d) {
            if (s->n[d] <= 1) continue;
            size_t st = stride6(s, d);
            const double *F = s->flux[d];
            const double *dd = s->dt_[d];
            LOOP6I(s) {
                int c[6] = {i, j, k, l, n, m};
                size_t p = H6(s, i, j, k, l, n, m);
                s->dE[p] = s->dE[p] - (F[p] - F[p - st]) / G1(dd, c[d]);
            }
        }
        /* meanfl, :202-252 */
        LOOP6I(s) { s->meanfl[H6(s, i, j, k, l, n, m)] = 0.0; }
        for (int d = 0; d < 3; ++d) {
            if (s->n[d] <= 1) continue;
            LOOP6I(s) {
                size_t p = H6(s, i, j, k, l, n, m);
                s->meanfl[p] = s->meanfl[p] + s->E[p] * s->lam[d][p];
            }
        }
        /* Euler update, :255-268 */
        LOOP6I(s) {
            size_t p = H6(s, i, j, k, l, n, m);
            s->dE[p] = s->dE[p] - s->meanfl[p];
            s->E[p] = s->E[p] + dt * s->dE[p];
        }
    }
}

/* ------------------------------------------------------------ diagnose */

/*
 * Scalars written to out[] (diagnose.f90:45-329):
 *  0 eint  1 mflint  2 mfl_max
 *  3..8  umax vmax wmax kmax lmax mmax
 *  9..20 bfl_xs bfl_xe bfl_ys bfl_ye bfl_zs bfl_ze bfl_ks bfl_ke bfl_ls bfl_le bfl_ms bfl_me
 *  21..26 cfl_x .. cfl_m
 *  27 residual
 * 3D outputs (nx*ny*nz interior, Fortran order), any may be NULL:
 *  E_int, meanfl_int, diss[6] = diss_ks diss_ke diss_ls diss_le diss_ms diss_me
 */
void or_diagnose(const or_state *s, double dt, double *out, double *E_int, double *meanfl_int, double **diss)
{
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2], nk = s->n[3], nl = s->n[4], nm = s->n[5];
    const double *dxt = s->dt_[0], *dyt = s->dt_[1], *dzt = s->dt_[2];
    const double *dkt = s->dt_[3], *dlt = s->dt_[4], *dmt = s->dt_[5];
    double umax = 0, vmax = 0, wmax = 0, kmax = 0, lmax = 0, mmax = 0, eint = 0, mflint = 0, mfl_max = 0;
    const double eps_r4 = (double)1e-12f;
    LOOP6I(s) {
        size_t p = H6(s, i, j, k, l, n, m);
        double fxa = G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
        eint = eint + s->dE[p] * fxa;
        mflint = mflint + s->meanfl[p] * fxa;
        double den = s->E[p] > eps_r4 ? s->E[p] : eps_r4;
        double r = fabs(s->meanfl[p]) / den;
        mfl_max = mfl_max > r ? mfl_max : r;
        double a;
        a = fabs(s->vel[0][p]); umax = umax > a ? umax : a;
        a = fabs(s->vel[1][p]); vmax = vmax > a ? vmax : a;
        a = fabs(s->vel[2][p]); wmax = wmax > a ? wmax : a;
        a = fabs(s->vel[3][p]); kmax = kmax > a ? kmax : a;
        a = fabs(s->vel[4][p]); lmax = lmax > a ? lmax : a;
        a = fabs(s->vel[5][p]); mmax = mmax > a ? mmax : a;
    }
    size_t n3 = (size_t)nx * ny * nz;
    if (E_int) memset(E_int, 0, n3 * sizeof(double));
    if (meanfl_int) memset(meanfl_int, 0, n3 * sizeof(double));
    if (E_int || meanfl_int) {
        LOOP6I(s) {
            size_t p = H6(s, i, j, k, l, n, m);
            size_t q = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
            if (E_int) E_int[q] = E_int[q] + s->E[p] * G1(dlt, n) * G1(dmt, m) * G1(dkt, l);
            if (meanfl_int) meanfl_int[q] = meanfl_int[q] + s->meanfl[p] * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
        }
    }
    double bfl[12] = {0};
    /* x */
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j)
        bfl[0] = bfl[0] + s->flux[0][H6(s, 0, j, k, l, n, m)] * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j)
        bfl[1] = bfl[1] + s->flux[0][H6(s, nx, j, k, l, n, m)] * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
    /* y */
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i)
        bfl[2] = bfl[2] + s->flux[1][H6(s, i, 0, k, l, n, m)] * G1(dxt, i) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i)
        bfl[3] = bfl[3] + s->flux[1][H6(s, i, ny, k, l, n, m)] * G1(dxt, i) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
    /* z */
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i)
        bfl[4] = bfl[4] + s->flux[2][H6(s, i, j, 0, l, n, m)] * G1(dxt, i) * G1(dyt, j) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i)
        bfl[5] = bfl[5] + s->flux[2][H6(s, i, j, nz, l, n, m)] * G1(dxt, i) * G1(dyt, j) * G1(dkt, l) * G1(dlt, n) * G1(dmt, m);
    /* k, with diss_ks / diss_ke */
    double *dks = diss ? diss[0] : NULL, *dke = diss ? diss[1] : NULL;
    double *dls = diss ? diss[2] : NULL, *dle = diss ? diss[3] : NULL;
    double *dms = diss ? diss[4] : NULL, *dme = diss ? diss[5] : NULL;
    if (diss) for (int q = 0; q < 6; ++q) if (diss[q]) memset(diss[q], 0, n3 * sizeof(double));
#define Q3 ((size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1)))
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
        double f = s->flux[3][H6(s, i, j, k, 0, n, m)];
        bfl[6] = bfl[6] + f * G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dlt, n) * G1(dmt, m);
        if (dks) dks[Q3] = dks[Q3] + f * G1(dlt, n) * G1(dmt, m);
    }
    for (int m = 1; m <= nm; ++m) for (int n = 1; n <= nl; ++n)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
        double f = s->flux[3][H6(s, i, j, k, nk, n, m)];
        bfl[7] = bfl[7] + f * G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dlt, n) * G1(dmt, m);
        if (dke) dke[Q3] = dke[Q3] + f * G1(dlt, n) * G1(dmt, m);
    }
    /* l */
    for (int m = 1; m <= nm; ++m) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
        double f = s->flux[4][H6(s, i, j, k, l, 0, m)];
        bfl[8] = bfl[8] + f * G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dmt, m);
        if (dls) dls[Q3] = dls[Q3] + f * G1(dkt, l) * G1(dmt, m);
    }
    for (int m = 1; m <= nm; ++m) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
        double f = s->flux[4][H6(s, i, j, k, l, nl, m)];
        bfl[9] = bfl[9] + f * G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dmt, m);
        if (dle) dle[Q3] = dle[Q3] + f * G1(dkt, l) * G1(dmt, m);
    }
    /* m */
    for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
        double f = s->flux[5][H6(s, i, j, k, l, n, 0)];
        bfl[10] = bfl[10] + f * G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n);
        if (dms) dms[Q3] = dms[Q3] + f * G1(dkt, l) * G1(dlt, n);
    }
    for (int n = 1; n <= nl; ++n) for (int l = 1; l <= nk; ++l)
    for (int k = 1; k <= nz; ++k) for (int j = 1; j <= ny; ++j) for (int i = 1; i <= nx; ++i) {
        double f = s->flux[5][H6(s, i, j, k, l, n, nm)];
        bfl[11] = bfl[11] + f * G1(dxt, i) * G1(dyt, j) * G1(dzt, k) * G1(dkt, l) * G1(dlt, n);
        if (dme) dme[Q3] = dme[Q3] + f * G1(dkt, l) * G1(dlt, n);
    }
#undef Q3
    /* CFL numbers, :294-313 */
    double cfl[6] = {0};
    double vmaxs[6] = {umax, vmax, wmax, kmax, lmax, mmax};
    for (int d = 0; d < 6; ++d)
        for (int i = 1; i <= s->n[d]; ++i) {
            double c = vmaxs[d] * dt / G1(s->dt_[d], i);
            cfl[d] = cfl[d] > c ? cfl[d] : c;
        }
    out[0] = eint; out[1] = mflint; out[2] = mfl_max;
    for (int d = 0; d < 6; ++d) out[3 + d] = vmaxs[d];
    for (int q = 0; q < 12; ++q) out[9 + q] = bfl[q];
    for (int d = 0; d < 6; ++d) out[21 + d] = cfl[d];
    /* residual, :327-328 */
    out[27] = eint + mflint + bfl[1] + bfl[3] + bfl[5] + bfl[7] + bfl[9] + bfl[11]
            - bfl[0] - bfl[2] - bfl[4] - bfl[6] - bfl[8] - bfl[10];
}

/* ------------------------------------------------------------ field access */

/*
 * Field ids: 0 E, 1 dE, 2 meanfl, 3 omega_i, 4 omega_e,
 *            5..10 xdot..mdot, 11..16 flux_x..flux_m, 17..19 lambda_x..z.
 * region 0: interior (n_0 x ... x n_5).
 * region 1: faces of the field's own axis (velocities and fluxes only):
 *           index 0..n_d along that axis, interior along the others.
 */
static double *field_ptr(or_state *s, int id)
{
    switch (id) {
    case 0: return s->E;
    case 1: return s->dE;
    case 2: return s->meanfl;
    case 3: return s->omega_i;
    case 4: return s->omega_e;
    default: break;
    }
    if (id >= 5 && id <= 10) return s->vel[id - 5];
    if (id >= 11 && id <= 16) return s->flux[id - 11];
    if (id >= 17 && id <= 19) return s->lam[id - 17];
    return NULL;
}

static int field_axis(int id)
{
    if (id >= 5 && id <= 10) return id - 5;
    if (id >= 11 && id <= 16) return id - 11;
    return -1;
}

int or_get_field(or_state *s, int id, int region, double *out)
{
    double *f = field_ptr(s, id);
    if (!f) return -1;
    int ax = region == 1 ? field_axis(id) : -1;
    if (region == 1 && ax < 0) return -2;
    int lo[6] = {1, 1, 1, 1, 1, 1};
    if (ax >= 0) lo[ax] = 0;
    size_t q = 0;
    for (int m = lo[5]; m <= s->n[5]; ++m)
    for (int n = lo[4]; n <= s->n[4]; ++n)
    for (int l = lo[3]; l <= s->n[3]; ++l)
    for (int k = lo[2]; k <= s->n[2]; ++k)
    for (int j = lo[1]; j <= s->n[1]; ++j)
    for (int i = lo[0]; i <= s->n[0]; ++i)
        out[q++] = f[H6(s, i, j, k, l, n, m)];
    return 0;
}

int or_set_field(or_state *s, int id, int region, const double *in)
{
    double *f = field_ptr(s, id);
    if (!f) return -1;
    int ax = region == 1 ? field_axis(id) : -1;
    if (region == 1 && ax < 0) return -2;
    int lo[6] = {1, 1, 1, 1, 1, 1};
    if (ax >= 0) lo[ax] = 0;
    size_t q = 0;
    for (int m = lo[5]; m <= s->n[5]; ++m)
    for (int n = lo[4]; n <= s->n[4]; ++n)
    for (int l = lo[3]; l <= s->n[3]; ++l)
    for (int k = lo[2]; k <= s->n[2]; ++k)
    for (int j = lo[1]; j <= s->n[1]; ++j)
    for (int i = lo[0]; i <= s->n[0]; ++i)
        f[H6(s, i, j, k, l, n, m)] = in[q++];
    return 0;
}
