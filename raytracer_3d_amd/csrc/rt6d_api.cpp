// rt6d_api.cpp -- C ABI of the 6D step (include/rt6d.h): context, device
// memory, call ordering and dispatch to the kernels of rt6d_exact.hip and
// rt6d_fast.hip.  Host-side only; every numerical operation on 6D data runs
// on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rt6d.h"
#include "rt6d_kernels.h"

using namespace rt6d;

// Scratch of rt6d_diagnose_noalloc, one copy per GPU in a file-scope device
// array: the per-block partials of the diagnose sums, the reduced slots,
// and the staging of the 3D outputs (E_int, meanfl_int, six diss fields)
// for grids of up to kNoAllocMaxN3 (x,y,z) columns.  4.5 MB per GPU.
constexpr size_t kNoAllocMaxN3 = 65536;
constexpr size_t kNoAllocDoubles = (size_t)kDiagParts * kDiagSlots + 16 + 8 * kNoAllocMaxN3;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(RT6D_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

Geo make_geo(const int *n)
{
    Geo g{};
    uint64_t cells = 1;
    for (int d = 0; d < kAxes; ++d) {
        g.n[d] = n[d];
        g.inner[d] = (int)cells;
        cells *= (uint64_t)n[d];
        if (n[d] > 1) g.act |= 1 << d;
    }
    for (int d = 0; d < 5; ++d) g.div[d] = make_fastdiv((uint32_t)n[d]);
    g.ncells = (uint32_t)cells;
    return g;
}

// its device address on each GPU, resolved by rt6d_create with that device
// current (the array's host-side name is not a device address)
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

int resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return 0;   // rt6d_diagnose_noalloc then refuses
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(RT6D_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[device] = (double *)p;
    return 0;
}

}  // namespace

struct rt6d_ctx {
    int n[kAxes] = {0};
    int up = 1, down = 1, prec = RT6D_FP64;
    size_t esz = 8;               // bytes per element of the step precision
    Geo g{};
    uint64_t cells = 0;
    int device = 0;
    int ncu = 256;   // compute units of the device (kernel choice)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double f0 = 1e-4, sign = 1.0;
    std::vector<double> t_h[kAxes], dt_h[kAxes];
    bool grid_set[kAxes] = {false, false, false, false, false, false};
    bool bg_set = false;
    // device buffers
    double *d_grid = nullptr;     // per axis: t (n+4) then dt (n+4)
    size_t grid_off[kAxes] = {0};
    double *d_bg = nullptr;       // Nsqr, U, V halo'd 3D
    size_t n3h = 0;
    double *d_coef = nullptr;     // per axis: dx, rdx, cfl (n+1 each)
    size_t coef_off[kAxes] = {0};
    double coef_dt = NAN;
    void *d_E[2] = {nullptr, nullptr};
    int cur = 0;
    bool E_set = false, have_prev = false;
    double prev_dt = 0.0;
    long long steps = 0;
    void *d_vel[kAxes] = {nullptr};
    void *d_lam = nullptr;
    size_t vel_count[kAxes] = {0};
    bool params_ready = false;
    bool homog = false;
    int opt_exact = 0, opt_path = 0, opt_block = 256, opt_kernel = 0, opt_sync = 1;
    void *d_hx = nullptr, *d_hy = nullptr, *d_hz = nullptr, *d_hm = nullptr, *d_hl = nullptr, *d_hrec = nullptr;
    uint32_t n4 = 0, inner4m = 0;
    double vmax[kAxes] = {0};
    double cflsum_max = 0.0;
    void *d_tdE = nullptr, *d_tmf = nullptr, *d_tfl[kAxes] = {nullptr};
    int fl_src = 0, de_src = 0;   // 0 stale, 1 from the last step's start state, 2 from current E
    double *d_partial = nullptr;
    uint32_t nparts = 0;
    double *d_dout = nullptr;
    double *d_d3 = nullptr;       // E_int, meanfl_int, 6 diss arrays
    size_t dev_bytes = 0;
    int *d_flag = nullptr;
    unsigned long long *d_bits = nullptr;
    rt6d_ensemble *owner = nullptr;  // ensemble this context is a member of
    int forcing = 0;                 // bottom forcing applied before each step (0 none)
    double *d_force = nullptr;       // its 5D field over (x,y,k,l,m)

    int dalloc(void **p, size_t bytes)
    {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8);
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(RT6D_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
        }
        dev_bytes += bytes;
        return 0;
    }
    void dfree(void *&p)
    {
        if (p) hipFree(p);
        p = nullptr;
    }
    // free a buffer dalloc counted with `bytes`
    void dfree_n(void *&p, size_t bytes)
    {
        if (p) dev_bytes -= bytes;
        dfree(p);
    }
};

namespace {

int ens_settle(rt6d_ensemble *e);

int check_ctx(rt6d_ctx *c)
{
    if (!c) return fail(RT6D_EINVAL, "null context");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(RT6D_EHIP, "hipSetDevice: %s", hipGetErrorString(e));
    // an ensemble member's state is current once the ensemble's enqueued steps are done
    if (c->owner) return ens_settle(c->owner);
    return 0;
}

AxisCoef axis_coef(const rt6d_ctx *c);

// The plane records of the homogeneous path for step dt, the dt of the axis
// coefficients (queued on the context's stream).
int build_records(rt6d_ctx *c, double dt)
{
    if (!c->d_hrec) return 0;
    const AxisCoef ax = axis_coef(c);
    const hipError_t e =
        c->prec == RT6D_FP64
            ? launch_build_records<double>(c->g, (const double *)c->d_hx, (const double *)c->d_hy,
                                           (const double *)c->d_hz, (const double *)c->d_hm, (const double *)c->d_hl,
                                           ax, dt, (double *)c->d_hrec, c->stream)
            : launch_build_records<float>(c->g, (const float *)c->d_hx, (const float *)c->d_hy,
                                          (const float *)c->d_hz, (const float *)c->d_hm, (const float *)c->d_hl, ax,
                                          dt, (float *)c->d_hrec, c->stream);
    HIPCHK(e);
    return 0;
}

// dx, rdx and dt/dx of every axis for step dt (binary64, on the device), and
// the plane records that depend on them
int ensure_coef(rt6d_ctx *c, double dt)
{
    if (c->coef_dt == dt) return 0;
    size_t tot = 0;
    for (int d = 0; d < kAxes; ++d) tot += 5 * (size_t)(c->n[d] + 1);
    std::vector<double> h(tot);
    for (int d = 0; d < kAxes; ++d) {
        const int n = c->n[d];
        double *dx = h.data() + c->coef_off[d];
        double *rdx = dx + (n + 1);
        double *cf = rdx + (n + 1);
        double *cr = cf + (n + 1);
        for (int f = 0; f <= n; ++f) {
            dx[f] = c->dt_h[d][f + 1];  // Fortran index f -> halo slot f+1
            rdx[f] = 1.0 / dx[f];
            cf[f] = dt / dx[f];
            cr[2 * f] = cf[f];
            cr[2 * f + 1] = rdx[f];
        }
    }
    HIPCHK(hipMemcpyAsync(c->d_coef, h.data(), tot * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int r;
    if ((r = build_records(c, dt))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->coef_dt = dt;
    return 0;
}

AxisCoef axis_coef(const rt6d_ctx *c)
{
    AxisCoef a{};
    for (int d = 0; d < kAxes; ++d) {
        a.dx[d] = c->d_coef + c->coef_off[d];
        a.rdx[d] = a.dx[d] + (c->n[d] + 1);
        a.cfl[d] = a.rdx[d] + (c->n[d] + 1);
        a.cr[d] = a.cfl[d] + (c->n[d] + 1);
    }
    return a;
}

ParamIn param_in(const rt6d_ctx *c)
{
    ParamIn pi{};
    for (int d = 0; d < kAxes; ++d) {
        pi.t[d] = c->d_grid + c->grid_off[d];
        pi.dt[d] = pi.t[d] + (c->n[d] + 4);
        pi.n[d] = c->n[d];
    }
    pi.Nsqr = c->d_bg;
    pi.U = c->d_bg + c->n3h;
    pi.V = c->d_bg + 2 * c->n3h;
    pi.f0 = c->f0;
    pi.sign = c->sign;
    pi.up = c->up;
    pi.down = c->down;
    return pi;
}

template <typename T>
StepPtrs<T> step_ptrs(const rt6d_ctx *c, const void *Ein, void *Eout, double dt)
{
    StepPtrs<T> p{};
    p.Ein = (const T *)Ein;
    p.Eout = (T *)Eout;
    for (int d = 0; d < kAxes; ++d) p.vel[d] = (const T *)c->d_vel[d];
    p.lam = (const T *)c->d_lam;
    p.hom.vx3 = (const T *)c->d_hx;
    p.hom.vy3 = (const T *)c->d_hy;
    p.hom.vz = (const T *)c->d_hz;
    p.hom.vm = (const T *)c->d_hm;
    p.hom.lam = (const T *)c->d_hl;
    p.hom.rec = (const T *)c->d_hrec;
    p.hom.n4 = c->n4;
    p.hom.inner4m = c->inner4m;
    p.ax = axis_coef(c);
    p.dt = dt;
    p.dE = (T *)c->d_tdE;
    p.meanfl = (T *)c->d_tmf;
    for (int d = 0; d < kAxes; ++d) p.flux[d] = (T *)c->d_tfl[d];
    return p;
}

// a context that belongs to an ensemble is stepped only by the ensemble
int not_member(const rt6d_ctx *c, const char *fn)
{
    if (c->owner) return fail(RT6D_ESTATE, "%s: context is an ensemble member (remove it first)", fn);
    return 0;
}

void ens_detach(rt6d_ensemble *e, rt6d_ctx *c);
int ens_settle(rt6d_ensemble *e);

int require_ready(rt6d_ctx *c)
{
    if (!c->params_ready) return fail(RT6D_ESTATE, "calc_parameter has not run");
    if (!c->E_set) return fail(RT6D_ESTATE, "E has not been uploaded");
    return 0;
}

int ensure_tendency_buffers(rt6d_ctx *c)
{
    if (c->d_tdE) return 0;
    int r;
    if ((r = c->dalloc(&c->d_tdE, c->cells * c->esz))) return r;
    if ((r = c->dalloc(&c->d_tmf, c->cells * c->esz))) return r;
    for (int d = 0; d < kAxes; ++d) {
        if ((r = c->dalloc(&c->d_tfl[d], c->vel_count[d] * c->esz))) return r;
        HIPCHK(hipMemsetAsync(c->d_tfl[d], 0, c->vel_count[d] * c->esz, c->stream));
    }
    HIPCHK(hipMemsetAsync(c->d_tdE, 0, c->cells * c->esz, c->stream));
    HIPCHK(hipMemsetAsync(c->d_tmf, 0, c->cells * c->esz, c->stream));
    return 0;
}

// Tendency (dE, meanfl, fluxes) of E buffer `which` with step dt.
int run_tendency(rt6d_ctx *c, const void *E, double dt)
{
    int r;
    if ((r = ensure_tendency_buffers(c))) return r;
    if ((r = ensure_coef(c, dt))) return r;
    hipError_t e;
    if (c->prec == RT6D_FP64) {
        StepPtrs<double> p = step_ptrs<double>(c, E, nullptr, dt);
        e = c->opt_exact ? launch_tendency_exact(c->g, p, c->homog, c->stream)
                         : launch_tendency_fast<double>(c->g, p, c->homog, c->stream);
    } else {
        StepPtrs<float> p = step_ptrs<float>(c, E, nullptr, dt);
        e = launch_tendency_fast<float>(c->g, p, c->homog, c->stream);
    }
    HIPCHK(e);
    return 0;
}

// dE/meanfl/fluxes as integrate leaves them after the last step
int ensure_step_byproducts(rt6d_ctx *c, bool need_flux)
{
    if (c->de_src == 1 && (!need_flux || c->fl_src != 0)) return 0;
    int r;
    if ((r = ensure_tendency_buffers(c))) return r;
    if (!c->have_prev) {
        // no step yet: the reference's zero-initialised arrays
        HIPCHK(hipMemsetAsync(c->d_tdE, 0, c->cells * c->esz, c->stream));
        HIPCHK(hipMemsetAsync(c->d_tmf, 0, c->cells * c->esz, c->stream));
        for (int d = 0; d < kAxes; ++d)
            HIPCHK(hipMemsetAsync(c->d_tfl[d], 0, c->vel_count[d] * c->esz, c->stream));
    } else {
        if ((r = run_tendency(c, c->d_E[c->cur ^ 1], c->prev_dt))) return r;
    }
    c->de_src = 1;
    c->fl_src = 1;
    return 0;
}

int download_typed(rt6d_ctx *c, const void *src, size_t count, double *dst)
{
    if (c->prec == RT6D_FP64) {
        HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(count);
        HIPCHK(hipMemcpyAsync(tmp.data(), src, count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < count; ++i) dst[i] = (double)tmp[i];
    }
    return 0;
}

// copy an interior-shaped array into a halo'd host array (interior part only)
void interior_to_halo(const int *n, const double *in, double *out)
{
    const int H0 = n[0] + 4, H1 = n[1] + 4, H2 = n[2] + 4, H3 = n[3] + 4, H4 = n[4] + 4;
    size_t q = 0;
    for (int m = 0; m < n[5]; ++m)
        for (int l5 = 0; l5 < n[4]; ++l5)
            for (int l = 0; l < n[3]; ++l)
                for (int k = 0; k < n[2]; ++k)
                    for (int j = 0; j < n[1]; ++j) {
                        size_t base = (size_t)2 + H0 * ((size_t)(j + 2) + H1 * ((size_t)(k + 2) +
                                      H2 * ((size_t)(l + 2) + H3 * ((size_t)(l5 + 2) + (size_t)H4 * (m + 2)))));
                        memcpy(out + base, in + q, (size_t)n[0] * sizeof(double));
                        q += n[0];
                    }
}

void halo_to_interior(const int *n, const double *in, double *out)
{
    const int H0 = n[0] + 4, H1 = n[1] + 4, H2 = n[2] + 4, H3 = n[3] + 4, H4 = n[4] + 4;
    size_t q = 0;
    for (int m = 0; m < n[5]; ++m)
        for (int l5 = 0; l5 < n[4]; ++l5)
            for (int l = 0; l < n[3]; ++l)
                for (int k = 0; k < n[2]; ++k)
                    for (int j = 0; j < n[1]; ++j) {
                        size_t base = (size_t)2 + H0 * ((size_t)(j + 2) + H1 * ((size_t)(k + 2) +
                                      H2 * ((size_t)(l + 2) + H3 * ((size_t)(l5 + 2) + (size_t)H4 * (m + 2)))));
                        memcpy(out + q, in + base, (size_t)n[0] * sizeof(double));
                        q += n[0];
                    }
}

// Faces 1..n_d of a face array (the reference's T-point values) -> interior.
void faces_to_interior(const int *n, int d, const double *in, double *out)
{
    size_t inner = 1, outer = 1;
    for (int e = 0; e < d; ++e) inner *= (size_t)n[e];
    for (int e = d + 1; e < kAxes; ++e) outer *= (size_t)n[e];
    const size_t nd = (size_t)n[d];
    for (size_t o = 0; o < outer; ++o)
        for (size_t f = 1; f <= nd; ++f)
            memcpy(out + (o * nd + (f - 1)) * inner, in + (o * (nd + 1) + f) * inner, inner * sizeof(double));
}

}  // namespace

// ====================================================================== ABI

extern "C" {

const char *rt6d_version(void) { return "rt6d 0.1 gfx950"; }

const char *rt6d_last_error(void) { return g_err.c_str(); }

int rt6d_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rt6d_build_grid(int axis, int n, const double *spacing, double fixed, double *t, double *dt, double *u,
                    double *du)
{
    if (axis < 0 || axis >= kAxes || n < 1 || !spacing) return fail(RT6D_EINVAL, "rt6d_build_grid: bad arguments");
    const int N = n + 4;
    std::vector<double> D(N), Du(N), Tt(N), Uu(N);
    // Fortran index f in [-1, n+2] at slot f+1; ghosts take the edge spacing (calc_grid.f90:104-111)
    for (int f = 1; f <= n; ++f) D[f + 1] = spacing[f - 1];
    for (int i = 1; i <= 2; ++i) {
        D[n + i + 1] = D[n + 1];
        D[1 - i + 1] = D[2];
    }
    // u_centered_grid (calc_grid.f90:300-322) over the halo'd array
    Uu[0] = 0.0;
    for (int p = 1; p < N; ++p) Uu[p] = Uu[p - 1] + D[p];
    Tt[0] = Uu[0] - D[0] * 0.5;
    for (int p = 1; p < N; ++p) Tt[p] = 2.0 * Uu[p - 1] - Tt[p - 1];
    for (int p = 0; p < N - 1; ++p) Du[p] = Tt[p + 1] - Tt[p];
    Du[N - 1] = 2.0 * D[N - 1] - Du[(N - 2) > 0 ? N - 2 : 0];
    double sh;
    if (axis <= 1) {
        sh = Du[2];           // dxu(1), calc_grid.f90:115,124
    } else if (axis == 2) {
        sh = Uu[n + 1];       // zu(nz), :133
    } else if (n > 1) {
        sh = Uu[n / 2 + 1];   // u(n/2), :143,156,169
    } else {
        sh = 0.0;
    }
    for (int p = 0; p < N; ++p) {
        if (axis >= 3 && n == 1) {
            Tt[p] = fixed;
            Uu[p] = fixed;
        } else {
            Tt[p] = Tt[p] - sh;
            Uu[p] = Uu[p] - sh;
        }
    }
    if (t) memcpy(t, Tt.data(), N * sizeof(double));
    if (dt) memcpy(dt, D.data(), N * sizeof(double));
    if (u) memcpy(u, Uu.data(), N * sizeof(double));
    if (du) memcpy(du, Du.data(), N * sizeof(double));
    return 0;
}

int rt6d_create(const int *n, int reflect_up, int reflect_down, int precision, rt6d_ctx **out)
{
    if (!n || !out) return fail(RT6D_EINVAL, "rt6d_create: null argument");
    *out = nullptr;
    uint64_t cells = 1;
    for (int d = 0; d < kAxes; ++d) {
        if (n[d] < 1) return fail(RT6D_EINVAL, "rt6d_create: n[%d] = %d < 1", d, n[d]);
        cells *= (uint64_t)n[d];
    }
    // face arrays hold cells*(n_d+1)/n_d <= 2*cells elements, indexed in 32 bits
    if (cells * 2 >= (1ull << 31)) return fail(RT6D_EINVAL, "rt6d_create: %llu cells exceed the 2^30 limit",
                                               (unsigned long long)cells);
    if (precision != RT6D_FP64 && precision != RT6D_FP32)
        return fail(RT6D_EINVAL, "rt6d_create: precision %d", precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RT6D_EHIP, "rt6d_create: no HIP device");
    rt6d_ctx *c = new rt6d_ctx();
    for (int d = 0; d < kAxes; ++d) c->n[d] = n[d];
    c->up = reflect_up ? 1 : 0;
    c->down = reflect_down ? 1 : 0;
    c->prec = precision;
    c->esz = precision == RT6D_FP64 ? 8 : 4;
    c->cells = cells;
    c->g = make_geo(n);
    c->g.up = c->up;
    c->g.down = c->down;
    hipGetDevice(&c->device);
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
            c->ncu = prop.multiProcessorCount;
    }
    int r = 0;
    auto bail = [&](int code) {
        rt6d_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(RT6D_EHIP, "hipStreamCreate failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(RT6D_EHIP, "hipEventCreate failed"));
    // set-up of rt6d_diagnose_noalloc: the scratch address on this device
    if ((r = resolve_scratch(c->device))) return bail(r);
    size_t gtot = 0, ctot = 0;
    for (int d = 0; d < kAxes; ++d) {
        c->grid_off[d] = gtot;
        gtot += 2 * (size_t)(n[d] + 4);
        c->coef_off[d] = ctot;
        ctot += 5 * (size_t)(n[d] + 1);
        c->t_h[d].assign(n[d] + 4, 0.0);
        c->dt_h[d].assign(n[d] + 4, 0.0);
        size_t fc = cells / (uint64_t)n[d] * (uint64_t)(n[d] + 1);
        c->vel_count[d] = fc;
    }
    c->n3h = (size_t)(n[0] + 4) * (n[1] + 4) * (n[2] + 4);
    if ((r = c->dalloc((void **)&c->d_grid, gtot * sizeof(double)))) return bail(r);
    if ((r = c->dalloc((void **)&c->d_coef, ctot * sizeof(double)))) return bail(r);
    if ((r = c->dalloc((void **)&c->d_bg, 3 * c->n3h * sizeof(double)))) return bail(r);
    if ((r = c->dalloc(&c->d_E[0], cells * c->esz))) return bail(r);
    if ((r = c->dalloc(&c->d_E[1], cells * c->esz))) return bail(r);
    if ((r = c->dalloc((void **)&c->d_flag, sizeof(int)))) return bail(r);
    if ((r = c->dalloc((void **)&c->d_bits, 8 * sizeof(unsigned long long)))) return bail(r);
    if (hipMemsetAsync(c->d_E[0], 0, cells * c->esz, c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_E[1], 0, cells * c->esz, c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_bg, 0, 3 * c->n3h * sizeof(double), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(RT6D_EHIP, "initial clear failed"));
    c->n4 = (uint32_t)(cells / ((uint64_t)n[0] * n[1]));
    c->inner4m = (uint32_t)n[2] * n[3] * n[4];
    *out = c;
    return 0;
}

int rt6d_destroy(rt6d_ctx *c)
{
    if (!c) return 0;
    if (c->owner) ens_detach(c->owner, c);
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    c->dfree((void *&)c->d_force);
    c->dfree((void *&)c->d_grid);
    c->dfree((void *&)c->d_coef);
    c->dfree((void *&)c->d_bg);
    c->dfree(c->d_E[0]);
    c->dfree(c->d_E[1]);
    for (int d = 0; d < kAxes; ++d) {
        c->dfree(c->d_vel[d]);
        c->dfree(c->d_tfl[d]);
    }
    c->dfree(c->d_lam);
    c->dfree(c->d_hx);
    c->dfree(c->d_hy);
    c->dfree(c->d_hz);
    c->dfree(c->d_hm);
    c->dfree(c->d_hl);
    c->dfree(c->d_hrec);
    c->dfree(c->d_tdE);
    c->dfree(c->d_tmf);
    c->dfree((void *&)c->d_partial);
    c->dfree((void *&)c->d_dout);
    c->dfree((void *&)c->d_d3);
    c->dfree((void *&)c->d_flag);
    c->dfree((void *&)c->d_bits);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int rt6d_set_grid(rt6d_ctx *c, int axis, const double *t, const double *dt)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = not_member(c, "rt6d_set_grid"))) return r;
    if (axis < 0 || axis >= kAxes || !t || !dt) return fail(RT6D_EINVAL, "rt6d_set_grid: bad arguments");
    const int N = c->n[axis] + 4;
    for (int p = 0; p < N; ++p) {
        if (!(dt[p] > 0.0) || !std::isfinite(dt[p]))
            return fail(RT6D_EINVAL, "rt6d_set_grid: axis %d spacing[%d] = %g is not positive", axis, p, dt[p]);
    }
    memcpy(c->t_h[axis].data(), t, N * sizeof(double));
    memcpy(c->dt_h[axis].data(), dt, N * sizeof(double));
    HIPCHK(hipMemcpyAsync(c->d_grid + c->grid_off[axis], t, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_grid + c->grid_off[axis] + N, dt, N * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->grid_set[axis] = true;
    if (axis < 2) {
        // uniform spacing over the faces and cells the step kernels read (halo slots 1..n+1)
        bool uni = true;
        for (int p = 2; p <= c->n[axis] + 1; ++p) uni = uni && dt[p] == dt[1];
        c->g.uxy = uni ? (c->g.uxy | (1 << axis)) : (c->g.uxy & ~(1 << axis));
    }
    c->coef_dt = NAN;
    c->params_ready = false;
    return 0;
}

int rt6d_set_background(rt6d_ctx *c, const double *Nsqr, const double *U, const double *V, double f0,
                        double sign_of_omega_i)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = not_member(c, "rt6d_set_background"))) return r;
    if (!Nsqr || !U || !V) return fail(RT6D_EINVAL, "rt6d_set_background: null array");
    c->f0 = f0;
    c->sign = sign_of_omega_i;
    const size_t b = c->n3h * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_bg, Nsqr, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_bg + c->n3h, U, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_bg + 2 * c->n3h, V, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->bg_set = true;
    c->params_ready = false;
    return 0;
}

int rt6d_calc_parameter(rt6d_ctx *c)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = not_member(c, "rt6d_calc_parameter"))) return r;
    for (int d = 0; d < kAxes; ++d)
        if (!c->grid_set[d]) return fail(RT6D_ESTATE, "rt6d_calc_parameter: grid of axis %d not set", d);
    if (!c->bg_set) return fail(RT6D_ESTATE, "rt6d_calc_parameter: background not set");
    // release previous coefficient fields
    for (int d = 0; d < kAxes; ++d) c->dfree_n(c->d_vel[d], c->vel_count[d] * c->esz);
    c->dfree_n(c->d_lam, c->cells * c->esz);
    {
        const size_t es = c->esz;
        const size_t nzf = (size_t)c->n4 / c->n[2] * (c->n[2] + 1), nmf = (size_t)c->inner4m * (c->n[5] + 1);
        c->dfree_n(c->d_hx, 3 * (size_t)c->n4 * es);
        c->dfree_n(c->d_hy, 3 * (size_t)c->n4 * es);
        c->dfree_n(c->d_hz, nzf * es);
        c->dfree_n(c->d_hm, nmf * es);
        c->dfree_n(c->d_hl, (size_t)c->n4 * es);
        c->dfree_n(c->d_hrec, (size_t)kRec * c->n4 * es);
    }
    c->homog = false;
    c->params_ready = false;
    if ((r = ensure_coef(c, 1.0))) return r;
    const ParamIn pi = param_in(c);
    // binary64 face velocities and lambda
    double *v64[kAxes] = {nullptr};
    double *l64 = nullptr;
    auto cleanup64 = [&]() {
        for (int d = 0; d < kAxes; ++d)
            if (v64[d] && (void *)v64[d] != c->d_vel[d]) hipFree(v64[d]);
        if (l64 && (void *)l64 != c->d_lam) hipFree(l64);
    };
    for (int d = 0; d < kAxes; ++d) {
        if ((r = c->dalloc((void **)&v64[d], c->vel_count[d] * sizeof(double)))) { cleanup64(); return r; }
        hipError_t e = launch_param_velocity(d, pi, v64[d], c->stream);
        if (e != hipSuccess) { cleanup64(); return fail(RT6D_EHIP, "param velocity %d: %s", d, hipGetErrorString(e)); }
    }
    if ((r = c->dalloc((void **)&l64, c->cells * sizeof(double)))) { cleanup64(); return r; }
    {
        hipError_t e = launch_param_lambda(pi, v64, l64, nullptr, c->stream);
        if (e != hipSuccess) { cleanup64(); return fail(RT6D_EHIP, "param lambda: %s", hipGetErrorString(e)); }
    }
    // velocity maxima (diagnose) and the CFL sum (rt6d_cfl_dt)
    HIPCHK(hipMemsetAsync(c->d_bits, 0, 8 * sizeof(unsigned long long), c->stream));
    HIPCHK(launch_vmax(c->g, v64, c->d_bits, c->stream));
    HIPCHK(launch_cfl_max(c->g, v64, axis_coef(c), c->d_bits + 6, c->stream));
    unsigned long long bits[8];
    HIPCHK(hipMemcpyAsync(bits, c->d_bits, sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int d = 0; d < kAxes; ++d) memcpy(&c->vmax[d], &bits[d], 8);
    memcpy(&c->cflsum_max, &bits[6], 8);
    if (c->prec == RT6D_FP64) {
        for (int d = 0; d < kAxes; ++d) c->d_vel[d] = v64[d];
        c->d_lam = l64;
    } else {
        for (int d = 0; d < kAxes; ++d) {
            if ((r = c->dalloc(&c->d_vel[d], c->vel_count[d] * sizeof(float)))) { cleanup64(); return r; }
            HIPCHK(launch_convert_d2f(v64[d], (float *)c->d_vel[d], c->vel_count[d], c->stream));
        }
        if ((r = c->dalloc(&c->d_lam, c->cells * sizeof(float)))) { cleanup64(); return r; }
        HIPCHK(launch_convert_d2f(l64, (float *)c->d_lam, c->cells, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        cleanup64();
        for (int d = 0; d < kAxes; ++d) c->dev_bytes -= c->vel_count[d] * sizeof(double);
        c->dev_bytes -= c->cells * sizeof(double);
    }
    // horizontally homogeneous coefficients: tables over (z,k,l,m), checked
    // bit for bit against every cell before the path is used
    if (c->opt_path != 1 && c->n[0] * c->n[1] > 1) {
        const size_t es = c->esz;
        const size_t nzf = (size_t)c->n4 / c->n[2] * (c->n[2] + 1);
        const size_t nmf = (size_t)c->inner4m * (c->n[5] + 1);
        if ((r = c->dalloc(&c->d_hx, 3 * (size_t)c->n4 * es))) return r;
        if ((r = c->dalloc(&c->d_hy, 3 * (size_t)c->n4 * es))) return r;
        if ((r = c->dalloc(&c->d_hz, nzf * es))) return r;
        if ((r = c->dalloc(&c->d_hm, nmf * es))) return r;
        if ((r = c->dalloc(&c->d_hl, (size_t)c->n4 * es))) return r;
        HIPCHK(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        hipError_t e;
        if (c->prec == RT6D_FP64) {
            const double *vv[kAxes];
            for (int d = 0; d < kAxes; ++d) vv[d] = (const double *)c->d_vel[d];
            e = launch_build_homog<double>(c->g, vv, (const double *)c->d_lam, (double *)c->d_hx, (double *)c->d_hy,
                                           (double *)c->d_hz, (double *)c->d_hm, (double *)c->d_hl, c->d_flag,
                                           c->stream);
        } else {
            const float *vv[kAxes];
            for (int d = 0; d < kAxes; ++d) vv[d] = (const float *)c->d_vel[d];
            e = launch_build_homog<float>(c->g, vv, (const float *)c->d_lam, (float *)c->d_hx, (float *)c->d_hy,
                                          (float *)c->d_hz, (float *)c->d_hm, (float *)c->d_hl, c->d_flag, c->stream);
        }
        HIPCHK(e);
        int bad = 1;
        HIPCHK(hipMemcpyAsync(&bad, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->homog = bad == 0;
        if (c->homog) {
            if ((r = c->dalloc(&c->d_hrec, (size_t)kRec * c->n4 * es))) return r;
            if ((r = build_records(c, c->coef_dt))) return r;
            HIPCHK(hipStreamSynchronize(c->stream));
            // no kernel of the homogeneous path reads the 6D coefficient
            // fields: free them (downloads recompute them on demand)
            for (int d = 0; d < kAxes; ++d) c->dfree_n(c->d_vel[d], c->vel_count[d] * es);
            c->dfree_n(c->d_lam, c->cells * es);
        } else {
            c->dfree_n(c->d_hx, 3 * (size_t)c->n4 * es);
            c->dfree_n(c->d_hy, 3 * (size_t)c->n4 * es);
            c->dfree_n(c->d_hz, nzf * es);
            c->dfree_n(c->d_hm, nmf * es);
            c->dfree_n(c->d_hl, (size_t)c->n4 * es);
        }
    }
    c->params_ready = true;
    c->fl_src = c->de_src = 0;
    return 0;
}

int rt6d_cfl_dt(rt6d_ctx *c, double cfl, double *dt)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if (!dt || !(cfl > 0.0)) return fail(RT6D_EINVAL, "rt6d_cfl_dt: bad arguments");
    if (!c->params_ready) return fail(RT6D_ESTATE, "rt6d_cfl_dt: calc_parameter has not run");
    if (!(c->cflsum_max > 0.0)) return fail(RT6D_EINVAL, "rt6d_cfl_dt: all velocities vanish");
    *dt = cfl / c->cflsum_max;
    return 0;
}

int rt6d_upload_E(rt6d_ctx *c, const double *E, int layout)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = not_member(c, "rt6d_upload_E"))) return r;
    if (!E) return fail(RT6D_EINVAL, "rt6d_upload_E: null array");
    const double *src = E;
    std::vector<double> tmp;
    if (layout == RT6D_LAYOUT_HALO) {
        tmp.resize(c->cells);
        halo_to_interior(c->n, E, tmp.data());
        src = tmp.data();
    } else if (layout != RT6D_LAYOUT_INTERIOR) {
        return fail(RT6D_EINVAL, "rt6d_upload_E: layout %d", layout);
    }
    if (c->prec == RT6D_FP64) {
        HIPCHK(hipMemcpyAsync(c->d_E[c->cur], src, c->cells * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> f(c->cells);
        for (size_t i = 0; i < c->cells; ++i) f[i] = (float)src[i];
        HIPCHK(hipMemcpyAsync(c->d_E[c->cur], f.data(), c->cells * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->E_set = true;
    c->have_prev = false;
    c->steps = 0;
    c->fl_src = c->de_src = 0;
    return 0;
}

int rt6d_init_separable(rt6d_ctx *c, const double *factors, double peak, double noise, unsigned long long seed,
                        unsigned long long member)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = not_member(c, "rt6d_init_separable"))) return r;
    if (!factors || !std::isfinite(peak) || !std::isfinite(noise))
        return fail(RT6D_EINVAL, "rt6d_init_separable: bad arguments");
    size_t nf = 0;
    for (int d = 0; d < kAxes; ++d) nf += (size_t)c->n[d];
    double *d_f = nullptr;
    if ((r = c->dalloc((void **)&d_f, nf * sizeof(double)))) return r;
    // the host generator's key for stream 0: seed*FNV + member*golden (mod 2^64)
    const uint64_t key = (uint64_t)seed * 0x100000001B3ull + (uint64_t)member * 0x2545F4914F6CDD1Dull;
    hipError_t e = hipMemcpyAsync(d_f, factors, nf * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = c->prec == RT6D_FP64
                ? launch_init_separable<double>(c->g, d_f, peak, noise, key, (double *)c->d_E[c->cur], c->stream)
                : launch_init_separable<float>(c->g, d_f, peak, noise, key, (float *)c->d_E[c->cur], c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    c->dfree((void *&)d_f);
    c->dev_bytes -= nf * sizeof(double);
    HIPCHK(e);
    c->E_set = true;
    c->have_prev = false;
    c->steps = 0;
    c->fl_src = c->de_src = 0;
    return 0;
}

int rt6d_set_forcing(rt6d_ctx *c, int kind, const double *field)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = not_member(c, "rt6d_set_forcing"))) return r;
    if (kind < 0 || kind > 2) return fail(RT6D_EINVAL, "rt6d_set_forcing: kind %d", kind);
    if (kind == 0) {
        c->forcing = 0;
        return 0;
    }
    if (!field) return fail(RT6D_EINVAL, "rt6d_set_forcing: null field");
    if (c->n[2] < 1) return fail(RT6D_EINVAL, "rt6d_set_forcing: no z axis");
    const size_t n5 = c->cells / (size_t)c->n[2];
    if (!c->d_force && (r = c->dalloc((void **)&c->d_force, n5 * sizeof(double)))) return r;
    HIPCHK(hipMemcpyAsync(c->d_force, field, n5 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->forcing = kind;
    return 0;
}

int rt6d_branch_flips(rt6d_ctx *a, rt6d_ctx *b, long long faces[6], long long flips[6])
{
    int r;
    if (!a || !b || !faces || !flips) return fail(RT6D_EINVAL, "rt6d_branch_flips: null argument");
    if (a->prec != RT6D_FP64 || b->prec != RT6D_FP32)
        return fail(RT6D_EINVAL, "rt6d_branch_flips: needs a binary64 and a binary32 context");
    for (int d = 0; d < kAxes; ++d)
        if (a->n[d] != b->n[d]) return fail(RT6D_EINVAL, "rt6d_branch_flips: grids differ");
    if (a->device != b->device) return fail(RT6D_EINVAL, "rt6d_branch_flips: contexts on different devices");
    if ((r = check_ctx(a))) return r;
    if ((r = require_ready(a)) || (r = require_ready(b))) return r;
    HIPCHK(hipStreamSynchronize(b->stream));
    unsigned long long *d_cnt = nullptr;
    if ((r = a->dalloc((void **)&d_cnt, 12 * sizeof(unsigned long long)))) return r;
    const double dt = std::isfinite(a->coef_dt) ? a->coef_dt : 1.0;
    const StepPtrs<double> pa = step_ptrs<double>(a, a->d_E[a->cur], nullptr, dt);
    const StepPtrs<float> pb = step_ptrs<float>(b, b->d_E[b->cur], nullptr, dt);
    unsigned long long h[12] = {0};
    hipError_t e = hipMemsetAsync(d_cnt, 0, sizeof h, a->stream);
    if (e == hipSuccess) e = launch_branch_flips(a->g, pa, a->homog, pb, b->homog, d_cnt, a->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h, d_cnt, sizeof h, hipMemcpyDeviceToHost, a->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
    a->dfree((void *&)d_cnt);
    a->dev_bytes -= sizeof h;
    HIPCHK(e);
    for (int d = 0; d < kAxes; ++d) {
        faces[d] = (long long)h[d];
        flips[d] = (long long)h[6 + d];
    }
    return 0;
}

static int do_steps(rt6d_ctx *c, double dt, int nsteps)
{
    int r;
    if ((r = not_member(c, "rt6d_step"))) return r;
    if ((r = require_ready(c))) return r;
    if (nsteps < 0 || !std::isfinite(dt)) return fail(RT6D_EINVAL, "rt6d_step: bad arguments");
    if (nsteps == 0) return 0;
    if ((r = ensure_coef(c, dt))) return r;
    const LaunchCfg lc{c->opt_block, c->opt_kernel, c->ncu, c->stream};
    const double dz1 = c->dt_h[2][2];   // dzt(1): halo slot f+1 holds Fortran index f
    for (int s = 0; s < nsteps; ++s) {
        void *in = c->d_E[c->cur], *out = c->d_E[c->cur ^ 1];
        hipError_t e;
        if (c->forcing) {   // set_forcing precedes integrate (main.f90:69-70)
            // the homogeneous path keeps zdot only in its (z,k,l,m) table
            const void *zd = c->homog ? c->d_hz : c->d_vel[2];
            e = c->prec == RT6D_FP64
                    ? launch_bottom_forcing<double>(c->g, c->forcing, (double *)in, c->d_force, (const double *)zd,
                                                    c->homog, dt, dz1, c->stream)
                    : launch_bottom_forcing<float>(c->g, c->forcing, (float *)in, c->d_force, (const float *)zd,
                                                   c->homog, dt, dz1, c->stream);
            HIPCHK(e);
        }
        if (c->prec == RT6D_FP64) {
            StepPtrs<double> p = step_ptrs<double>(c, in, out, dt);
            e = c->opt_exact ? launch_step_exact(c->g, p, c->homog, lc)
                             : launch_step_fast<double>(c->g, p, c->homog, lc);
        } else {
            StepPtrs<float> p = step_ptrs<float>(c, in, out, dt);
            e = launch_step_fast<float>(c->g, p, c->homog, lc);
        }
        HIPCHK(e);
        c->cur ^= 1;
    }
    c->have_prev = true;
    c->prev_dt = dt;
    c->steps += nsteps;
    c->fl_src = c->de_src = 0;
    return 0;
}

int rt6d_step(rt6d_ctx *c, double dt, int nsteps)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = do_steps(c, dt, nsteps))) return r;
    if (c->opt_sync) HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int rt6d_step_timed(rt6d_ctx *c, double dt, int nsteps, double *ms)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = require_ready(c))) return r;
    if ((r = ensure_coef(c, dt))) return r;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if ((r = do_steps(c, dt, nsteps))) return r;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ms) *ms = (double)f;
    return 0;
}

int rt6d_fluxes(rt6d_ctx *c, double dt)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = require_ready(c))) return r;
    if ((r = run_tendency(c, c->d_E[c->cur], dt))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->fl_src = 2;
    c->de_src = 0;
    return 0;
}

int rt6d_download(rt6d_ctx *c, int field, int layout, double *out)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if (!out) return fail(RT6D_EINVAL, "rt6d_download: null output");
    const size_t cells = c->cells;
    if (field == RT6D_FIELD_E) {
        if (!c->E_set) return fail(RT6D_ESTATE, "rt6d_download: E not uploaded");
        if (layout == RT6D_LAYOUT_INTERIOR) return download_typed(c, c->d_E[c->cur], cells, out);
        if (layout == RT6D_LAYOUT_HALO) {
            std::vector<double> tmp(cells);
            if ((r = download_typed(c, c->d_E[c->cur], cells, tmp.data()))) return r;
            interior_to_halo(c->n, tmp.data(), out);
            return 0;
        }
        return fail(RT6D_EINVAL, "rt6d_download: layout %d for E", layout);
    }
    if (field == RT6D_FIELD_DE || field == RT6D_FIELD_MEANFL) {
        if ((r = require_ready(c))) return r;
        if (layout != RT6D_LAYOUT_INTERIOR && layout != RT6D_LAYOUT_HALO)
            return fail(RT6D_EINVAL, "rt6d_download: layout %d for dE/meanfl", layout);
        if ((r = ensure_step_byproducts(c, false))) return r;
        const void *src = field == RT6D_FIELD_DE ? c->d_tdE : c->d_tmf;
        if (layout == RT6D_LAYOUT_INTERIOR) return download_typed(c, src, cells, out);
        std::vector<double> tmp(cells);   // halo layout: the interior is written, the ghosts are left
        if ((r = download_typed(c, src, cells, tmp.data()))) return r;
        interior_to_halo(c->n, tmp.data(), out);
        return 0;
    }
    if (field >= RT6D_FIELD_FLUX_X && field < RT6D_FIELD_FLUX_X + 6) {
        if ((r = require_ready(c))) return r;
        const int d = field - RT6D_FIELD_FLUX_X;
        if (c->fl_src == 0 && (r = ensure_step_byproducts(c, true))) return r;
        if (layout == RT6D_LAYOUT_FACES) return download_typed(c, c->d_tfl[d], c->vel_count[d], out);
        if (layout == RT6D_LAYOUT_INTERIOR) {
            std::vector<double> tmp(c->vel_count[d]);
            if ((r = download_typed(c, c->d_tfl[d], c->vel_count[d], tmp.data()))) return r;
            faces_to_interior(c->n, d, tmp.data(), out);
            return 0;
        }
        return fail(RT6D_EINVAL, "rt6d_download: layout %d for a flux", layout);
    }
    if (!c->params_ready) return fail(RT6D_ESTATE, "rt6d_download: calc_parameter has not run");
    const ParamIn pi = param_in(c);
    if (field >= RT6D_FIELD_XDOT && field < RT6D_FIELD_XDOT + 6) {
        const int d = field - RT6D_FIELD_XDOT;
        if (layout != RT6D_LAYOUT_FACES && layout != RT6D_LAYOUT_INTERIOR)
            return fail(RT6D_EINVAL, "rt6d_download: layout %d for a velocity", layout);
        std::vector<double> tmp(c->vel_count[d]);
        if (c->d_vel[d]) {
            if ((r = download_typed(c, c->d_vel[d], c->vel_count[d], tmp.data()))) return r;
        } else {
            // freed on the homogeneous path: recompute (binary64, rounded to
            // the context's precision as the stored field was)
            double *v = nullptr;
            HIPCHK(hipMalloc((void **)&v, c->vel_count[d] * sizeof(double)));
            hipError_t e = launch_param_velocity(d, pi, v, c->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(tmp.data(), v, c->vel_count[d] * sizeof(double), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            hipFree(v);
            HIPCHK(e);
            if (c->prec == RT6D_FP32)
                for (double &x : tmp) x = (double)(float)x;
        }
        if (layout == RT6D_LAYOUT_FACES) memcpy(out, tmp.data(), tmp.size() * sizeof(double));
        else faces_to_interior(c->n, d, tmp.data(), out);
        return 0;
    }
    if (layout != RT6D_LAYOUT_INTERIOR) return fail(RT6D_EINVAL, "rt6d_download: field %d is interior-only", field);
    if (field == RT6D_FIELD_LAMBDA_SUM && c->d_lam) return download_typed(c, c->d_lam, cells, out);
    if (field == RT6D_FIELD_LAMBDA_SUM || (field >= RT6D_FIELD_LAMBDA_X && field < RT6D_FIELD_LAMBDA_X + 3)) {
        // recompute in binary64 from binary64 velocities (the sum rounded to
        // the context's precision as the stored field was)
        double *v64[kAxes] = {nullptr}, *tmp = nullptr;
        double *lsum = nullptr;
        hipError_t e = hipSuccess;
        for (int d = 0; d < 3 && e == hipSuccess; ++d) {
            e = hipMalloc((void **)&v64[d], c->vel_count[d] * sizeof(double));
            if (e == hipSuccess) e = launch_param_velocity(d, pi, v64[d], c->stream);
        }
        if (e == hipSuccess) e = hipMalloc((void **)&tmp, 3 * cells * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&lsum, cells * sizeof(double));
        if (e == hipSuccess) e = launch_param_lambda(pi, v64, lsum, tmp, c->stream);
        const double *src = field == RT6D_FIELD_LAMBDA_SUM ? lsum : tmp + (size_t)(field - RT6D_FIELD_LAMBDA_X) * cells;
        if (e == hipSuccess) e = hipMemcpyAsync(out, src, cells * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess && field == RT6D_FIELD_LAMBDA_SUM && c->prec == RT6D_FP32)
            for (size_t i = 0; i < cells; ++i) out[i] = (double)(float)out[i];
        for (int d = 0; d < 3; ++d) if (v64[d]) hipFree(v64[d]);
        if (tmp) hipFree(tmp);
        if (lsum) hipFree(lsum);
        HIPCHK(e);
        return 0;
    }
    if (field == RT6D_FIELD_OMEGA_I || field == RT6D_FIELD_OMEGA_E) {
        double *tmp = nullptr;
        HIPCHK(hipMalloc((void **)&tmp, cells * sizeof(double)));
        hipError_t e = launch_param_omega(pi, field == RT6D_FIELD_OMEGA_I ? tmp : nullptr,
                                          field == RT6D_FIELD_OMEGA_E ? tmp : nullptr, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, cells * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        hipFree(tmp);
        HIPCHK(e);
        return 0;
    }
    return fail(RT6D_EINVAL, "rt6d_download: unknown field %d", field);
}

// The diagnose kernels on buffers the caller provides: partial (nparts x
// kDiagSlots doubles), dout (16 doubles) and, with 3D outputs, d3 (8 x n3
// doubles); copies the results out and completes the reference's scalars.
static int diag_run(rt6d_ctx *c, double dt, rt6d_diag *out, double *E_int, double *meanfl_int, double *diss6,
                    double *partial, uint32_t nparts, double *dout, double *d3)
{
    const size_t n3 = (size_t)c->n[0] * c->n[1] * c->n[2];
    const bool want3 = E_int || meanfl_int || diss6;
    const void *Eprev = c->d_E[c->cur ^ 1];
    const void *Ecur = c->d_E[c->cur];
    double *Ei = want3 ? d3 : nullptr, *Mi = want3 ? d3 + n3 : nullptr, *Di = want3 ? d3 + 2 * n3 : nullptr;
    hipError_t e;
    if (c->prec == RT6D_FP64) {
        StepPtrs<double> p = step_ptrs<double>(c, Eprev, nullptr, c->prev_dt);
        e = c->opt_exact ? launch_diag_exact(c->g, p, c->homog, (const double *)Ecur, partial, nparts, dout, Ei, Mi,
                                             Di, c->stream)
                         : launch_diag_fast<double>(c->g, p, c->homog, (const double *)Ecur, partial, nparts, dout,
                                                    Ei, Mi, Di, c->stream);
    } else {
        StepPtrs<float> p = step_ptrs<float>(c, Eprev, nullptr, c->prev_dt);
        e = launch_diag_fast<float>(c->g, p, c->homog, (const float *)Ecur, partial, nparts, dout, Ei, Mi, Di,
                                    c->stream);
    }
    HIPCHK(e);
    double s[kDiagSlots];
    HIPCHK(hipMemcpyAsync(s, dout, sizeof s, hipMemcpyDeviceToHost, c->stream));
    if (want3) {
        if (E_int) HIPCHK(hipMemcpyAsync(E_int, Ei, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (meanfl_int) HIPCHK(hipMemcpyAsync(meanfl_int, Mi, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (diss6) HIPCHK(hipMemcpyAsync(diss6, Di, 6 * n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof *out);
    out->eint = s[0];
    out->mflint = s[1];
    out->mfl_max = s[2];
    for (int q = 0; q < 12; ++q) out->bfl[q] = s[3 + q];
    for (int d = 0; d < kAxes; ++d) {
        out->vmax[d] = c->vmax[d];
        double m = 0.0;
        for (int i = 1; i <= c->n[d]; ++i) {
            const double x = c->vmax[d] * dt / c->dt_h[d][i + 1];
            m = m > x ? m : x;
        }
        out->cfl[d] = m;
    }
    const double *b = out->bfl;
    out->residual = out->eint + out->mflint + b[1] + b[3] + b[5] + b[7] + b[9] + b[11] - b[0] - b[2] - b[4] - b[6] -
                    b[8] - b[10];
    return 0;
}

int rt6d_diagnose(rt6d_ctx *c, double dt, rt6d_diag *out, double *E_int, double *meanfl_int, double *diss6)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = require_ready(c))) return r;
    if (!out) return fail(RT6D_EINVAL, "rt6d_diagnose: null output");
    if (!c->have_prev) return fail(RT6D_ESTATE, "rt6d_diagnose: no step has been taken since E was set");
    if (!c->d_partial) {
        if ((r = c->dalloc((void **)&c->d_partial, (size_t)kDiagParts * kDiagSlots * sizeof(double)))) return r;
        c->nparts = kDiagParts;
    }
    if (!c->d_dout && (r = c->dalloc((void **)&c->d_dout, 16 * sizeof(double)))) return r;
    const size_t n3 = (size_t)c->n[0] * c->n[1] * c->n[2];
    const bool want3 = E_int || meanfl_int || diss6;
    if (want3 && !c->d_d3 && (r = c->dalloc((void **)&c->d_d3, 8 * n3 * sizeof(double)))) return r;
    if ((r = ensure_coef(c, c->prev_dt))) return r;
    return diag_run(c, dt, out, E_int, meanfl_int, diss6, c->d_partial, c->nparts, c->d_dout, c->d_d3);
}

int rt6d_diagnose_noalloc(rt6d_ctx *c, double dt, rt6d_diag *out, double *E_int, double *meanfl_int, double *diss6)
{
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = require_ready(c))) return r;
    if (!out) return fail(RT6D_EINVAL, "rt6d_diagnose_noalloc: null output");
    if (!c->have_prev) return fail(RT6D_ESTATE, "rt6d_diagnose_noalloc: no step has been taken since E was set");
    const size_t n3 = (size_t)c->n[0] * c->n[1] * c->n[2];
    if ((E_int || meanfl_int || diss6) && n3 > kNoAllocMaxN3)
        return fail(RT6D_EINVAL,
                    "rt6d_diagnose_noalloc: nx*ny*nz = %zu exceeds the %zu columns its scratch holds for the 3D "
                    "outputs (pass NULL for them, or use rt6d_diagnose)", n3, kNoAllocMaxN3);
    // the coefficients of the last step's dt are on the device from that step;
    // uploading others would take a host buffer
    if (c->coef_dt != c->prev_dt)
        return fail(RT6D_ESTATE, "rt6d_diagnose_noalloc: the device coefficients are not those of the last step's dt "
                    "(a later rt6d_fluxes changed them); step again or use rt6d_diagnose");
    double *base = c->device < kMaxDevices ? g_scratch[c->device] : nullptr;
    if (!base) return fail(RT6D_ESTATE, "rt6d_diagnose_noalloc: no scratch address for device %d", c->device);
    double *partial = base, *dout = base + (size_t)kDiagParts * kDiagSlots, *d3 = dout + 16;
    return diag_run(c, dt, out, E_int, meanfl_int, diss6, partial, kDiagParts, dout, d3);
}

int rt6d_copy_bandwidth(long long bytes, int reps, double *ms)
{
    if (bytes < 16 || reps < 1 || !ms) return fail(RT6D_EINVAL, "rt6d_copy_bandwidth: bad arguments");
    const size_t nb = (size_t)bytes & ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, nb);
    if (e == hipSuccess) e = hipMalloc(&b, nb);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, nb, s);
    if (e == hipSuccess) e = launch_copy(a, b, nb, s);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(a, b, nb, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) return fail(RT6D_EHIP, "rt6d_copy_bandwidth: %s", hipGetErrorString(e));
    *ms = (double)f / reps;
    return 0;
}

int rt6d_set_option(rt6d_ctx *c, int option, int value)
{
    if (!c) return fail(RT6D_EINVAL, "null context");
    if (int r = not_member(c, "rt6d_set_option")) return r;
    switch (option) {
    case RT6D_OPT_EXACT:
        if (value && c->prec != RT6D_FP64) return fail(RT6D_EINVAL, "exact arithmetic needs RT6D_FP64");
        c->opt_exact = value ? 1 : 0;
        c->fl_src = c->de_src = 0;
        return 0;
    case RT6D_OPT_PATH:
        if (value < 0 || value > 1) return fail(RT6D_EINVAL, "path option %d", value);
        c->opt_path = value;
        c->params_ready = false;  // takes effect at the next calc_parameter
        return 0;
    case RT6D_OPT_BLOCK:
        if (value < 64 || value > 1024 || (value & 63)) return fail(RT6D_EINVAL, "block %d", value);
        c->opt_block = value;
        return 0;
    case RT6D_OPT_KERNEL:
        if (value < 0 || value > 23) return fail(RT6D_EINVAL, "kernel option %d", value);
        if (kernel_option_removed(value))
            return fail(RT6D_EINVAL, "kernel option %d names a removed kernel (slower than the column family)", value);
        if ((value == 16 && c->n[2] % 2 != 0) || (value == 17 && c->n[2] % 4 != 0))
            return fail(RT6D_EINVAL, "pair kernel: nz (%d) not divisible by its z cells per thread", c->n[2]);
        if ((value == 5 || value == 10) && c->n[2] % 2 != 0)
            return fail(RT6D_EINVAL, "two z cells per thread need an even nz (have %d)", c->n[2]);
        if ((value == 6 || value == 11) && c->n[2] % 4 != 0)
            return fail(RT6D_EINVAL, "four z cells per thread need nz divisible by 4 (have %d)", c->n[2]);
        c->opt_kernel = value;
        return 0;
    case RT6D_OPT_SYNC:
        if (value < 0 || value > 1) return fail(RT6D_EINVAL, "sync option %d", value);
        c->opt_sync = value;
        return 0;
    default:
        return fail(RT6D_EINVAL, "unknown option %d", option);
    }
}

int rt6d_query(rt6d_ctx *c, int what, long long *value)
{
    if (!c || !value) return fail(RT6D_EINVAL, "rt6d_query: null argument");
    switch (what) {
    case RT6D_Q_HOMOGENEOUS: *value = c->homog ? 1 : 0; return 0;
    case RT6D_Q_DEVICE_BYTES: *value = (long long)c->dev_bytes; return 0;
    case RT6D_Q_CELLS: *value = (long long)c->cells; return 0;
    case RT6D_Q_STEPS: *value = c->steps; return 0;
    case RT6D_Q_UNIFORM_XY: *value = c->g.uxy & 3; return 0;
    case RT6D_Q_KERNEL:
        // the kernel a step runs: the dispatch's own resolution of the option
        *value = c->prec == RT6D_FP64 ? resolve_kernel<double>(c->g, c->homog, c->opt_kernel, c->ncu)
                                      : resolve_kernel<float>(c->g, c->homog, c->opt_kernel, c->ncu);
        return 0;
    case RT6D_Q_X_HANDOVER: {
        const int k = c->prec == RT6D_FP64 ? resolve_kernel<double>(c->g, c->homog, c->opt_kernel, c->ncu)
                                           : resolve_kernel<float>(c->g, c->homog, c->opt_kernel, c->ncu);
        const int zc = k == 21 ? 1 : 2;
        const bool wl = c->homog && (k == 21 || k == 22);
        *value = wl && (c->prec == RT6D_FP64 ? colx2_hands_over<double>(c->g, c->opt_exact, zc)
                                             : colx2_hands_over<float>(c->g, c->opt_exact, zc));
        return 0;
    }
    case RT6D_Q_BYTES_PER_CELL: {
        // E read + E write, plus the streamed coefficient fields on the general path
        double b = 2.0 * c->esz;
        if (!c->homog) {
            for (int d = 0; d < kAxes; ++d)
                if (c->n[d] > 1) b += (double)c->esz * (c->n[d] + 1) / c->n[d];
            b += c->esz;  // lambda
        }
        *value = (long long)llround(b * 1000.0);  // milli-bytes per cell
        return 0;
    }
    default: return fail(RT6D_EINVAL, "rt6d_query: unknown item %d", what);
    }
}

int rt6d_synchronize(rt6d_ctx *c)
{
    int r;
    if ((r = check_ctx(c))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}


// ================================================================= ensembles

}  // extern "C"

struct rt6d_ensemble {
    int prec = RT6D_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<rt6d_ctx *> slot;     // nullptr = free
    std::vector<double> dt;
    std::vector<int> cur0;            // member's current buffer at the last rebuild
    int blocks_per_cu = 0, order = 0, ncu = 256;
    int tiles = 0;                    // member tile shapes: RT6D_TILES_* of rt6d.h
    void *d_mem = nullptr;
    size_t mem_cap = 0;
    EnsTask *d_tasks = nullptr;
    size_t task_cap = 0;
    uint32_t ntasks = 0;
    uint32_t off[kEnsVariants] = {0}, cnt[kEnsVariants] = {0};
    long long nxal = 0;               // members whose step passes x faces between lanes
    uint32_t *d_ctr = nullptr;
    bool dirty = true;
    bool pending = false;             // steps enqueued, not yet waited for
    long long k = 0;                  // steps since the last rebuild
    long long steps = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // two steps starting at parity 0 / 1
};

namespace {

void ens_drop_graphs(rt6d_ensemble *e)
{
    for (int q = 0; q < 2; ++q)
        if (e->graph[q]) { hipGraphExecDestroy(e->graph[q]); e->graph[q] = nullptr; }
}

template <typename T>
int ens_rebuild_t(rt6d_ensemble *e)
{
    std::vector<EnsMember<T>> mem;
    std::vector<uint32_t> nbl;
    std::vector<int> var;
    std::vector<std::pair<size_t, int>> rowal_cand;   // (member, z cells per thread)
    bool rowal_all = true;
    for (size_t i = 0; i < e->slot.size(); ++i) {
        EnsMember<T> M{};
        rt6d_ctx *c = e->slot[i];
        int v = 0;
        if (c) {
            int zc = 1;
            if (!col_config<T>(c->g, M.cc, zc)) return fail(RT6D_EINVAL, "ensemble member too large");
            M.g = c->g;
            const int a = c->cur;
            M.p[0] = step_ptrs<T>(c, c->d_E[a], c->d_E[a ^ 1], e->dt[i]);
            M.p[1] = step_ptrs<T>(c, c->d_E[a ^ 1], c->d_E[a], e->dt[i]);
            const bool full = c->homog ? is_full<true>(c->g) : is_full<false>(c->g);
            // the member's tiles: the wave-local pair kernel where nx is even
            // and <= 32 (two z cells per thread where nz is even, unless the
            // ensemble's tiles ask for one), with uniform x/y spacings as
            // scalars; else quads in binary32 where they fit; else pairs
            // where nx is even, else the column kernel
            const bool wl = e->tiles != RT6D_TILES_PAIRS && colx2wl_fits(c->g, c->homog, 1);
            if (wl && e->tiles == RT6D_TILES_WAVE_LOCAL_1Z) zc = 1;
            const bool uxy = wl && full && (c->g.uxy & 3) == 3;
            const bool q4 = !wl && sizeof(T) == 4 && zc == 2 && colq_fits(c->g, c->homog, 2);
            const bool x2 = !wl && !q4 && colx2_fits(c->g, c->homog, zc);
            if (x2 || q4 || wl) M.cx = colx2_config(c->g, zc, q4 ? 4 : 2, false);   // x cells per thread: 4 or 2
            // binary64 wave-local members with uniform spacings that can deal
            // whole rows to wavefronts without one more of them (decided for
            // the whole ensemble below)
            if (uxy && sizeof(T) == 8 && colx2_rowal<T>(c->g, false, zc)) {
                rowal_cand.push_back({i, zc});
                rowal_all = rowal_all && colx2_config(c->g, zc, 2, true).xal;
            }
            v = (c->homog ? 1 : 0) | (zc == 2 ? 2 : 0) | (full ? 4 : 0) | (x2 ? 8 : 0) | (q4 ? 16 : 0) |
                (wl ? 32 : 0) | (uxy ? 64 : 0);
            M.variant = (uint32_t)v;
            e->cur0[i] = a;
            if (x2 || q4 || wl) {
                const uint64_t nthreads = (uint64_t)M.cx.nxyp * M.cx.nzb * c->g.n[3] * c->g.n[4];
                nbl.push_back((uint32_t)((nthreads + 255) / 256));
            } else {
                nbl.push_back(col_blocks<T>(c->g, M.cc));
            }
        } else {
            nbl.push_back(0);
        }
        var.push_back(v);
        mem.push_back(M);
    }
    // Whole rows per wavefront with the x faces passed between lanes (variant
    // bit 7), as the single-run kernel, where every such member can take it:
    // one member that cannot would split the step into two launches.
    e->nxal = rowal_all ? (long long)rowal_cand.size() : 0;
    if (rowal_all) {
        for (const auto &[i, zc] : rowal_cand) {
            const rt6d_ctx *c = e->slot[i];
            EnsMember<T> &M = mem[i];
            M.cx = colx2_config(c->g, zc, 2, true);
            var[i] |= 128;
            M.variant = (uint32_t)var[i];
            const uint64_t nthreads = (uint64_t)M.cx.nxyp * M.cx.nzb * c->g.n[3] * c->g.n[4];
            nbl[i] = (uint32_t)((nthreads + 255) / 256);
        }
    }
    // the work list, grouped by variant; within a variant member by member or interleaved
    std::vector<EnsTask> tasks;
    for (int v = 0; v < kEnsVariants; ++v) {
        e->off[v] = (uint32_t)tasks.size();
        if (e->order == 0) {
            for (size_t i = 0; i < nbl.size(); ++i)
                if (var[i] == v)
                    for (uint32_t b = 0; b < nbl[i]; ++b) tasks.push_back(EnsTask{(uint32_t)i, b});
        } else {
            uint32_t mx = 0;
            for (size_t i = 0; i < nbl.size(); ++i)
                if (var[i] == v && nbl[i] > mx) mx = nbl[i];
            for (uint32_t b = 0; b < mx; ++b)
                for (size_t i = 0; i < nbl.size(); ++i)
                    if (var[i] == v && b < nbl[i]) tasks.push_back(EnsTask{(uint32_t)i, b});
        }
        e->cnt[v] = (uint32_t)tasks.size() - e->off[v];
    }
    const size_t mb = mem.size() * sizeof(EnsMember<T>), tb = tasks.size() * sizeof(EnsTask);
    if (mb > e->mem_cap) {
        if (e->d_mem) hipFree(e->d_mem);
        e->d_mem = nullptr;
        e->mem_cap = 0;
        HIPCHK(hipMalloc(&e->d_mem, mb));
        e->mem_cap = mb;
    }
    if (tb > e->task_cap) {
        if (e->d_tasks) hipFree(e->d_tasks);
        e->d_tasks = nullptr;
        e->task_cap = 0;
        HIPCHK(hipMalloc((void **)&e->d_tasks, tb));
        e->task_cap = tb;
    }
    if (mb) HIPCHK(hipMemcpyAsync(e->d_mem, mem.data(), mb, hipMemcpyHostToDevice, e->stream));
    if (tb) HIPCHK(hipMemcpyAsync(e->d_tasks, tasks.data(), tb, hipMemcpyHostToDevice, e->stream));
    // both parity sets of work counters start at zero (outside any captured graph)
    HIPCHK(hipMemsetAsync(e->d_ctr, 0, kEnsCounters * sizeof(uint32_t), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->ntasks = (uint32_t)tasks.size();
    ens_drop_graphs(e);
    e->k = 0;
    e->dirty = false;
    return 0;
}

// write the ensemble's progress back into the member contexts
void ens_sync_members(rt6d_ensemble *e)
{
    for (size_t i = 0; i < e->slot.size(); ++i) {
        rt6d_ctx *c = e->slot[i];
        if (!c) continue;
        c->cur = e->cur0[i] ^ (int)(e->k & 1);
    }
}

int ens_launch(rt6d_ensemble *e, uint32_t parity)
{
    const uint32_t pb = e->blocks_per_cu > 0 ? (uint32_t)(e->blocks_per_cu * e->ncu) : 0u;
    hipError_t r;
    if (e->prec == RT6D_FP64)
        r = launch_ensemble<double>((const EnsMember<double> *)e->d_mem, e->d_tasks, e->off, e->cnt, parity, pb,
                                    e->d_ctr, e->stream);
    else
        r = launch_ensemble<float>((const EnsMember<float> *)e->d_mem, e->d_tasks, e->off, e->cnt, parity, pb,
                                   e->d_ctr, e->stream);
    HIPCHK(r);
    return 0;
}

int ens_graph(rt6d_ensemble *e, int start)
{
    if (e->graph[start]) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int r1 = ens_launch(e, (uint32_t)start);
    int r2 = r1 ? r1 : ens_launch(e, (uint32_t)(start ^ 1));
    hipError_t ec = hipStreamEndCapture(e->stream, &g);
    if (r2) {
        if (g) hipGraphDestroy(g);
        return r2;
    }
    HIPCHK(ec);
    hipError_t ei = hipGraphInstantiate(&e->graph[start], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIPCHK(ei);
    return 0;
}

// Wait for enqueued steps (if any) and settle the members' state.
int ens_settle(rt6d_ensemble *e)
{
    if (e->pending) {
        HIPCHK(hipStreamSynchronize(e->stream));
        e->pending = false;
    }
    ens_sync_members(e);
    return 0;
}

int ens_enqueue(rt6d_ensemble *e, int nsteps)
{
    int r;
    if (e->dirty) {
        if ((r = ens_settle(e))) return r;
        r = e->prec == RT6D_FP64 ? ens_rebuild_t<double>(e) : ens_rebuild_t<float>(e);
        if (r) return r;
    }
    if (e->ntasks == 0 || nsteps == 0) return 0;
    if (!e->pending) {
        // a member's dx/dt coefficients follow the last dt used on it (diagnose may use another)
        for (size_t i = 0; i < e->slot.size(); ++i)
            if (e->slot[i] && (r = ensure_coef(e->slot[i], e->dt[i]))) return r;
    }
    int left = nsteps;
    while (left >= 2) {
        const int start = (int)(e->k & 1);
        if ((r = ens_graph(e, start))) return r;
        HIPCHK(hipGraphLaunch(e->graph[start], e->stream));
        e->k += 2;
        left -= 2;
    }
    if (left == 1) {
        if ((r = ens_launch(e, (uint32_t)(e->k & 1)))) return r;
        e->k += 1;
    }
    e->pending = true;
    e->steps += nsteps;
    for (size_t i = 0; i < e->slot.size(); ++i) {
        rt6d_ctx *c = e->slot[i];
        if (!c) continue;
        c->steps += nsteps;
        c->have_prev = true;
        c->prev_dt = e->dt[i];
        c->fl_src = c->de_src = 0;
    }
    return 0;
}

// take c out of e (its context is being destroyed)
void ens_detach(rt6d_ensemble *e, rt6d_ctx *c)
{
    for (size_t i = 0; i < e->slot.size(); ++i)
        if (e->slot[i] == c) rt6d_ensemble_remove(e, (int)i);
}

}  // namespace

extern "C" {

int rt6d_ensemble_create(int precision, rt6d_ensemble **out)
{
    if (!out) return fail(RT6D_EINVAL, "rt6d_ensemble_create: null output");
    if (precision != RT6D_FP64 && precision != RT6D_FP32) return fail(RT6D_EINVAL, "precision %d", precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RT6D_EHIP, "rt6d_ensemble_create: no HIP device");
    rt6d_ensemble *e = new rt6d_ensemble();
    e->prec = precision;
    hipGetDevice(&e->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) == hipSuccess && prop.multiProcessorCount > 0)
        e->ncu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
        hipMalloc((void **)&e->d_ctr, kEnsCounters * sizeof(uint32_t)) != hipSuccess) {
        rt6d_ensemble_destroy(e);
        return fail(RT6D_EHIP, "rt6d_ensemble_create: stream/event/counter allocation failed");
    }
    *out = e;
    return 0;
}

int rt6d_ensemble_destroy(rt6d_ensemble *e)
{
    if (!e) return 0;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    e->pending = false;
    ens_sync_members(e);
    for (rt6d_ctx *c : e->slot)
        if (c) c->owner = nullptr;
    ens_drop_graphs(e);
    if (e->d_mem) hipFree(e->d_mem);
    if (e->d_tasks) hipFree(e->d_tasks);
    if (e->d_ctr) hipFree(e->d_ctr);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

int rt6d_ensemble_add(rt6d_ensemble *e, rt6d_ctx *c, double dt, int *slot)
{
    if (!e || !c) return fail(RT6D_EINVAL, "rt6d_ensemble_add: null argument");
    if (c->prec != e->prec) return fail(RT6D_EINVAL, "rt6d_ensemble_add: member precision differs");
    if (c->device != e->device) return fail(RT6D_EINVAL, "rt6d_ensemble_add: member on another device");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(RT6D_EINVAL, "rt6d_ensemble_add: dt = %g", dt);
    int r;
    if ((r = check_ctx(c))) return r;
    if ((r = require_ready(c))) return r;
    if (c->owner) return fail(RT6D_ESTATE, "rt6d_ensemble_add: context already in an ensemble");
    if (c->opt_exact) return fail(RT6D_EINVAL, "rt6d_ensemble_add: ensembles run the fast kernels (exact option set)");
    if (c->forcing) return fail(RT6D_EINVAL, "rt6d_ensemble_add: ensembles do not apply forcing (set_forcing kind 0 first)");
    ColCfg cc;
    int zc;
    const bool fits = e->prec == RT6D_FP64 ? col_config<double>(c->g, cc, zc) : col_config<float>(c->g, cc, zc);
    if (!fits) return fail(RT6D_EINVAL, "rt6d_ensemble_add: member too large for the ensemble kernel");
    // finish enqueued ensemble steps before the member set changes
    if ((r = ens_settle(e))) return r;
    if ((r = ensure_coef(c, dt))) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    int s = -1;
    for (size_t i = 0; i < e->slot.size(); ++i)
        if (!e->slot[i]) { s = (int)i; break; }
    if (s < 0) {
        s = (int)e->slot.size();
        e->slot.push_back(nullptr);
        e->dt.push_back(0.0);
        e->cur0.push_back(0);
    }
    e->slot[s] = c;
    e->dt[s] = dt;
    e->cur0[s] = c->cur ^ (int)(e->k & 1);  // so that ens_sync_members leaves c->cur as it is
    c->owner = e;
    e->dirty = true;
    if (slot) *slot = s;
    return 0;
}

int rt6d_ensemble_remove(rt6d_ensemble *e, int s)
{
    if (!e || s < 0 || s >= (int)e->slot.size() || !e->slot[s])
        return fail(RT6D_EINVAL, "rt6d_ensemble_remove: bad slot %d", s);
    int r;
    if ((r = ens_settle(e))) return r;
    e->slot[s]->owner = nullptr;
    e->slot[s] = nullptr;
    e->dirty = true;
    return 0;
}

int rt6d_ensemble_step_async(rt6d_ensemble *e, int nsteps)
{
    if (!e || nsteps < 0) return fail(RT6D_EINVAL, "rt6d_ensemble_step_async: bad arguments");
    HIPCHK(hipSetDevice(e->device));
    if (!e->pending) HIPCHK(hipEventRecord(e->ev0, e->stream));
    int r = ens_enqueue(e, nsteps);
    if (r) return r;
    if (e->pending) HIPCHK(hipEventRecord(e->ev1, e->stream));
    return 0;
}

int rt6d_ensemble_wait(rt6d_ensemble *e, double *ms)
{
    if (!e) return fail(RT6D_EINVAL, "rt6d_ensemble_wait: null ensemble");
    HIPCHK(hipSetDevice(e->device));
    double t = 0.0;
    if (e->pending) {
        HIPCHK(hipEventSynchronize(e->ev1));
        float f = 0.f;
        HIPCHK(hipEventElapsedTime(&f, e->ev0, e->ev1));
        t = (double)f;
    }
    int r = ens_settle(e);
    if (r) return r;
    if (ms) *ms = t;
    return 0;
}

int rt6d_ensemble_step(rt6d_ensemble *e, int nsteps, double *ms)
{
    int r;
    if ((r = rt6d_ensemble_step_async(e, nsteps))) return r;
    return rt6d_ensemble_wait(e, ms);
}

int rt6d_ensemble_set_schedule(rt6d_ensemble *e, int blocks_per_cu, int order)
{
    if (!e || blocks_per_cu < 0 || blocks_per_cu > 16 || order < 0 || order > 1)
        return fail(RT6D_EINVAL, "rt6d_ensemble_set_schedule: bad arguments");
    int r;
    if ((r = ens_settle(e))) return r;
    if (order != e->order) e->dirty = true;
    if (blocks_per_cu != e->blocks_per_cu) ens_drop_graphs(e);
    e->blocks_per_cu = blocks_per_cu;
    e->order = order;
    return 0;
}

int rt6d_ensemble_set_tiles(rt6d_ensemble *e, int tiles)
{
    if (!e || tiles < 0 || tiles > RT6D_TILES_PAIRS) return fail(RT6D_EINVAL, "rt6d_ensemble_set_tiles: bad arguments");
    int r;
    if ((r = ens_settle(e))) return r;
    if (tiles != e->tiles) e->dirty = true;
    e->tiles = tiles;
    return 0;
}

int rt6d_ensemble_tune(rt6d_ensemble *e, int trial_steps, int *bpc_out, int *order_out, int *tiles_out)
{
    if (!e || trial_steps < 2) return fail(RT6D_EINVAL, "rt6d_ensemble_tune: need >= 2 trial steps");
    const int bpcs[] = {0, 2, 4, 8};
    double best = 1e300;
    int bb = e->blocks_per_cu, bo = e->order, bt = e->tiles;
    for (int tiles = 0; tiles <= RT6D_TILES_PAIRS; ++tiles)
        for (int order = 0; order <= 1; ++order)
            for (int bpc : bpcs) {
                int r = rt6d_ensemble_set_tiles(e, tiles);
                if (!r) r = rt6d_ensemble_set_schedule(e, bpc, order);
                if (r) return r;
                double ms = 0.0;
                if ((r = rt6d_ensemble_step(e, 2, nullptr))) return r;   // build + warm the graph
                if ((r = rt6d_ensemble_step(e, trial_steps, &ms))) return r;
                if (ms < best) { best = ms; bb = bpc; bo = order; bt = tiles; }
            }
    int r = rt6d_ensemble_set_tiles(e, bt);
    if (!r) r = rt6d_ensemble_set_schedule(e, bb, bo);
    if (r) return r;
    if (bpc_out) *bpc_out = bb;
    if (order_out) *order_out = bo;
    if (tiles_out) *tiles_out = bt;
    return 0;
}

int rt6d_ensemble_query(rt6d_ensemble *e, int what, long long *value)
{
    if (!e || !value) return fail(RT6D_EINVAL, "rt6d_ensemble_query: null argument");
    int r;
    if (e->dirty) {
        HIPCHK(hipSetDevice(e->device));
        if ((r = ens_settle(e))) return r;
        r = e->prec == RT6D_FP64 ? ens_rebuild_t<double>(e) : ens_rebuild_t<float>(e);
        if (r) return r;
    }
    switch (what) {
    case RT6D_EQ_LAUNCHES: {
        long long n = 0;
        for (int v = 0; v < kEnsVariants; ++v) n += e->cnt[v] ? 1 : 0;
        *value = n;
        return 0;
    }
    case RT6D_EQ_TASKS: *value = (long long)e->ntasks; return 0;
    case RT6D_EQ_X_HANDOVER: *value = e->nxal; return 0;
    default: return fail(RT6D_EINVAL, "rt6d_ensemble_query: unknown query %d", what);
    }
}

int rt6d_ensemble_info(rt6d_ensemble *e, int *members, long long *cells, long long *steps)
{
    if (!e) return fail(RT6D_EINVAL, "null ensemble");
    int n = 0;
    long long cc = 0;
    for (rt6d_ctx *c : e->slot)
        if (c) { ++n; cc += (long long)c->cells; }
    if (members) *members = n;
    if (cells) *cells = cc;
    if (steps) *steps = e->steps;
    return 0;
}

}  // extern "C"
