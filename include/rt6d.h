/*
 * rt6d.h -- C ABI of the MI355X (gfx950) 6D wave-energy time step.
 *
 * The library holds one resident device copy of the reference model's state
 * (for_src/main_module.f90:41-77): E, the group velocities and refraction
 * rates, the mean-flow growth rate, the 1D grids and the background N^2, U, V.
 * Every entry point below stands in for one reference routine or for the
 * module state it works on; the reference routines are argument-less Fortran
 * subroutines over main_module globals, so the arguments here are exactly the
 * globals each one reads or writes.
 *
 * Conventions
 *  - Every call returns 0 on success and a negative RT6D_E* code on failure;
 *    rt6d_last_error() then gives a message (per host thread).  Nothing
 *    aborts; the Fortran bindings turn a non-zero status into halt_stop, as
 *    the reference does on fatal errors (for_src/parallel_mpi.f90:223-236).
 *  - Host arrays are borrowed for the duration of a call only (copied in or
 *    out).  Device buffers are owned by the context and carry no halos:
 *    ghost cells are index arithmetic inside the kernels.
 *  - Arrays are Fortran order (x fastest) over (x,y,z,k,l,m); axis numbers
 *    0..5 follow that order.  "Halo" layouts span (1-2 : n+2) per axis, the
 *    reference's onx=2 allocation (for_src/main_module.f90:114,135-136).
 *  - One context is used by one host thread at a time; calls on one context
 *    are serialised on the context's own HIP stream.
 */
#ifndef RT6D_H
#define RT6D_H

#ifdef __cplusplus
extern "C" {
#endif

#define RT6D_OK 0
#define RT6D_EINVAL (-1)   /* bad argument or call order */
#define RT6D_EHIP (-2)     /* HIP runtime error */
#define RT6D_ENOMEM (-3)   /* device allocation failed */
#define RT6D_ESTATE (-4)   /* required setup step missing */

/* precision of the step */
#define RT6D_FP64 0        /* E, fluxes and velocities in binary64 */
#define RT6D_FP32 1        /* E, fluxes, velocities in binary32; coefficients
                              derived in binary64, energy sums compensated binary64 */

/* field ids for rt6d_download */
#define RT6D_FIELD_E 0
#define RT6D_FIELD_DE 1
#define RT6D_FIELD_MEANFL 2
#define RT6D_FIELD_OMEGA_I 3
#define RT6D_FIELD_OMEGA_E 4
#define RT6D_FIELD_XDOT 5    /* 5..10: xdot ydot zdot kdot ldot mdot */
#define RT6D_FIELD_FLUX_X 11 /* 11..16: flux_x .. flux_m */
#define RT6D_FIELD_LAMBDA_X 17 /* 17..19: lambda_x lambda_y lambda_z */
#define RT6D_FIELD_LAMBDA_SUM 20 /* lambda summed over active x,y,z */

/* layouts for host arrays */
#define RT6D_LAYOUT_INTERIOR 0 /* n_x*...*n_m */
#define RT6D_LAYOUT_FACES 1    /* velocities/fluxes: 0..n_d along their own axis */
#define RT6D_LAYOUT_HALO 2     /* (n+4) per axis, onx = 2 */

/* options for rt6d_set_option */
#define RT6D_OPT_EXACT 1     /* 1: reference operation order, IEEE divides, no FMA.
                                0 (default): reciprocals, division-free limiter, FMA */
#define RT6D_OPT_PATH 2      /* 0 auto (default), 1 force the general path
                                (6D coefficient fields streamed from HBM) */
#define RT6D_OPT_BLOCK 3     /* threads per block of the per-cell step kernel (64..1024) */
#define RT6D_OPT_KERNEL 4    /* step kernel: 0 auto (default: on the homogeneous path with
                                nx even and <= 32, the pair kernel with the wave-local
                                neighbour exchange, two z cells per thread (option 22)
                                where nz is even and, in binary64, where that gives each
                                SIMD 2.5 wavefronts, else one (option 21); binary32 grids
                                that do not give each SIMD 2.5 wavefronts with one z cell
                                per thread, and binary64 rows wider than 32, run the pair
                                kernel (16/15); binary32 rows wider than 32 four x cells
                                per thread (option 19) where nx is divisible by 4; odd nx
                                the column kernel with two z cells per thread when nz is
                                even; the general path the column kernel with the y
                                neighbours staged in LDS);
                                1 per-cell kernel;
                                4/5/6 column kernel, 1/2/4 z cells per thread;
                                9/10/11 column kernel with the y neighbours staged in
                                LDS, 1/2/4 z cells per thread;
                                15/16/17 column kernel with two x cells per thread (paired
                                accesses), 1/2/4 z cells per thread: homogeneous path with
                                nx even;
                                18/19/20 column kernel with four x cells per thread
                                (16-byte accesses), 1/2/4 z cells per thread: binary32
                                homogeneous path with nx divisible by 4;
                                21/22/23 the pair kernel with the x and y neighbour pairs
                                exchanged through wave-local LDS slots: homogeneous path,
                                nx even and <= 32.
                                Options 2, 3, 7, 8, 12-14 and 24-26 named kernels that were
                                slower than these on every grid measured and were removed:
                                setting one is RT6D_EINVAL.
                                An option whose kernel does not fit the grid runs the
                                column kernel (two z cells per thread when nz is even), or
                                the per-cell kernel where one field reaches 2 GiB; the pair
                                kernels (15-17, 21-23) address E one m plane at a time and
                                fit until one m plane reaches 2 GiB */
#define RT6D_OPT_SYNC 5      /* 1 (default): rt6d_step returns once its steps are done;
                                0: rt6d_step returns once they are enqueued on the context's
                                stream, and the next call that reads the device state
                                (download, diagnose, rt6d_synchronize, ...) waits for them;
                                a fault of those steps is reported by that call */

/* queries for rt6d_query */
#define RT6D_Q_HOMOGENEOUS 1 /* 1 if the horizontally homogeneous path is active */
#define RT6D_Q_DEVICE_BYTES 2
#define RT6D_Q_CELLS 3
#define RT6D_Q_STEPS 4       /* steps taken since the last upload of E */
#define RT6D_Q_BYTES_PER_CELL 5 /* algorithmic HBM bytes per cell-update of the active step kernel */
#define RT6D_Q_UNIFORM_XY 7  /* bit 0 (1): every x (y) spacing is the same, so the wave-local
                                kernels take that axis's coefficients as scalars */
#define RT6D_Q_KERNEL 6      /* the kernel rt6d_step runs, as an RT6D_OPT_KERNEL option: the
                                set option or what the auto choice resolves to, after the
                                fallback for a grid the option does not fit (the dispatch's
                                own resolution; valid after calc_parameter, which decides
                                the homogeneous path) */
#define RT6D_Q_X_HANDOVER 8  /* 1 if that kernel is the wave-local pair kernel whose wavefronts
                                start at row starts and pass the x face between two pairs
                                from lane to lane (fast mode, uniform x/y spacing) */

typedef struct rt6d_ctx rt6d_ctx;

/* Scalars of diagnose.f90:45-329 (no netCDF). */
typedef struct rt6d_diag {
    double eint;       /* sum dE * vol */
    double mflint;     /* sum meanfl * vol */
    double mfl_max;    /* max |meanfl| / max(1e-12, E) */
    double vmax[6];    /* umax vmax wmax kmax lmax mmax */
    double bfl[12];    /* bfl_xs bfl_xe bfl_ys bfl_ye bfl_zs bfl_ze bfl_ks bfl_ke bfl_ls bfl_le bfl_ms bfl_me */
    double cfl[6];     /* cfl_x .. cfl_m */
    double residual;   /* eint + mflint + sum(bfl_e - bfl_s) */
} rt6d_diag;

/* Library build/version string. */
const char *rt6d_version(void);

/* Message for the last failing call on this host thread. */
const char *rt6d_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int rt6d_device_count(void);

/*
 * allocate_main_module (for_src/main_module.f90:131-238) + the switches of
 * set_parameter (main_module.f90:24-31).  n[6] = nx ny nz nk nl nm, each >= 1;
 * precision RT6D_FP64 or RT6D_FP32.  Uses the current HIP device.
 */
int rt6d_create(const int *n, int reflect_up, int reflect_down, int precision, rt6d_ctx **out);
int rt6d_destroy(rt6d_ctx *ctx);

/*
 * Host helper restating calc_grid for one axis (for_src/calc_grid.f90:104-177,
 * u_centered_grid :300-322).  spacing: n interior T spacings.  axis 0..5.
 * fixed: k_fixed/l_fixed/m_fixed for a wavenumber axis with n = 1.
 * Outputs t, dt, u, du: n+4 values each (halo'd), any may be NULL.
 */
int rt6d_build_grid(int axis, int n, const double *spacing, double fixed,
                    double *t, double *dt, double *u, double *du);

/* The 1D grid of one axis as calc_grid leaves it: T positions and T spacings,
 * n+4 values each (halo'd).  Replaces main_module's xt/dxt .. mt/dmt. */
int rt6d_set_grid(rt6d_ctx *ctx, int axis, const double *t, const double *dt);

/* set_initial_conditions' background: Nsqr, U, V as halo'd 3D arrays
 * ((nx+4)(ny+4)(nz+4)), f0 and sign_of_omega_i (main_module.f90:24-26,62-64). */
int rt6d_set_background(rt6d_ctx *ctx, const double *Nsqr, const double *U, const double *V,
                        double f0, double sign_of_omega_i);

/* calc_parameter (for_src/calc_parameter.f90:2-264) on the device. */
int rt6d_calc_parameter(rt6d_ctx *ctx);

/* dt = cfl / max_cell sum_d max(|v_d(c-1/2)|, |v_d(c+1/2)|) / dxt_d(c), over
 * active axes, reduced in binary64 on the device.  (No reference counterpart:
 * the reference takes dt from its config; diagnose.f90:294-313 only prints CFL.) */
int rt6d_cfl_dt(rt6d_ctx *ctx, double cfl, double *dt);

/* Set E (main_module.f90:41); layout INTERIOR or HALO (ghosts ignored, as
 * integrate refills them every step). */
int rt6d_upload_E(rt6d_ctx *ctx, const double *E, int layout);

/* E on the device from a separable product (no reference counterpart: the
 * reference's config plug-ins fill E on the host in set_initial_conditions,
 * e.g. for_config/gm1.f90).  factors = f_x(1..nx), f_y(1..ny), ..., f_m(1..nm)
 * (n0+...+n5 values); E = peak*fx*fy*fz*fk*fl*fm*(1 + noise*(2u-1)) with the
 * products taken left to right in binary64 and u = U[0,1) from splitmix64
 * keyed by (seed, member, linear cell index), then rounded to the context's
 * precision.  Replaces E like rt6d_upload_E. */
int rt6d_init_separable(rt6d_ctx *ctx, const double *factors, double peak, double noise,
                        unsigned long long seed, unsigned long long member);

/* Superbee branch flips between a binary64 context and a binary32 context of
 * the same problem (no reference counterpart: a precision study of the FP32
 * form).  For the upper face of every cell and each advected axis, the
 * limiter branch (phi = 0, 2r, 1, r or 2) that each context selects on its
 * own current state in its own arithmetic; faces[d] counts the faces
 * compared, flips[d] those whose branches differ. */
int rt6d_branch_flips(rt6d_ctx *fp64_ctx, rt6d_ctx *fp32_ctx, long long faces[6], long long flips[6]);

/* set_forcing on the device, applied to E before every step (for_src/main.f90:69
 * calls set_forcing before integrate).  kind 0 none; 1 bottom flux injection,
 * E(z=1) += dt*field/dzt(1) (for_config/lee1.f90:199-212); 2 bottom relaxation,
 * E(z=1) -= dt*max(0,zdot)*(E(z=1)-field)/dzt(1) (for_config/critical_gm1.f90:125-139).
 * field: 5D over (x,y,k,l,m), nx*ny*nk*nl*nm values, Fortran order; copied. */
int rt6d_set_forcing(rt6d_ctx *ctx, int kind, const double *field);

/* integrate (for_src/integrate.f90:4-271) called nsteps times with step dt. */
int rt6d_step(rt6d_ctx *ctx, double dt, int nsteps);

/* As rt6d_step; *ms receives the device time of the nsteps launches, from HIP
 * events recorded on the context's stream around them. */
int rt6d_step_timed(rt6d_ctx *ctx, double dt, int nsteps, double *ms);

/* superbee_fluxes (for_src/superbee_fluxes.f90:4-174) on the current E, with
 * integrate's ghost treatment; results readable with rt6d_download(FLUX_*). */
int rt6d_fluxes(rt6d_ctx *ctx, double dt);

/* Copy a field to the host.  E: current state.  DE, MEANFL, FLUX_*: those of
 * the last integrate (as the reference leaves them after integrate), or of
 * the last rt6d_fluxes call if it came later.  Velocities/lambdas/omegas:
 * calc_parameter outputs.  layout INTERIOR, FACES (velocities, fluxes) or
 * HALO (E, DE, MEANFL: the interior entries are written, ghost entries of the
 * host array are left as they were). */
int rt6d_download(rt6d_ctx *ctx, int field, int layout, double *out);

/*
 * diagnose reductions (for_src/diagnose.f90:45-329) for the state after the
 * last step: E is the current state, dE/meanfl/fluxes those of that step.
 * 3D outputs (nx*ny*nz each, may be NULL): E_int, meanfl_int, and diss6 =
 * 6 consecutive arrays diss_ks diss_ke diss_ls diss_le diss_ms diss_me.
 */
int rt6d_diagnose(rt6d_ctx *ctx, double dt, rt6d_diag *out,
                  double *E_int, double *meanfl_int, double *diss6);

/*
 * rt6d_diagnose without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  It
 * runs rt6d_diagnose's kernels, with the same results bit for bit, on a
 * file-scope device array of 4.5 MB per GPU (held until the process ends)
 * instead of buffers of the context: the per-block partial sums, the reduced
 * scalars and the staging of the 3D outputs.  Limits: with any 3D output
 * requested, nx*ny*nz <= 65536 (RT6D_EINVAL beyond); the last step's dt must
 * still be the context's coefficient dt (RT6D_ESTATE after an rt6d_fluxes
 * with another dt).  All calls on one GPU share its array: calls that
 * overlap on one GPU (other threads, other contexts) race; call it from one
 * host thread per GPU at a time.
 */
int rt6d_diagnose_noalloc(rt6d_ctx *ctx, double dt, rt6d_diag *out,
                          double *E_int, double *meanfl_int, double *diss6);

/* Measurement helper with no reference counterpart: copies `bytes` between
 * two scratch device buffers with a plain streaming kernel `reps` times and
 * returns the average kernel time (ms).  Gives the achievable copy bandwidth
 * and the reference point for the HBM byte counters of rocprofv3. */
int rt6d_copy_bandwidth(long long bytes, int reps, double *ms);

int rt6d_set_option(rt6d_ctx *ctx, int option, int value);
int rt6d_query(rt6d_ctx *ctx, int what, long long *value);
int rt6d_synchronize(rt6d_ctx *ctx);

/*
 * Ensembles (no reference counterpart: the reference runs one integration
 * per MPI job).  An ensemble advances many independent integrations -- each
 * a context set up as above (calc_parameter done, E uploaded), with its own
 * grid, background and dt -- with one launch per step and kernel variant
 * over a (member, block) work list (one block per task, or persistent blocks
 * taking tasks from work counters), two steps replayed from a captured HIP
 * graph.  Members whose grids take the same kernel variant share one launch,
 * so an ensemble of one variant steps with one launch per step.
 * Members keep their own state; while in an ensemble a context must not be
 * stepped on its own.  rt6d_ensemble_remove (or destroy) hands the context
 * back with its state current, so rt6d_download/rt6d_diagnose work on it.
 * All members share the ensemble's precision and device.
 */
typedef struct rt6d_ensemble rt6d_ensemble;
int rt6d_ensemble_create(int precision, rt6d_ensemble **out);
int rt6d_ensemble_destroy(rt6d_ensemble *ens);
/* add a member stepping with dt; *slot receives its slot number */
int rt6d_ensemble_add(rt6d_ensemble *ens, rt6d_ctx *member, double dt, int *slot);
/* take a member out; its context reflects every step the ensemble took */
int rt6d_ensemble_remove(rt6d_ensemble *ens, int slot);
/* advance every member nsteps steps; *ms (may be NULL) gets the device time */
int rt6d_ensemble_step(rt6d_ensemble *ens, int nsteps, double *ms);
/* the same in two halves: enqueue the steps on the ensemble's stream and
 * return, so that several ensembles run on the device at once; wait for
 * them (*ms: device time since the first enqueue).  Any call on a member
 * context first waits for its ensemble's enqueued steps. */
int rt6d_ensemble_step_async(rt6d_ensemble *ens, int nsteps);
int rt6d_ensemble_wait(rt6d_ensemble *ens, double *ms);
/* schedule: blocks_per_cu 0 runs one hardware block per (member, block) task
 * (default); 1..16 runs that many persistent blocks per compute unit, which
 * take tasks from per-XCD work counters.  order 0: the work list member by
 * member; 1: members interleaved block by block */
int rt6d_ensemble_set_schedule(rt6d_ensemble *ens, int blocks_per_cu, int order);
/* tiles: the shape of the work a thread takes in each member (the results
 * do not depend on it).  RT6D_TILES_WAVE_LOCAL (default): members with nx
 * even and <= 32 run the pair kernel with the wave-local neighbour exchange,
 * two z cells per thread where nz is even; RT6D_TILES_WAVE_LOCAL_1Z: the same
 * with one z cell per thread; RT6D_TILES_PAIRS: no wave-local exchange
 * (pairs, or quads in binary32).  Members the wave-local kernel does not fit
 * take quads (binary32, nx divisible by 4), pairs (nx even) or columns. */
#define RT6D_TILES_WAVE_LOCAL 0
#define RT6D_TILES_WAVE_LOCAL_1Z 1
#define RT6D_TILES_PAIRS 2
int rt6d_ensemble_set_tiles(rt6d_ensemble *ens, int tiles);
/* autotuner: times each tile choice x schedule (blocks_per_cu 0, 2, 4, 8 x
 * both orders) on 2 + trial_steps real steps of the current members -- the
 * members advance by 24*(2 + trial_steps) steps; results do not depend on
 * tiles or schedule -- and keeps the fastest (any output may be NULL) */
int rt6d_ensemble_tune(rt6d_ensemble *ens, int trial_steps, int *blocks_per_cu, int *order, int *tiles);
/* number of members, total cells, steps taken by the ensemble */
int rt6d_ensemble_info(rt6d_ensemble *ens, int *members, long long *cells, long long *steps);
/* queries for rt6d_ensemble_query (after bringing the work list up to date
 * with the current members and tiles) */
#define RT6D_EQ_LAUNCHES 1   /* kernel launches per ensemble step: one per kernel variant
                                among the members, nothing else (members whose grids take
                                the same variant share one launch) */
#define RT6D_EQ_TASKS 2      /* (member, block) tasks in the work list */
#define RT6D_EQ_X_HANDOVER 3 /* members stepped by the wave-local pair kernel with whole rows per
                                wavefront and the x faces passed between lanes (binary64, where
                                every wave-local member with uniform spacings can take it, so
                                that they keep one launch) */
int rt6d_ensemble_query(rt6d_ensemble *ens, int what, long long *value);

#ifdef __cplusplus
}
#endif
#endif /* RT6D_H */
