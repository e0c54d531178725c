/*
 * qasmhd.h -- C ABI of the MI355X (gfx950) ideal-MHD step of qasMHD.
 *
 * The reference has no FFI: its hot path sits behind the protected driver
 * methods ideal(double dt), bound() and setdt(int) of MHD1D / MHD2D / MHD3D
 * (1D/SERIAL/shock/mhd1d_class.hpp:35-36, 2D/SERIAL/OTvortex/mhd2d_class.hpp:
 * 43-45, 3D/MPI/OTvortex/mhd3d_class.hpp:58-60), reached only from exec_()
 * (mhd2d_class.cpp:35-61).  A driver that swaps this library in overrides
 * those three methods with the calls below; its field
 * layout (common/mhd_class.hpp:32-42) and solver switches (RMN, ODR, R_K,
 * CTW of mymacros.hpp, dnxs/dnys/dnzs and stxs/stys/stzs) carry over as the
 * runtime fields of qas_grid / qas_scheme.
 *
 * Conventions
 *  - State = 8 SoA double arrays val[0..7] = {ro, mx, my, mz, bx, by, bz, en}
 *    of nx*ny*nz cells each, ghosts included, index ss = nx*(ny*k+j)+i;
 *    bx at the (i-1/2) face, by at (j-1/2), bz at (k-1/2).
 *  - Every call returns 0 on success or a negative QAS_E* code; the message
 *    of the last failure is qas_last_error().  Nothing throws across the ABI.
 *  - One host thread per context.  Calls are ordered on the context's HIP
 *    stream; qas_download and qas_cfl_dt synchronise, the others are
 *    asynchronous.
 */
#ifndef QASMHD_H
#define QASMHD_H

#ifdef __cplusplus
extern "C" {
#endif

#define QAS_OK 0
#define QAS_E_ARG (-1)      /* invalid argument / unsupported configuration */
#define QAS_E_HIP (-2)      /* HIP runtime error (message carries the HIP name) */
#define QAS_E_NOMEM (-3)    /* device allocation failed */
#define QAS_E_STATE (-4)    /* call out of order (e.g. step before upload) */

typedef struct qas_ctx qas_ctx;

/* Grid: cells per axis INCLUDE the ghosts, as nx/ny/nz of mhd*_class.hpp.
 * ndim 1: ny = nz = 1; ndim 2: nz = 1.  xoff/yoff/zoff must be 4 on the
 * active axes (mhd2d_class.hpp:10) and 0 on the others. */
typedef struct {
  int ndim;
  int nx, ny, nz;
  int xoff, yoff, zoff;
  double dx, dy, dz;
} qas_grid;

/* Scheme: the reference's compile-time switches as runtime fields. */
typedef struct {
  double gam;       /* specific-heat ratio (MHD::gam, mhd_class.hpp:32)       */
  double cfl;       /* CFL number (mymacros.hpp CFL)                         */
  int rmn;          /* Riemann solver: 0 Roe, 1 HLLD, 2 LHLLD, 3 MLAU (RMN)  */
  int odr;          /* spatial order 1..4 (ODR)                              */
  int r_k;          /* SSP-RK stages 1..3 (R_K)                              */
  int ctw;          /* CUCT upwind weighting on/off (CTW)                    */
  int dn[3][8];     /* boundary flags dnxs/dnys/dnzs: 0 periodic, +-1
                       Neumann/Dirichlet, +-2 open/zero-fix                 */
  int st[3][8];     /* staggering flags stxs/stys/stzs (bx:x, by:y, bz:z = 1) */
  int precision;    /* 0 = FP64 (bit-identical to the reference),
                       1 = mixed (FP32 reconstruction and fluxes; FP64 state,
                           RK stage sum and pressure recovery),
                       2 = FP64 with fused multiply-adds in the sweeps (same
                           arithmetic, one rounding per multiply-add; not
                           bit-identical; 2D and 3D)                        */
} qas_scheme;

/* Fill a scheme with the reference defaults: gam 5/3, cfl 0.4, HLLD, ODR 4,
 * R_K 3, CTW 1, periodic, stxs[4] = stys[5] = stzs[6] = 1, FP64. */
void qas_scheme_default(qas_scheme *s);

/* Allocate device state for one domain on the current HIP device.
 * Replaces MHD1D/2D/3D's constructor-side arrays (mhd2d_class.cpp:116-164)
 * on the device. */
int qas_create(qas_ctx **out, const qas_grid *g, const qas_scheme *s);
void qas_destroy(qas_ctx *c);
const char *qas_last_error(const qas_ctx *c);

/* Host <-> device copies of the state (val[] order above). */
int qas_upload(qas_ctx *c, const double *const val[8]);
int qas_download(qas_ctx *c, double *const val[8]);
/* Same for one contiguous host block of 8*nd doubles (variable-major). */
int qas_upload_packed(qas_ctx *c, const double *state);
int qas_download_packed(qas_ctx *c, double *state);
/* Device-to-device copies: `state` is a device pointer to 8*nd doubles. */
int qas_upload_device(qas_ctx *c, const double *state);
int qas_download_device(qas_ctx *c, double *state);

/* bound(val): ghost fill of the 8 variables -- replaces MHD2D::bound /
 * bc1d / bc2d / bc3d (common/common_func.c:34-278) and, on one device, the
 * MPI halo exchange mpi_sdrv3d + mpi_[xyz]bc3d (mpi/common_mpi.c:198-525). */
int qas_bound(qas_ctx *c);

/* ideal(dt): one SSP-RK step including the per-stage bound -- replaces
 * MHD1D/2D/3D::ideal (1D/SERIAL/shock/mhd1d_solve.cpp:3-124,
 * 2D/SERIAL/OTvortex/mhd2d_solve.cpp:3-400, 3D/MPI/OTvortex/mhd3d_solve.cpp:
 * 3-729). */
int qas_step(qas_ctx *c, double dt);

/* setdt(1): CFL time step of the current state (vmax floored at 1.0,
 * 2nd-order cell-centred B) -- replaces MHD*::setdt (mhd2d_class.cpp:11-33,
 * mhd1d_class.cpp:11-30, mhd3d_class.cpp:15-46).  Synchronises. */
int qas_cfl_dt(qas_ctx *c, double *dt_out);

/* exec_-style loop on the device (mhd2d_class.cpp:44-57 without I/O):
 * nsteps x {bound; ideal(dt); if flg and n even: dt = setdt}, with dt kept
 * in device memory between steps (no host synchronisation).  On entry dt is
 * *dt_inout if *dt_inout > 0, else setdt(1) of the current state; on exit
 * *dt_inout holds the current dt (this synchronises once).  dt_hist, if not
 * NULL, receives the dt used by each step.  The step counter n continues
 * across calls (n % 2 is taken on the running count). */
int qas_run(qas_ctx *c, int nsteps, int flg, double *dt_inout, double *dt_hist);
/* Running step count used by qas_run's n % 2 schedule; reset to 0 here. */
int qas_reset_counter(qas_ctx *c);
/* qas_run launches every kernel from the host by default; on = 1 replays
 * pairs of steps as one captured HIP graph instead (R_K 2 and 3, timing
 * hooks off), on = 0 goes back to host launches.  Results are identical;
 * on one MI355X host launches ran as fast as replay at 1024^2 and 1-2 %
 * faster at 256^2 and 512^2 (the host enqueues ahead of the GPU). */
int qas_use_graphs(qas_ctx *c, int on);

/* 2D launch shape (no counterpart in the reference: its loops have no
 * launch geometry): x-sweep wavefronts per workgroup (xwpb 1, 2, 4) and the
 * workgroup size of the element-wise kernels (ewb 64, 128, 256).  Every
 * shape gives the same results.  qas_tune times RK stage 1 of the uploaded
 * state in each of the 9 shapes and keeps the fastest (the state is not
 * changed); the first qas_run of a 2D context calls it unless a shape was
 * set.  qas_set_shape fixes a shape, qas_shape reports the current one
 * (default 4, 256).  1D and 3D have one shape. */
int qas_tune(qas_ctx *c);
int qas_set_shape(qas_ctx *c, int xwpb, int ewb);
int qas_shape(const qas_ctx *c, int *xwpb, int *ewb);

/* Block until all work queued on the context has finished. */
int qas_sync(qas_ctx *c);

/* Device memory held by the context (bytes). */
long long qas_device_bytes(const qas_ctx *c);

/* ---- allocation-free run (no counterpart in the reference) -------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Runs what
 * qas_run runs with graphs off -- the bound() before the first step, then
 * nsteps x {ideal(dt); if flg and n even: dt = setdt(1)} with n = nstep0 + 1
 * .. nstep0 + nsteps -- on a state the caller keeps in device memory:
 * d_state = 8 * nx*ny*nz doubles (variable-major val[] order, ghosts
 * included), read and updated in place.  The RK stage buffers and every
 * scratch array come from one static device array of
 * QAS_STATIC_SCRATCH_BYTES per GPU; a domain needing more is rejected with
 * QAS_E_ARG (roughly 2D up to 660^2 and 3D up to 60^3 interior cells).
 * dt > 0 is the first step's dt, dt <= 0 takes setdt(1) of the state.
 * d_dt_out (device pointer, may be NULL) receives the dt after the last
 * step.  stream: a hipStream_t, NULL for the null stream; the call only
 * enqueues (no synchronisation, no allocation, no hipMemset).  Messages go
 * to err (errlen bytes, may be NULL).  FP64 or mixed; no gravity, no
 * dissipation, no dt history.  All calls on one GPU share its array, whose
 * used part each call clears first: calls that overlap in time on one GPU
 * race, so order them (one stream). */
#define QAS_STATIC_SCRATCH_BYTES (128ull << 20)
int qas_run_static(const qas_grid *g, const qas_scheme *s, double *d_state, int nsteps, long long nstep0, int flg,
                   double dt, double *d_dt_out, void *stream, char *err, int errlen);

/* ---- dissipative step (2D, FP64) ---------------------------------------
 * qas_set_dissipation replaces the arrays DMHD2D keeps and DMHD2D::setdc
 * fills (common/mhd_class.hpp:38 nu, eta; 2D/SERIAL/MRX/dmhd2d_init_.cpp:86-103):
 * per-cell kinematic viscosity and resistivity, nx*ny doubles each (ghosts
 * included, host memory), and dcmax = max over cells of max(2 nu, eta), the
 * value setdc returns.  nu = NULL switches the dissipative step off.
 * With it set, qas_run follows DMHD2D::exec_ (dmhd2d_class.cpp:15-33): after
 * ideal(dt), ncmax = 1 + (int)(2 (dcmax dt) / dr^2) sub-steps of
 * {bound; dsptv(dt / ncmax)} (dt is read back each step; no graph replay).
 * qas_dsptv is one DMHD2D::dsptv(dt) call (dmhd2d_solve.cpp:3-151), after
 * bound() when the state is not bounded. */
int qas_set_dissipation(qas_ctx *c, const double *nu, const double *eta, double dcmax);
int qas_dsptv(qas_ctx *c, double dt);

/* ---- gravity (2D, FP64) -------------------------------------------------
 * qas_set_gravity gives the step the gravitational potential of the
 * reference's RTI driver (common/mhd_class.hpp:39 phi_g, filled by
 * 2D/SERIAL/RTI/mhd2d_init_.cpp:55): nx*ny doubles, ghosts included, host
 * memory; the state's energy includes ro phi.  With it, ideal(), bound()
 * and setdt() are those of 2D/SERIAL/RTI/mhd2d_{solve,class}.cpp: potential
 * round trips around the pressure recovery, work done by gravity in the
 * energy fluxes, momentum sources, hydrostatic energy ghosts along y.
 * phi = NULL switches gravity off. */
int qas_set_gravity(qas_ctx *c, const double *phi);


/* ---- batched domains (no counterpart in the reference) -----------------
 * Advances many independent 2D domains of uneven sizes with one launch per
 * stage kernel for all of them; a run of up to 256 steps is one launch of a
 * captured hipGraph (cached by run length; longer runs replay it).  Each
 * domain is a qas_ctx with its own grid, state and dt; all
 * must share rmn, odr, r_k (2 or 3), ctw, precision and gam, live on one
 * device, hold an uploaded state and have run the same number of steps.
 * qas_batch_run(b, n, flg, dt, hist) gives every domain i the result of
 * qas_run(ctx_i, n, flg, &dt[i], hist + i*n) (dt[i] <= 0: from the CFL
 * condition; hist may be NULL).  The contexts stay owned by the caller and
 * must outlive the batch; calls on one batch are not thread-safe.
 * Launch shape (autotuned): qas_batch_tune times RK stage 1 of the batch
 * in each candidate shape -- instance block order (0 as given, 1 largest
 * domain first), x-sweep wavefronts per workgroup (xwpb 2 or 4), element-
 * wise workgroup size (ewb 128 or 256) -- and keeps the fastest; the first
 * qas_batch_run tunes unless a shape was set.  Results are bit-identical in
 * every shape.  qas_batch_set_shape fixes one (xwpb 1/2/4, ewb 64/128/256);
 * qas_batch_shape returns it (1 when tuned or set, 0 for the default). */
typedef struct qas_batch qas_batch;
int qas_batch_create(qas_batch **b, qas_ctx *const *ctxs, int n);
int qas_batch_run(qas_batch *b, int nsteps, int flg, double *dt_inout, double *dt_hist);
int qas_batch_tune(qas_batch *b);
int qas_batch_set_shape(qas_batch *b, int order, int xwpb, int ewb);
int qas_batch_shape(const qas_batch *b, int *order, int *xwpb, int *ewb);
/* qas_batch_run replays pairs of steps of all domains as one captured HIP
 * graph (on = 1, the default) or launches every kernel from the host
 * (on = 0); the same results either way. */
int qas_batch_use_graphs(qas_batch *b, int on);
const char *qas_batch_last_error(const qas_batch *b);
void qas_batch_destroy(qas_batch *b);
/* Batched discrete adjoint (no counterpart in the reference): the adjoint
 * entry points below for every domain of a batch at once, each kernel
 * launched once for all domains (same scope as qas_adj_*, r_k 3).
 * qas_batch_adj_forward(b, n, dts): every domain i advances n steps with
 *   step sizes dts[i*n .. i*n+n-1], keeping its checkpoints (as
 *   qas_adj_forward on its context).
 * qas_batch_adj_backward(b, g_final, g_init): g_final[i], g_init[i] are host
 *   arrays (packed 8 x nd_i) of domain i, as for qas_adj_backward; every
 *   domain must hold the same number of recorded steps. */
int qas_batch_adj_forward(qas_batch *b, int nsteps, const double *dts);
int qas_batch_adj_backward(qas_batch *b, const double *const *g_final, double *const *g_init);


/* ---- discrete adjoint (no counterpart in the reference) ----------------
 * Reverse-mode derivative of `nsteps` steps of ideal(dt) (with the bound()
 * before the first step) at frozen step sizes.  Scope: 2D, FP64, HLLD,
 * LHLLD or MLAU (rmn 1-3), odr 4, r_k 3, ctw 0 or 1; HLLD with any boundary
 * kinds (periodic, open, mirror, staggered flags), LHLLD and MLAU periodic
 * on both axes; other schemes are rejected with QAS_E_ARG.
 * qas_adj_forward   advances the current state by nsteps steps with the
 *                   given dt values (as qas_step would), keeping each step's
 *                   input state on the device (8 x nd doubles per step).
 * qas_adj_backward  given the adjoint of the final state (packed 8 x nd,
 *                   ghost entries included: they are added onto the cells
 *                   bound() copies them from), returns the adjoint of the
 *                   state passed to the first step (packed 8 x nd; ghost
 *                   entries zero, as the forward overwrites them).  Host
 *                   arrays; the _device variant takes device pointers.
 *                   Non-smooth operations (WCNS weights, wave-speed and
 *                   HLLD region selections, MLAU polynomial branches and
 *                   shock sensor, positivity fallback) are
 *                   differentiated at the branch the forward takes. */
int qas_adj_forward(qas_ctx *c, int nsteps, const double *dts);
int qas_adj_backward(qas_ctx *c, const double *g_final, double *g_init);
int qas_adj_backward_device(qas_ctx *c, const double *g_final, double *g_init);
/* on = 1: the following qas_adj_forward calls keep, besides each step's
 * input state, every stage's intermediates (the two inner stage states,
 * the primitives and face EMF parts of all three stages: 46 x nd doubles
 * more per step), and qas_adj_backward reads them instead of recomputing
 * the three forward stages of each step.  Same results bit for bit; more
 * memory, less time.  on = 0 (the default): checkpoints only.  The batched
 * adjoint (qas_batch_adj_*) always recomputes. */
int qas_adj_keep_stages(qas_ctx *c, int on);


/* ---- instrumentation (no counterpart in the reference) -----------------
 * Kernel identifiers of one stage, for the timing hooks and the work model. */
#define QAS_K_PRIM 0      /* cell-centred B + primitives (2D: in SWEEP_Y)   */
#define QAS_K_SWEEP_X 1   /* reconstruction + Riemann + divergence along x  */
#define QAS_K_SWEEP_Y 2
#define QAS_K_SWEEP_Z 3
#define QAS_K_EMF 4       /* CUCT edge EMFs                                 */
#define QAS_K_CT 5        /* staggered-B update (2D: with the ghost fill of  */
                          /* the stage output)                              */
#define QAS_K_BOUND 6     /* ghost fill (2D: before the first step only)    */
#define QAS_K_CFL 7       /* CFL reduction                                  */
#define QAS_K_COUNT 8

/* Record HIP events around every launch of the kernels whose bit is set in
 * `mask` (1 << QAS_K_*), for up to `capacity` launches; resets earlier
 * records.  mask 0 turns recording off. */
int qas_prof_enable(qas_ctx *c, unsigned mask, int capacity);
/* Total device time (ms) and launch count recorded for kernel `kid`.
 * Synchronises. */
int qas_prof_query(qas_ctx *c, int kid, double *total_ms, long long *launches);
/* Algorithmic work of one launch of kernel `kid` for this context: flops
 * (+ - * / sqrt counted 1 each, the cheaper characteristic projection) and
 * compulsory HBM bytes. */
int qas_kernel_model(const qas_ctx *c, int kid, double *flops, double *bytes);
/* Per-point flop counts for (rmn, odr): [0] reconstruction (fast/Alfven),
 * [1] reconstruction (Alfven/entropy), [2] Riemann solver, [3] B-flux face
 * states, [4] one-sided linear interpolation, [5] 4th-order difference,
 * [6] face-to-centre, [7] CUCT weight. */
int qas_point_flops(int rmn, int odr, double out[8]);

/* Host-side dot-product check of a hand-written per-point adjoint against
 * central finite differences of its forward function, on `ntrials` random
 * inputs; writes one relative error per trial.  No device call. */
#define QAS_ADJ_WCNS4 0   /* WCNS-4 face values                             */
#define QAS_ADJ_LIN4 1    /* linear 4th-order face values                   */
#define QAS_ADJ_LRSTATE 2 /* characteristic reconstruction (ODR 4)          */
#define QAS_ADJ_LRFB 3    /* B-flux face states                             */
#define QAS_ADJ_CUCT 4    /* CUCT upwind weight                             */
#define QAS_ADJ_HLLD 5    /* HLLD flux                                      */
#define QAS_ADJ_MLAU 6    /* MLAU flux (incl. the shock-sensor inputs dv)   */
#define QAS_ADJ_LHLLD 7   /* LHLLD flux (incl. the shock-sensor inputs dv)  */
int qas_adjoint_selfcheck(int which, int ntrials, unsigned long long seed, double *relerr);

/* ---- batched per-point evaluation on the device (tests) -----------------
 * One GPU thread per item through the same device functions the stage
 * kernels use; host arrays in and out; each call allocates its own device
 * buffers.  They replace, for testing, direct calls of the reference's
 * per-point functions:
 *   qas_test_interp   interpol[ODR-1] (kind 0..3) / l_interp[ODR-1] (kind
 *                     10..13), common/common_interp.c:7-150; f5 = n x 5
 *                     stencil, fl at i+1/2, fr at i-1/2
 *   qas_test_f4       fcen[ODR-1] (op 0) / df1[ODR-1] (op 1),
 *                     common/common_interp.h:10-24; f4 = n x (i-1, i, i+1, i+2)
 *   qas_test_lrstate  mhd_lrstate, common/mhd_func.c:108-161; st = n x 7 x 5
 *                     (ro, vn, vt, vu, bt, bu, pr), bn = cell-centred normal B
 *   qas_test_flux     riemann[RMN], common/mhd_flux.c:7-698; n x 7 states,
 *                     dv = n x 2 shock sensor
 *   qas_test_cuct     mhd_cuct_weight, common/mhd_func.c:56-76; cells = n x
 *                     (ro[4], vx[4], vy[4]) of cells (i-1,j-1),(i,j-1),(i-1,j),
 *                     (i,j); faces = n x (bx(i,j-1), bx(i,j), by(i-1,j), by(i,j))
 *   qas_test_rcp      no reference counterpart: the in-range reciprocal the
 *                     WCNS weights use (fast) and the IEEE division 1.0 / d
 *                     (div), to show they agree bit for bit where the
 *                     kernels take the former (2^-767 <= |d| <= 2^1022) */
int qas_test_interp(int kind, int n, const double *f5, double *fl, double *fr);
int qas_test_f4(int op, int odr, int n, const double *f4, double *out);
int qas_test_lrstate(int odr, int n, const double *st, const double *bn, double gam, double *vl, double *vr);
int qas_test_flux(int rmn, int n, const double *ql, const double *qr, const double *bn, double gam,
                  const double *dv, double *f);
int qas_test_cuct(int n, const double *cells, const double *faces, double *out);
int qas_test_rcp(int n, const double *d, double *fast, double *div);

/* ---- host runs of the 2D ghost fill (tests; no GPU) ----------------------
 * The same __host__ __device__ functions the kernels use, run on the host
 * one thread after another; dn, st = 2 x 8 boundary flags (x, y), arrays
 * 8 x (nx * ny) doubles with ghosts, x fastest, nx, ny >= 16.
 *   qas_test_bound_host  bound() of f in place (common/common_func.c:34-138
 *                        via the k_bound body)
 *   qas_test_ctb2d_host  the fused CT + ghost-fill launch of one stage with
 *                        zero EMF (each CT face becomes a u0 + b in; u0 is
 *                        out itself when alias, as in the last RK stage);
 *                        the CT threads run first, or the ghost threads
 *                        (ghosts_first) */
int qas_test_bound_host(int nx, int ny, const int *dn, const int *st, double *f);
int qas_test_ctb2d_host(int nx, int ny, const int *dn, const int *st, const double *in, const double *u0,
                        double *out, double a, double b, int alias, int ghosts_first);

#ifdef __cplusplus
}
#endif
#endif
