/*
 * qas_oracle.h -- CPU restatement of qasMHD's ideal-MHD step (TEST INFRASTRUCTURE).
 *
 * This header and qas_oracle.c are the parity oracle and the "port" CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load the library built from them; the product (qasmhd_amd) never
 * does.
 *
 * Field layout follows common/mhd_class.hpp:32-42: eight SoA double arrays
 * val[0..7] = {ro, mx, my, mz, bx, by, bz, en}, index ss = nx*(ny*k+j)+i with x
 * fastest, 4 ghost cells per side on every active axis, bx stored at the
 * (i-1/2) face, by at (j-1/2), bz at (k-1/2) (stxs[4]=stys[5]=stzs[6]=1).
 */
#ifndef QAS_ORACLE_H
#define QAS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int ndim;               /* 1, 2 or 3                                     */
  int nx, ny, nz;         /* cells per axis INCLUDING ghosts (1 if unused) */
  int xoff, yoff, zoff;   /* ghost layers (4 on active axes, 0 otherwise)  */
  double dx, dy, dz;
  double gam, cfl;
  int rmn;                /* 0 Roe, 1 HLLD, 2 LHLLD, 3 MLAU  (RMN)         */
  int odr;                /* spatial order 1..4                 (ODR)      */
  int r_k;                /* RK stages 1..3                     (R_K)      */
  int ctw;                /* CUCT upwind weight on/off          (CTW)      */
  int dn[3][8];           /* dnxs/dnys/dnzs boundary flags                 */
  int st[3][8];           /* stxs/stys/stzs staggering flags               */
  const double *phi;      /* 2D gravitational potential per cell, or NULL  */
} qor_cfg;

/* bound(val): ghost fill of the 8 variables (bc1d/bc2d/bc3d).               */
void qor_bound(const qor_cfg *c, double *const val[8]);
/* ideal(dt): one SSP-RK step with per-stage bound.                          */
void qor_step(const qor_cfg *c, double *const val[8], double dt);
/* setdt(1): CFL time step of the current state (floor 1.0 on the speed).    */
double qor_cfl_dt(const qor_cfg *c, double *const val[8]);
/* exec_-style loop: nsteps of {bound; ideal(dt); if flg and n even: setdt}.
 * *dt is the dt of the first step on entry and the current dt on exit.
 * dt_hist (may be NULL) receives the dt used by each step.                   */
void qor_run(const qor_cfg *c, double *const val[8], int nsteps, int flg,
             double *dt, double *dt_hist);

/* dsptv(dt): one visco-resistive SSP-RK step (2D), nu / eta per cell.      */
void qor_dsptv(const qor_cfg *c, double *const val[8], const double *nu, const double *eta, double dt);
/* DMHD2D::exec_ loop: ideal(dt) then ncmax sub-steps of dsptv(dt/ncmax).    */
void qor_run_dissipative(const qor_cfg *c, double *const val[8], const double *nu, const double *eta,
                         double dcmax, int nsteps, int flg, double *dt, double *dt_hist);

/* ---- per-function entry points (batched, for golden-vector tests) ------- */
/* kind 0..3: interpol[ODR-1] (1st, MUSCL-mm, WCNS3, WCNS4);
 * kind 10..13: l_interp[ODR-1] (linear 1st..4th).  f5: n x 5 stencil.       */
void qor_interp_batch(int kind, int n, const double *f5, double *fl, double *fr);
/* fcen[ODR-1] / df1[ODR-1] on n x 4 stencils (f[-1],f[0],f[1],f[2]).        */
void qor_fcen_batch(int odr, int n, const double *f4, double *out);
void qor_df_batch(int odr, int n, const double *f4, double *out);
/* mhd_lrstate: st is n x 7 x 5 (ro,vn,vt,vu,bt,bu,pr stencils), bn n.       */
void qor_lrstate_batch(int odr, int n, const double *st, const double *bn,
                       double gam, double *vl, double *vr);
/* calc_flux_*: ql,qr n x 7 (ro,vn,vt,vu,bt,bu,pr), bn n, dv n x 2 -> f n x 7 */
void qor_flux_batch(int rmn, int n, const double *ql, const double *qr,
                    const double *bn, double gam, const double *dv, double *f);
/* mhd_cuct_weight on n corners: c is n x 12 = ro[4], vx[4], vy[4] at the
 * corner's cells (i-1,j-1),(i,j-1),(i-1,j),(i,j); b is n x 4 =
 * bx(i,j-1), bx(i,j), by(i-1,j), by(i,j).                                   */
void qor_cuct_batch(int n, const double *c, const double *b, double *out);

#ifdef __cplusplus
}
#endif
#endif
