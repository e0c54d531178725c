/*
 * qas_oracle.c -- CPU restatement of qasMHD's per-step ideal-MHD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP path
 * and the "port" CPU baseline.  Nothing in the product loads it.
 *
 * It restates, function by function, the reference's arithmetic in the SAME
 * floating-point operation order (no FMA: build with -ffp-contract=off), so
 * that its FP64 results are bit-identical to the reference's own code:
 *   interpolation ........ common/common_interp.c:7-150, common_interp.h:6-24
 *   characteristic recon . common/mhd_eigen.c:175-434 (eigen_mi/ae),
 *                          :513-665 (m_reconst/a_reconst)
 *   lrstate / lr_fb ...... common/mhd_func.c:108-205
 *   Riemann solvers ...... common/mhd_flux.c:7-192 (MLAU), 194-433 (Roe),
 *                          435-561 (HLLD), 563-698 (LHLLD)
 *   CUCT weight .......... common/mhd_func.c:56-76
 *   updates .............. common/mhd_func.c:207-325, common_func.h:28-32
 *   boundaries ........... common/common_func.c:34-278
 *   stage sequences ...... 1D/SERIAL/shock/mhd1d_solve.cpp:3-124,
 *                          2D/SERIAL/OTvortex/mhd2d_solve.cpp:3-400,
 *                          3D/MPI/OTvortex/mhd3d_solve.cpp:3-729
 *   CFL .................. mhd1d_class.cpp:11-30, mhd2d_class.cpp:11-33,
 *                          mhd3d_class.cpp:15-46
 * Pinning: tests/test_oracle_golden.py compares every function here against
 * vectors written by the reference's own code (oracle/build.py), and the
 * 1D/2D steps against N-step reference runs, bit for bit.  The 3D step has no
 * buildable reference (its driver needs mpi.h); it is pinned by reduction to
 * the 2D reference and by invariants (tests/test_oracle_props.py); its CFL
 * (setdt of mhd3d_class.cpp:15-46) by closed forms on uniform states moving
 * along z with a z-staggered Bz, by rotation invariance and by an
 * independent numpy evaluation (same file).
 */
#include "qas_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ternary max/min of common/common_func.h:4-14 (NOT fmax/fmin) */
static inline double omax(double a, double b) { return (a >= b) ? a : b; }
static inline double omin(double a, double b) { return (a >= b) ? b : a; }
static inline double ominmod(double a, double b)
{
  return (+omin(omax(a, b), 0.0) + omax(omin(a, b), 0.0));
}

/* ---------------------------------------------------------------------- */
/* 1-D interpolation kernels.  f points at the centre of a 5-point stencil;
 * fl = left state at i+1/2, fr = right state at i-1/2.                     */
typedef void (*lr_fn)(const double *f, double *fl, double *fr);

static void flr1(const double *f, double *fl, double *fr)
{
  *fl = f[0];
  *fr = f[0];
}
static void flr2(const double *f, double *fl, double *fr)
{
  *fl = (-f[-1] + 5.0 * f[0] + 2.0 * f[+1]) / 6.0;
  *fr = (-f[+1] + 5.0 * f[0] + 2.0 * f[-1]) / 6.0;
}
static void flr3(const double *f, double *fl, double *fr)
{
  *fl = 0.125 * (-f[-1] + 6.0 * f[0] + 3.0 * f[+1]);
  *fr = 0.125 * (-f[+1] + 6.0 * f[0] + 3.0 * f[-1]);
}
static void flr4(const double *f, double *fl, double *fr)
{
  *fl = 0.003125 * (9.0 * f[-2] - 56.0 * f[-1] + 234.0 * f[0] + 144.0 * f[+1] - 11.0 * f[+2]);
  *fr = 0.003125 * (9.0 * f[+2] - 56.0 * f[+1] + 234.0 * f[0] + 144.0 * f[-1] - 11.0 * f[-2]);
}
static void muscl_mm(const double *f, double *fl, double *fr)
{
  double d = 0.5 * ominmod(f[0] - f[-1], f[+1] - f[0]);
  *fl = f[0] + d;
  *fr = f[0] - d;
}
static void wcns3(const double *f, double *fl, double *fr)
{
  const double eps = 1e-40;
  double a, b, t, s0, s1, w0, w1;
  a = f[-1] - f[0];
  s0 = a * a;
  a = f[+1] - f[0];
  s1 = a * a;
  t = f[+1] - 2.0 * f[0] + f[-1];
  t = t * t;
  s0 = (1.0 + t / (s0 + eps));
  s1 = (1.0 + t / (s1 + eps));
  /* linear weights 1/4, 3/4 */
  w0 = 0.25 * s0;
  w1 = 0.75 * s1;
  w0 /= (w0 + w1);
  w1 = 1.0 - w0;
  *fl = 0.5 * (w0 * (-f[-1] + 3.0 * f[0]) + w1 * (f[0] + f[+1]));
  w0 = 0.75 * s0;
  w1 = 0.25 * s1;
  w0 /= (w0 + w1);
  w1 = 1.0 - w0;
  *fr = 0.5 * (w1 * (-f[+1] + 3.0 * f[0]) + w0 * (f[0] + f[-1]));
}
static void wcns4(const double *f, double *fl, double *fr)
{
  const double eps = 1e-40;
  double a, b, s0, s1, s2, w0, w1, w2, den;
  a = f[-2] - 4.0 * f[-1] + 3.0 * f[0];
  b = f[-2] - 2.0 * f[-1] + f[0];
  s0 = (13. / 12.) * b * b + 0.25 * a * a;
  a = f[-1] - f[+1];
  b = f[-1] - 2.0 * f[0] + f[+1];
  s1 = (13. / 12.) * b * b + 0.25 * a * a;
  a = f[+2] - 4.0 * f[+1] + 3.0 * f[0];
  b = f[+2] - 2.0 * f[+1] + f[0];
  s2 = (13. / 12.) * b * b + 0.25 * a * a;
  s0 = 1.0 / (s0 * s0 + eps);
  s1 = 1.0 / (s1 * s1 + eps);
  s2 = 1.0 / (s2 * s2 + eps);
  /* left value at i+1/2: linear weights 0.075, 0.650, 0.275 */
  w0 = 0.075 * s0;
  w1 = 0.650 * s1;
  w2 = 0.275 * s2;
  den = 1.0 / (w0 + w1 + w2);
  w0 *= den;
  w1 *= den;
  w2 = 1.0 - w0 - w1;
  *fl = 0.125 * (+w0 * (3.0 * f[-2] - 10.0 * f[-1] + 15.0 * f[0])
                 + w1 * (-f[-1] + 6.0 * f[0] + 3.0 * f[+1])
                 + w2 * (+3.0 * f[0] + 6.0 * f[+1] - f[+2]));
  /* right value at i-1/2: mirrored weights */
  w0 = 0.275 * s0;
  w1 = 0.650 * s1;
  w2 = 0.075 * s2;
  den = 1.0 / (w0 + w1 + w2);
  w0 *= den;
  w1 *= den;
  w2 = 1.0 - w0 - w1;
  *fr = 0.125 * (+w2 * (3.0 * f[+2] - 10.0 * f[+1] + 15.0 * f[0])
                 + w1 * (-f[+1] + 6.0 * f[0] + 3.0 * f[-1])
                 + w0 * (+3.0 * f[0] + 6.0 * f[-1] - f[-2]));
}

/* face->centre and 1st difference; f points at the (i-1/2) value */
static double fcen2(const double *f) { return (0.5 * (f[0] + f[1])); }
static double fcen4(const double *f) { return (0.0625 * (9.0 * (f[0] + f[1]) - (f[-1] + f[2]))); }
static double df2(const double *f) { return (f[1] - f[0]); }
static double df4(const double *f) { return ((27.0 * (f[1] - f[0]) - (f[2] - f[-1])) / 24.0); }

typedef double (*c_fn)(const double *f);

/* selection tables of common/mhd_class.hpp:107-113 (index ODR-1) */
static lr_fn interp_tab[4] = {flr1, muscl_mm, wcns3, wcns4};
static lr_fn linear_tab[4] = {flr1, flr2, flr3, flr4};
static c_fn fcen_tab[4] = {fcen2, fcen2, fcen4, fcen4};
static c_fn df_tab[4] = {df2, df2, df4, df4};

/* ---------------------------------------------------------------------- */
/* Characteristic reconstruction.  s[0..6][0..4] = (ro, vn, vt, vu, bt, bu, pr)
 * stencils; bn = cell-centred normal B.  Returns primitive face states in
 * the same 7-variable order.  The eigenvector matrices are applied densely,
 * as in the reference, so the operation sequence is identical.             */
static void eig_fast_alfven(double ro, double bx, double by, double bz, double pr, double gamma,
                            double l[7][7], double r[7][7])
{
  /* mhd_eigen_mi: fast + Alfven modes */
  double iro = 1.0 / ro;
  double cs2 = gamma * pr * iro;
  double ca2 = bx * bx * iro;
  double ct2 = (by * by + bz * bz) * iro;
  double cb2 = cs2 + ca2 + ct2;
  double cf2 = 0.5 * (cb2 + sqrt(fabs(cb2 * cb2 - 4.0 * cs2 * ca2)));
  double cf = sqrt(cf2);
  double rcf = ro * cf;
  double sro = sqrt(ro);
  double isroh = 0.5 / sro;
  double icf2 = 1.0 / cf2;
  double icf2h = 0.5 * icf2;
  double ircfh = 0.5 / rcf;
  double byh = 0.5 * by;
  double bzh = 0.5 * bz;
  const double L[7][7] = {
      {0, -rcf, 0, 0, +by, +bz, +1.0},
      {cf2, 0, 0, 0, -by, -bz, -1.0},
      {0, +rcf, 0, 0, +by, +bz, +1.0},
      {0, 0, +sro, 0, +1.0, 0, 0},
      {0, 0, +sro, 0, -1.0, 0, 0},
      {0, 0, 0, +sro, 0, +1.0, 0},
      {0, 0, 0, +sro, 0, -1.0, 0}};
  const double R[7][7] = { /* R[k][n]: k-th right eigenvector, component n */
      {+icf2h, -ircfh, 0, 0, 0, 0, +0.5},
      {+icf2, 0, 0, 0, 0, 0, 0},
      {+icf2h, +ircfh, 0, 0, 0, 0, +0.5},
      {0, 0, +isroh, 0, +0.5, 0, -byh},
      {0, 0, +isroh, 0, -0.5, 0, +byh},
      {0, 0, 0, +isroh, 0, +0.5, -bzh},
      {0, 0, 0, +isroh, 0, -0.5, +bzh}};
  memcpy(l, L, sizeof(L));
  memcpy(r, R, sizeof(R));
}

static void eig_alfven_entropy(double ro, double bx, double by, double bz, double pr, double gamma,
                               double l[7][7], double r[7][7])
{
  /* mhd_eigen_ae: Alfven + entropy modes */
  double iro = 1.0 / ro;
  double cs2 = gamma * pr * iro;
  double ca2 = bx * bx * iro;
  double ct2 = (by * by + bz * bz) * iro;
  double cb2 = cs2 + ca2 + ct2;
  double cf2 = 0.5 * (cb2 + sqrt(fabs(cb2 * cb2 - 4.0 * cs2 * ca2)));
  double cf = sqrt(cf2);
  double sro = sqrt(ro);
  double isroh = 0.5 / sro;
  double icf2 = 1.0 / cf2;
  double byh = 0.5 * by;
  double bzh = 0.5 * bz;
  (void)cf;
  const double L[7][7] = {
      {0, +1.0, 0, 0, 0, 0, 0},
      {cf2, 0, 0, 0, -by, -bz, -1.0},
      {0, 0, 0, 0, +by, +bz, +1.0},
      {0, 0, +sro, 0, +1.0, 0, 0},
      {0, 0, +sro, 0, -1.0, 0, 0},
      {0, 0, 0, +sro, 0, +1.0, 0},
      {0, 0, 0, +sro, 0, -1.0, 0}};
  const double R[7][7] = {
      {0, +1.0, 0, 0, 0, 0, 0},
      {+icf2, 0, 0, 0, 0, 0, 0},
      {+icf2, 0, 0, 0, 0, 0, +1.0},
      {0, 0, +isroh, 0, +0.5, 0, -byh},
      {0, 0, +isroh, 0, -0.5, 0, +byh},
      {0, 0, 0, +isroh, 0, +0.5, -bzh},
      {0, 0, 0, +isroh, 0, -0.5, +bzh}};
  memcpy(l, L, sizeof(L));
  memcpy(r, R, sizeof(R));
}

static void char_reconst(int fast, const double s[7][5], double bn, double gamma, lr_fn flr,
                         double *ul, double *ur)
{
  double l[7][7], r[7][7], w[7][5], wl[7], wr[7];
  int m, n;
  if (fast)
    eig_fast_alfven(s[0][2], bn, s[4][2], s[5][2], s[6][2], gamma, l, r);
  else
    eig_alfven_entropy(s[0][2], bn, s[4][2], s[5][2], s[6][2], gamma, l, r);
  for (m = 0; m < 5; m++)
    for (n = 0; n < 7; n++)
      w[n][m] = (+l[n][0] * s[0][m] + l[n][1] * s[1][m] + l[n][2] * s[2][m] + l[n][3] * s[3][m]
                 + l[n][4] * s[4][m] + l[n][5] * s[5][m] + l[n][6] * s[6][m]);
  for (n = 0; n < 7; n++) flr(&w[n][2], &wl[n], &wr[n]);
  for (n = 0; n < 7; n++)
    ul[n] = r[0][n] * wl[0] + r[1][n] * wl[1] + r[2][n] * wl[2] + r[3][n] * wl[3]
            + r[4][n] * wl[4] + r[5][n] * wl[5] + r[6][n] * wl[6];
  for (n = 0; n < 7; n++)
    ur[n] = r[0][n] * wr[0] + r[1][n] * wr[1] + r[2][n] * wr[2] + r[3][n] * wr[3]
            + r[4][n] * wr[4] + r[5][n] * wr[5] + r[6][n] * wr[6];
  /* keep exactly-zero v and B exactly zero (vn, vt, vu, bt, bu) */
  for (n = 1; n <= 5; n++) {
    int zero = 1;
    for (m = 0; m < 5; m++)
      if (s[n][m] != 0.0) zero = 0;
    if (zero) ul[n] = ur[n] = 0.0;
  }
}

/* mhd_lrstate: supersonic stencil centre -> fast/Alfven projection, else
 * Alfven/entropy; then the positivity fallback. */
static void lrstate(const double s[7][5], double bn, double gamma, lr_fn flr, double *vl, double *vr)
{
  const double eps = 1e-12;
  int sup = (s[0][2] * (s[1][2] * s[1][2] + s[2][2] * s[2][2] + s[3][2] * s[3][2]) > gamma * s[6][2]);
  char_reconst(sup, s, bn, gamma, flr, vl, vr);
  if (vl[0] <= eps || vl[6] <= eps) {
    vl[0] = s[0][2];
    vl[1] = s[1][2];
    vl[6] = s[6][2];
  }
  if (vr[0] <= eps || vr[6] <= eps) {
    vr[0] = s[0][2];
    vr[1] = s[1][2];
    vr[6] = s[6][2];
  }
}

/* ---------------------------------------------------------------------- */
/* Riemann solvers.  q = (ro, vn, vt, vu, bt, bu, pr); F = (ro, mn, mt, mu,
 * bt, bu, en) fluxes; bnc = interface normal B; dv = shock sensor pair.     */
#define FL_EPS (1e-8)

static void flux_mlau(const double *L, const double *R, double bnc, double gamma, const double *dv, double *F)
{
  const double rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const double ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  double gammam1i = 1.0 / (gamma - 1.0);
  double bnc2 = bnc * bnc;
  double sgn = (bnc > 0) ? (1.0) : (-1.0);
  double bnca = sgn * bnc;
  double roli = 1.0 / rol;
  double vl2 = vnl * vnl + vtl * vtl + vul * vul;
  double pml = 0.5 * (btl * btl + bul * bul);
  double ptl = prl + pml;
  double enl = gammam1i * prl + pml + 0.5 * rol * vl2;
  double hnl = (enl + ptl) * roli;
  double rori = 1.0 / ror;
  double vr2 = vnr * vnr + vtr * vtr + vur * vur;
  double pmr = 0.5 * (btr * btr + bur * bur);
  double ptr = prr + pmr;
  double enr = gammam1i * prr + pmr + 0.5 * ror * vr2;
  double hnr = (enr + ptr) * rori;
  double cl2 = gamma * prl * roli;
  double cr2 = gamma * prr * rori;
  double cal2 = bnc2 * roli;
  double car2 = bnc2 * rori;
  double cbl2 = cl2 + cal2 + 2.0 * pml * roli;
  double cbr2 = cr2 + car2 + 2.0 * pmr * rori;
  double cfl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * cl2 * cal2)));
  double cfr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * cr2 * car2)));
  cbl2 = vl2 + cal2 + 2.0 * pml * roli;
  cbr2 = vr2 + car2 + 2.0 * pmr * rori;
  double ccl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * vl2 * cal2)));
  double ccr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * vr2 * car2)));
  double cmax = sqrt(omax(cfl2, cfr2));
  double cfn = cmax;
  double cfni = 1.0 / cfn;
  double sl = omin(0.0, omin(vnl, vnr) - cmax);
  double sr = omax(0.0, omax(vnl, vnr) + cmax);
  double mal = vnl * cfni;
  double mar = vnr * cfni;
  double snl = (mal > 0) ? 1.0 : (-1.0);
  double snr = (mar > 0) ? 1.0 : (-1.0);
  double theta = omin(1.0, (cfn - omin(dv[0], 0.0)) / (cfn - omin(dv[1], 0.0)));
  theta *= theta;
  theta *= theta;
  double alpha = 0.1875;
  double beta = 0.125;
  double malp1 = mal + 1.0;
  double marm1 = mar - 1.0;
  double mal2m1 = mal * mal - 1.0;
  double mar2m1 = mar * mar - 1.0;
  double bfl, ffl, bfr, ffr;
  if (snl * mal > 1.0) {
    bfl = 0.5 * (1.0 + snl);
    ffl = bfl * mal;
  } else {
    bfl = +0.25 * malp1 * malp1 * (2.0 - mal) + alpha * mal * mal2m1 * mal2m1;
    ffl = +0.25 * malp1 * malp1 + beta * mal2m1 * mal2m1;
  }
  if (snr * mar > 1.0) {
    bfr = 0.5 * (1.0 - snr);
    ffr = bfr * mar;
  } else {
    bfr = +0.25 * marm1 * marm1 * (2.0 + mar) - alpha * mar * mar2m1 * mar2m1;
    ffr = -0.25 * marm1 * marm1 - beta * mar2m1 * mar2m1;
  }
  double mnc = ffl + ffr;
  double mflux = mnc * cfn;
  mflux -= omax(1.0 - fabs(mnc), 0.0) * theta * cfni * (ptr - ptl) / (rol + ror);
  double dir = (mflux > 0) ? 1.0 : (-1.0);
  double dirl = 0.5 * (1.0 + dir);
  double dirr = 1.0 - dirl;
  mflux *= (dirl * rol + dirr * ror);
  double mfl = mflux * dirl;
  double mfr = mflux - mfl;
  double ptot = bfl * ptl + bfr * ptr;
  double cfmod = sqrt(omax(ccl2, ccr2));
  ptot -= 0.25 * bfl * bfr * (rol + ror) * cfmod * (vnr - vnl);
  ptot -= 0.5 * (1.0 - cfmod * cfni) * (bfl + bfr - 1.0) * (ptl + ptr);
  double slvl = sl - vnl;
  double srvr = sr - vnr;
  double snc = dirl * sl + dirr * sr;
  double vndnm = mflux + dirl * rol * slvl + dirr * ror * srvr;
  double vnc = ((fabs(vndnm) > FL_EPS) && (snc != 0)) ? (mflux * snc / vndnm) : (dirl * vnl + dirr * vnr);
  double slvc = sl - vnc;
  double srvc = sr - vnc;
  double slvci = 1.0 / slvc;
  double srvci = 1.0 / srvc;
  double rhdl = rol * slvl * slvc - bnc2;
  double rhdr = ror * srvr * srvc - bnc2;
  double dvtl = 0.0, dvul = 0.0, dbtl = 0.0, dbul = 0.0;
  double dvtr = 0.0, dvur = 0.0, dbtr = 0.0, dbur = 0.0;
  if (fabs(rhdl) > FL_EPS) {
    double q = (vnc - vnl) / rhdl;
    dvtl = -bnc * btl * q;
    dvul = -bnc * bul * q;
    dbtl = bnc2 * btl * q * slvci;
    dbul = bnc2 * bul * q * slvci;
  }
  if (fabs(rhdr) > FL_EPS) {
    double q = (vnc - vnr) / rhdr;
    dvtr = -bnc * btr * q;
    dvur = -bnc * bur * q;
    dbtr = bnc2 * btr * q * srvci;
    dbur = bnc2 * bur * q * srvci;
  }
  double cmpl = slvl * slvci;
  double cmpr = srvr * srvci;
  double ro2l = rol * cmpl;
  double srol = sqrt(ro2l);
  double vt2l = dvtl + vtl;
  double vu2l = dvul + vul;
  double bt2l = dbtl + btl * cmpl;
  double bu2l = dbul + bul * cmpl;
  double ro2r = ror * cmpr;
  double sror = sqrt(ro2r);
  double vt2r = dvtr + vtr;
  double vu2r = dvur + vur;
  double bt2r = dbtr + btr * cmpr;
  double bu2r = dbur + bur * cmpr;
  double sroi = 1.0 / (srol + sror);
  double srollri = srol * sroi;
  double srorlri = sror * sroi;
  double bnflb = sgn * omin(omax(0.0, srollri * (bnca + sror * vnc)), bnca);
  double bnfrb = sgn * omin(omax(0.0, srorlri * (bnca - srol * vnc)), bnca);
  double bnflv = sgn * omin(omax(0.0, srorlri * (bnca + mflux / sror)), bnca);
  double bnfrv = sgn * omin(omax(0.0, srollri * (bnca - mflux / srol)), bnca);
  double bnsrv = omax(bnca - dir * vnc * (dirl * srol + dirr * sror), 0.0);
  double ccb = sroi * bnsrv;
  double ccv = srol * sror * ccb;
  double bnvt = bnflb * vt2l + bnfrb * vt2r + ccb * (bt2r - bt2l);
  double bnvu = bnflb * vu2l + bnfrb * vu2r + ccb * (bu2r - bu2l);
  double bnbt = bnflv * bt2l + bnfrv * bt2r + ccv * (vt2r - vt2l);
  double bnbu = bnflv * bu2l + bnfrv * bu2r + ccv * (vu2r - vu2l);
  bnvt -= vnc * (dirl * dbtl + dirr * dbtr);
  bnvu -= vnc * (dirl * dbul + dirr * dbur);
  bnbt -= mflux * (dirl * dvtl + dirr * dvtr);
  bnbu -= mflux * (dirl * dvul + dirr * dvur);
  double fro = mflux;
  double fmn = mfl * vnl + mfr * vnr + ptot - 0.5 * bnc2;
  double fmt = mfl * vtl + mfr * vtr - bnbt;
  double fmu = mfl * vul + mfr * vur - bnbu;
  double fbt = mfl * btl * roli + mfr * btr * rori - bnvt;
  double fbu = mfl * bul * roli + mfr * bur * rori - bnvu;
  double vb1 = dirl * (vtl * btl + vul * bul) + dirr * (vtr * btr + vur * bur);
  double vb2 = dirl * (vt2l * bt2l + vu2l * bu2l) + dirr * (vt2r * bt2r + vu2r * bu2r);
  double rv2b2 = (mflux * vnc - bnc2);
  rv2b2 *= rv2b2;
  double vb3 = 0.0;
  if (rv2b2 > FL_EPS * FL_EPS) {
    vb3 = (+(fmt * vnc + fbt * bnc) * (fmt * bnc + fbt * mflux)
           + (fmu * vnc + fbu * bnc) * (fmu * bnc + fbu * mflux)) / rv2b2;
  }
  double bnvb = sgn * (bnca * (snc * vb2 - vnc * vb1) * (dirl * slvci + dirr * srvci) + bnsrv * (vb3 - vb2));
  F[0] = fro;
  F[1] = fmn;
  F[2] = fmt;
  F[3] = fmu;
  F[4] = fbt;
  F[5] = fbu;
  F[6] = mfl * hnl + mfr * hnr - bnvb;
}

static void flux_roe(const double *L, const double *R, double bnc, double gamma, const double *dv, double *F)
{
  const double rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const double ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  (void)dv;
  double gammam1 = gamma - 1.0;
  double gammam2 = gamma - 2.0;
  double gammam1i = 1.0 / gammam1;
  double roli = 1.0 / rol;
  double rori = 1.0 / ror;
  double rrol = sqrt(rol);
  double rror = sqrt(ror);
  double rroi = 1.0 / (rrol + rror);
  double robar = rrol * rror;
  double robari = 1.0 / robar;
  double rrobar = sqrt(robar);
  double rrobai = sqrt(robari);
  double vnbar = (rrol * vnl + rror * vnr) * rroi;
  double vtbar = (rrol * vtl + rror * vtr) * rroi;
  double vubar = (rrol * vul + rror * vur) * rroi;
  double bnbar = bnc;
  double btbar = (rrol * btr + rror * btl) * rroi;
  double bubar = (rrol * bur + rror * bul) * rroi;
  double hl = (+gamma * gammam1i * prl * roli + 0.5 * (vnl * vnl + vtl * vtl + vul * vul)
               + (bnbar * bnbar + btl * btl + bul * bul) * roli);
  double hr = (+gamma * gammam1i * prr * rori + 0.5 * (vnr * vnr + vtr * vtr + vur * vur)
               + (bnbar * bnbar + btr * btr + bur * bur) * rori);
  double hbar = (rrol * hl + rror * hr) * rroi;
  double btave = 0.5 * (btl + btr);
  double buave = 0.5 * (bul + bur);
  double sgn = (bnbar > 0) ? (1.0) : (-1.0);
  double db2 = 0.5 * gammam2 * gammam1i * ((btr - btl) * (btr - btl) + (bur - bul) * (bur - bul)) * rroi * rroi;
  double c2 = gammam1 * (+hbar - 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar) - db2
                         - (bnbar * bnbar + btbar * btbar + bubar * bubar) * robari);
  double a2 = (gammam1 * (+hbar - 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar) - db2)
               - gammam2 * (bnbar * bnbar + btbar * btbar + bubar * bubar) * robari);
  double ca2 = bnbar * bnbar * robari;
  double cb2 = (btbar * btbar + bubar * bubar) * robari;
  double cf2 = 0.5 * (a2 + sqrt(cb2 * (a2 + c2 + ca2) + (c2 - ca2) * (c2 - ca2)));
  double cs2 = c2 * ca2 / cf2;
  double cf = sqrt(cf2);
  double cs = sqrt(cs2);
  double ca = sqrt(ca2);
  double c = sqrt(c2);
  double cfi = 1.0 / cf;
  double sgr1 = btbar * btbar + bubar * bubar - FL_EPS;
  double sf11 = (sgr1 > 0) ? 1.0 : 0.0;
  double sf21 = 1.0 - sf11;
  double beti = 1.0 / sqrt(btbar * btbar + bubar * bubar + sf21);
  double bett = sf11 * btbar * beti + sf21 * sqrt(0.5);
  double betu = sf11 * bubar * beti + sf21 * sqrt(0.5);
  double sgr2 = (btbar * btbar + bubar * bubar) * robari + fabs(ca2 - c2) - FL_EPS;
  double sf12 = (sgr2 > 0) ? 1.0 : 0.0;
  double sf22 = 1.0 - sf12;
  double cfca = omax(0.0, cf2 - ca2);
  double cfcs = omax(0.0, cf2 - cs2);
  double cfa = omax(0.0, cf2 - c2);
  double alpi = 1.0 / sqrt(cfcs + sf22);
  double alpf = sf12 * sqrt(cfca) * alpi + sf22;
  double alps = sf12 * sqrt(cfa) * alpi;
  double eeps = 0.0;
  double evpf = omax(fabs(vnbar + cf), eeps);
  double evmf = omax(fabs(vnbar - cf), eeps);
  double evps = omax(fabs(vnbar + cs), eeps);
  double evms = omax(fabs(vnbar - cs), eeps);
  double evpa = omax(fabs(vnbar + ca), eeps);
  double evma = omax(fabs(vnbar - ca), eeps);
  double ev00 = omax(fabs(vnbar), eeps);
  double dltro = ror - rol;
  double dltmn = robar * (vnr - vnl);
  double dltmt = robar * (vtr - vtl);
  double dltmu = robar * (vur - vul);
  double dltbt = btr - btl;
  double dltbu = bur - bul;
  double t1 = bett * dltbt + betu * dltbu;
  double t2 = (prr - prl + (btave * dltbt + buave * dltbu) + gammam2 * (btbar * dltbt + bubar * dltbu)) * gammam1i;
  double t3 = betu * dltbt - bett * dltbu;
  double s1 = dltmn;
  double s2 = bett * dltmt + betu * dltmu;
  double s3 = betu * dltmt - bett * dltmu;
  double p11 = alps * cf * rrobai;
  double p12 = -alpf * c2 * cfi * rrobai;
  double p21 = alpf * (cf2 - c2 * gammam2 * gammam1i);
  double p22 = alps * (cs2 - c2 * gammam2 * gammam1i);
  double q11 = alpf * cf;
  double q12 = alps * cs;
  double q21 = -alps * ca * sgn;
  double q22 = alpf * c * sgn;
  double detpi = 1.0 / (p11 * p22 - p12 * p21);
  double detqi = 1.0 / (q11 * q22 - q12 * q21);
  double wvpf = 0.5 * ((p22 * t1 - p12 * t2) * detpi + (q22 * s1 - q12 * s2) * detqi);
  double wvmf = 0.5 * ((p22 * t1 - p12 * t2) * detpi - (q22 * s1 - q12 * s2) * detqi);
  double wvps = 0.5 * ((-p21 * t1 + p11 * t2) * detpi + (-q21 * s1 + q11 * s2) * detqi);
  double wvms = 0.5 * ((-p21 * t1 + p11 * t2) * detpi - (-q21 * s1 + q11 * s2) * detqi);
  double wvpa = 0.5 * (rrobar * t3 - sgn * s3);
  double wvma = 0.5 * (rrobar * t3 + sgn * s3);
  double wv00 = dltro - alpf * (wvpf + wvmf) - alps * (wvps + wvms);
  double frol = rol * vnl;
  double fmnl = rol * vnl * vnl + prl + 0.5 * (-bnbar * bnbar + btl * btl + bul * bul);
  double fmtl = rol * vnl * vtl - bnbar * btl;
  double fmul = rol * vnl * vul - bnbar * bul;
  double fbtl = vnl * btl - vtl * bnbar;
  double fbul = vnl * bul - vul * bnbar;
  double fenl = rol * vnl * hl - bnbar * (bnbar * vnl + btl * vtl + bul * vul);
  double fror = ror * vnr;
  double fmnr = ror * vnr * vnr + prr + 0.5 * (-bnbar * bnbar + btr * btr + bur * bur);
  double fmtr = ror * vnr * vtr - bnbar * btr;
  double fmur = ror * vnr * vur - bnbar * bur;
  double fbtr = vnr * btr - vtr * bnbar;
  double fbur = vnr * bur - vur * bnbar;
  double fenr = ror * vnr * hr - bnbar * (bnbar * vnr + btr * vtr + bur * vur);
  double v2b = 0.5 * (vnbar * vnbar + vtbar * vtbar + vubar * vubar);
  double rvpfro = alpf;
  double rvpfmn = alpf * (vnbar + cf);
  double rvpfmt = alpf * vtbar - alps * bett * ca * sgn;
  double rvpfmu = alpf * vubar - alps * betu * ca * sgn;
  double rvpfbt = alps * bett * cf * rrobai;
  double rvpfbu = alps * betu * cf * rrobai;
  double rvpfen = alpf * (v2b + db2 + cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i)
                  - alps * ca * (bett * vtbar + betu * vubar) * sgn;
  double rvmfro = alpf;
  double rvmfmn = alpf * (vnbar - cf);
  double rvmfmt = alpf * vtbar + alps * bett * ca * sgn;
  double rvmfmu = alpf * vubar + alps * betu * ca * sgn;
  double rvmfbt = rvpfbt;
  double rvmfbu = rvpfbu;
  double rvmfen = alpf * (v2b + db2 - cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i)
                  + alps * ca * (bett * vtbar + betu * vubar) * sgn;
  double rvpsro = alps;
  double rvpsmn = alps * (vnbar + cs);
  double rvpsmt = alps * vtbar + c * sgn * alpf * bett;
  double rvpsmu = alps * vubar + c * sgn * alpf * betu;
  double rvpsbt = -rrobai * c2 * alpf * bett * cfi;
  double rvpsbu = -rrobai * c2 * alpf * betu * cfi;
  double rvpsen = alps * (v2b + db2 + cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i)
                  + alpf * c * (bett * vtbar + betu * vubar) * sgn;
  double rvmsro = alps;
  double rvmsmn = alps * (vnbar - cs);
  double rvmsmt = alps * vtbar - c * sgn * alpf * bett;
  double rvmsmu = alps * vubar - c * sgn * alpf * betu;
  double rvmsbt = rvpsbt;
  double rvmsbu = rvpsbu;
  double rvmsen = alps * (v2b + db2 - cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i)
                  - alpf * c * (bett * vtbar + betu * vubar) * sgn;
  double rvparo = 0.0;
  double rvpamn = 0.0;
  double rvpamt = -sgn * betu;
  double rvpamu = sgn * bett;
  double rvpabt = rrobai * betu;
  double rvpabu = -rrobai * bett;
  double rvpaen = -(betu * vtbar - bett * vubar) * sgn;
  double rvmaro = 0.0;
  double rvmamn = 0.0;
  double rvmamt = -rvpamt;
  double rvmamu = -rvpamu;
  double rvmabt = rvpabt;
  double rvmabu = rvpabu;
  double rvmaen = -rvpaen;
  double rv00ro = 1.0;
  double rv00mn = vnbar;
  double rv00mt = vtbar;
  double rv00mu = vubar;
  double rv00bt = 0.0;
  double rv00bu = 0.0;
  double rv00en = v2b + db2;
  F[0] = 0.5 * (+frol + fror - evpf * wvpf * rvpfro - evmf * wvmf * rvmfro - evps * wvps * rvpsro
                - evms * wvms * rvmsro - evpa * wvpa * rvparo - evma * wvma * rvmaro - ev00 * wv00 * rv00ro);
  F[1] = 0.5 * (+fmnl + fmnr - evpf * wvpf * rvpfmn - evmf * wvmf * rvmfmn - evps * wvps * rvpsmn
                - evms * wvms * rvmsmn - evpa * wvpa * rvpamn - evma * wvma * rvmamn - ev00 * wv00 * rv00mn);
  F[2] = 0.5 * (fmtl + fmtr - evpf * wvpf * rvpfmt - evmf * wvmf * rvmfmt - evps * wvps * rvpsmt
                - evms * wvms * rvmsmt - evpa * wvpa * rvpamt - evma * wvma * rvmamt - ev00 * wv00 * rv00mt);
  F[3] = 0.5 * (fmul + fmur - evpf * wvpf * rvpfmu - evmf * wvmf * rvmfmu - evps * wvps * rvpsmu
                - evms * wvms * rvmsmu - evpa * wvpa * rvpamu - evma * wvma * rvmamu - ev00 * wv00 * rv00mu);
  F[4] = 0.5 * (fbtl + fbtr - evpf * wvpf * rvpfbt - evmf * wvmf * rvmfbt - evps * wvps * rvpsbt
                - evms * wvms * rvmsbt - evpa * wvpa * rvpabt - evma * wvma * rvmabt - ev00 * wv00 * rv00bt);
  F[5] = 0.5 * (fbul + fbur - evpf * wvpf * rvpfbu - evmf * wvmf * rvmfbu - evps * wvps * rvpsbu
                - evms * wvms * rvmsbu - evpa * wvpa * rvpabu - evma * wvma * rvmabu - ev00 * wv00 * rv00bu);
  F[6] = 0.5 * (fenl + fenr - evpf * wvpf * rvpfen - evmf * wvmf * rvmfen - evps * wvps * rvpsen
                - evms * wvms * rvmsen - evpa * wvpa * rvpaen - evma * wvma * rvmaen - ev00 * wv00 * rv00en);
}

/* HLLD (lowdiss=0) and the quasi-all-speed LHLLD (lowdiss=1): the two
 * solvers share every line except the HLL normal velocity and total
 * pressure, where LHLLD applies the shock sensor theta and the all-speed
 * factor mcc*(2-mcc). */
static void flux_hlld_family(int lowdiss, const double *L, const double *R, double bnc, double gamma,
                             const double *dv, double *F)
{
  const double rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const double ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  double gammam1i = 1.0 / (gamma - 1.0);
  double bnc2 = bnc * bnc;
  double sgn = (bnc > 0) ? (1.0) : (-1.0);
  double roli = 1.0 / rol;
  double pml = 0.5 * (btl * btl + bul * bul);
  double ptl = prl + pml;
  double enl, enr, vl2 = 0, vr2 = 0;
  if (lowdiss) {
    vl2 = vnl * vnl + vtl * vtl + vul * vul;
    enl = gammam1i * prl + pml + 0.5 * rol * vl2;
  } else {
    enl = gammam1i * prl + pml + 0.5 * rol * (vnl * vnl + vtl * vtl + vul * vul);
  }
  double vbl = vtl * btl + vul * bul;
  double rori = 1.0 / ror;
  double pmr = 0.5 * (btr * btr + bur * bur);
  double ptr = prr + pmr;
  if (lowdiss) {
    vr2 = vnr * vnr + vtr * vtr + vur * vur;
    enr = gammam1i * prr + pmr + 0.5 * ror * vr2;
  } else {
    enr = gammam1i * prr + pmr + 0.5 * ror * (vnr * vnr + vtr * vtr + vur * vur);
  }
  double vbr = vtr * btr + vur * bur;
  double cl2 = gamma * prl * roli;
  double cr2 = gamma * prr * rori;
  double cal2 = bnc2 * roli;
  double car2 = bnc2 * rori;
  double cbl2 = cl2 + cal2 + 2.0 * pml * roli;
  double cbr2 = cr2 + car2 + 2.0 * pmr * rori;
  double cfl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * cl2 * cal2)));
  double cfr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * cr2 * car2)));
  double cmax, ccl2 = 0, ccr2 = 0;
  if (lowdiss) {
    cbl2 = vl2 + cal2 + 2.0 * pml * roli;
    cbr2 = vr2 + car2 + 2.0 * pmr * rori;
    ccl2 = 0.5 * (cbl2 + sqrt(fabs(cbl2 * cbl2 - 4.0 * vl2 * cal2)));
    ccr2 = 0.5 * (cbr2 + sqrt(fabs(cbr2 * cbr2 - 4.0 * vr2 * car2)));
    cmax = sqrt(omax(cfl2, cfr2));
  } else {
    double cfl = sqrt(cfl2);
    double cfr = sqrt(cfr2);
    cmax = omax(cfl, cfr);
  }
  double sl = omin(0.0, omin(vnl, vnr) - cmax);
  double sr = omax(0.0, omax(vnl, vnr) + cmax);
  double slvl = sl - vnl;
  double srvr = sr - vnr;
  double rslvl = rol * slvl;
  double rsrvr = ror * srvr;
  double drsvi = 1.0 / (rsrvr - rslvl);
  double vnc, ptc;
  if (lowdiss) {
    double theta = omin(1.0, (cmax - omin(dv[0], 0.0)) / (cmax - omin(dv[1], 0.0)));
    theta *= theta;
    theta *= theta;
    vnc = (rsrvr * vnr - rslvl * vnl - theta * (ptr - ptl)) * drsvi;
    double mcc = omin(1.0, sqrt(omax(ccl2, ccr2)) / cmax);
    ptc = (rsrvr * ptl - rslvl * ptr + (mcc * (2.0 - mcc)) * rsrvr * rslvl * (vnr - vnl)) * drsvi;
  } else {
    vnc = (rsrvr * vnr - rslvl * vnl - ptr + ptl) * drsvi;
    ptc = (rsrvr * ptl - rslvl * ptr + rsrvr * rslvl * (vnr - vnl)) * drsvi;
  }
  double slvc = sl - vnc;
  double srvc = sr - vnc;
  double ro2l = rslvl / slvc;
  double ro2r = rsrvr / srvc;
  double rhdl = rslvl * slvc - bnc2;
  double rhdr = rsrvr * srvc - bnc2;
  double vt2l, vu2l, bt2l, bu2l, vt2r, vu2r, bt2r, bu2r;
  if (fabs(rhdl) > FL_EPS) {
    double rhdli = 1.0 / rhdl;
    double rhnvl = (vnl - vnc) * bnc;
    double rhnbl = rslvl * slvl - bnc2;
    vt2l = vtl + rhnvl * rhdli * btl;
    vu2l = vul + rhnvl * rhdli * bul;
    bt2l = rhnbl * rhdli * btl;
    bu2l = rhnbl * rhdli * bul;
  } else {
    vt2l = vtl;
    vu2l = vul;
    bt2l = btl;
    bu2l = bul;
  }
  if (fabs(rhdr) > FL_EPS) {
    double rhdri = 1.0 / rhdr;
    double rhnvr = (vnr - vnc) * bnc;
    double rhnbr = rsrvr * srvr - bnc2;
    vt2r = vtr + rhnvr * rhdri * btr;
    vu2r = vur + rhnvr * rhdri * bur;
    bt2r = rhnbr * rhdri * btr;
    bu2r = rhnbr * rhdri * bur;
  } else {
    vt2r = vtr;
    vu2r = vur;
    bt2r = btr;
    bu2r = bur;
  }
  double vb2l = vt2l * bt2l + vu2l * bu2l;
  double vb2r = vt2r * bt2r + vu2r * bu2r;
  double en2l = (slvl * enl - ptl * vnl + ptc * vnc + bnc * (vbl - vb2l)) / slvc;
  double en2r = (srvr * enr - ptr * vnr + ptc * vnc + bnc * (vbr - vb2r)) / srvc;
  double rro2l = sqrt(ro2l);
  double rro2r = sqrt(ro2r);
  double rro2i = 1.0 / (rro2r + rro2l);
  double vt3m = (rro2r * vt2r + rro2l * vt2l + (bt2r - bt2l) * sgn) * rro2i;
  double vu3m = (rro2r * vu2r + rro2l * vu2l + (bu2r - bu2l) * sgn) * rro2i;
  double bt3m = (rro2l * bt2r + rro2r * bt2l + rro2r * rro2l * (vt2r - vt2l) * sgn) * rro2i;
  double bu3m = (rro2l * bu2r + rro2r * bu2l + rro2r * rro2l * (vu2r - vu2l) * sgn) * rro2i;
  double vb3m = vt3m * bt3m + vu3m * bu3m;
  double en3l = en2l - rro2l * (vb2l - vb3m) * sgn;
  double en3r = en2r + rro2r * (vb2r - vb3m) * sgn;
  double hl = (vnc > 0) ? (1.0) : (0.0);
  double hr = 1.0 - hl;
  double h2l = (vnc - fabs(bnc) / rro2l > 0) ? (1.0) : (0.0);
  double h3l = (1.0 - h2l) * hl;
  double h2r = (vnc + fabs(bnc) / rro2r > 0) ? (0.0) : (1.0);
  double h3r = (1.0 - h2r) * hr;
  double rou = ro2l * hl + ro2r * hr;
  double vtu = vt2l * h2l + vt3m * (h3l + h3r) + vt2r * h2r;
  double vuu = vu2l * h2l + vu3m * (h3l + h3r) + vu2r * h2r;
  double btu = bt2l * h2l + bt3m * (h3l + h3r) + bt2r * h2r;
  double buu = bu2l * h2l + bu3m * (h3l + h3r) + bu2r * h2r;
  double enu = en2l * h2l + en3l * h3l + en3r * h3r + en2r * h2r;
  F[0] = rou * vnc;
  F[1] = rou * vnc * vnc + ptc - bnc2 * 0.5;
  F[2] = rou * vtu * vnc - bnc * btu;
  F[3] = rou * vuu * vnc - bnc * buu;
  F[4] = btu * vnc - bnc * vtu;
  F[5] = buu * vnc - bnc * vuu;
  F[6] = (enu + ptc) * vnc - bnc * (vtu * btu + vuu * buu);
}

static void riemann(int rmn, const double *L, const double *R, double bn, double gamma, const double *dv, double *F)
{
  switch (rmn) {
  case 0: flux_roe(L, R, bn, gamma, dv, F); break;
  case 1: flux_hlld_family(0, L, R, bn, gamma, dv, F); break;
  case 2: flux_hlld_family(1, L, R, bn, gamma, dv, F); break;
  default: flux_mlau(L, R, bn, gamma, dv, F); break;
  }
}

/* mhd_cuct_weight; xo/yo are the strides of the two in-plane axes. */
static double cuct_weight(const double *ro, const double *va, const double *vb, const double *ba,
                          const double *bb, int xo, int yo)
{
  const double eps = 1e-12;
  int s0 = -xo - yo, s1 = -yo, s2 = -xo, s3 = 0;
  double roc = 0.25 * (ro[s0] + ro[s1] + ro[s2] + ro[s3]);
  roc = 1.0 / sqrt(roc);
  double vac = 0.25 * (va[s0] + va[s1] + va[s2] + va[s3]);
  double vbc = 0.25 * (vb[s0] + vb[s1] + vb[s2] + vb[s3]);
  double bac = 0.5 * (ba[s1] + ba[s3]);
  double bbc = 0.5 * (bb[s2] + bb[s3]);
  vac = fabs(vac) + fabs(bac * roc);
  vbc = fabs(vbc) + fabs(bbc * roc);
  return (vac + 0.5 * eps) / (vac + vbc + eps);
}

/* mhd_lr_fb: face states of the B-flux  b_t*v_n - b_n*v_t  (linear interp) */
static void lr_fb(const double *vn, const double *vt, const double *bn, const double *bt, int o,
                  lr_fn lin, double *vl, double *vr)
{
  double p[5], q[5], fl0, fr0, fl1, fr1;
  int m;
  for (m = 0; m < 5; m++) {
    p[m] = bt[(m - 2) * o] * vn[(m - 2) * o];
    q[m] = vt[(m - 2) * o];
  }
  lin(&p[2], &fl0, &fr0);
  lin(&q[2], &fl1, &fr1);
  *vl = fl0 - bn[o] * fl1;
  *vr = fr0 - bn[0] * fr1;
}

static void lr_single(const double *f, int o, lr_fn lin, double *vl, double *vr)
{
  double p[5];
  int m;
  for (m = 0; m < 5; m++) p[m] = f[(m - 2) * o];
  lin(&p[2], vl, vr);
}

static inline void rk_upd(double *f1, double f0, double df, double a, double b)
{
  (*f1) = a * f0 + b * ((*f1) + df);
}

/* ---------------------------------------------------------------------- */
/* Boundary conditions (bc1d / bc2d / bc3d).  dn: 0 periodic, +-1 Neumann /
 * Dirichlet mirror, +-2 open / zero-fix; st: staggered (face) variable.     */
/* fill ghosts of one line of n points at stride s (pointer at index 0) */
static void bc_line(double *f, int n, long s, int off, int st, int dn)
{
  int i;
  if (dn == 0) {
    for (i = 0; i < off; i++) {
      f[s * (n - 1 - i)] = f[s * (2 * off - 1 - i)];
      f[s * i] = f[s * (n - 2 * off + i)];
    }
  } else if (abs(dn) == 1) {
    for (i = 0; i < off; i++) f[s * i] = dn * f[s * ((2 * off - 1 + st) - i)];
    for (i = 0; i < off - st; i++) f[s * (n - 1 - i)] = dn * f[s * ((n - 2 * off + st) + i)];
  } else if (abs(dn) == 2) {
    for (i = 0; i < off; i++) f[s * i] = 0.25 * (2 + dn) * f[s * off];
    for (i = 0; i < off - st; i++) f[s * (n - 1 - i)] = 0.25 * (2 + dn) * f[s * (n - 1 - off + st)];
  }
}

/* one array; x pass over all (j,k), then y pass over all (i,k), then z */
static void bc_array(const qor_cfg *c, double *f, const int st[3], const int dn[3])
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const long sy = nx, sz = (long)nx * ny;
  int i, j, k;
  if (c->ndim >= 1)
    for (k = 0; k < nz; k++)
      for (j = 0; j < ny; j++) bc_line(f + sz * k + sy * j, nx, 1, c->xoff, st[0], dn[0]);
  if (c->ndim >= 2)
    for (k = 0; k < nz; k++)
      for (i = 0; i < nx; i++) bc_line(f + sz * k + i, ny, sy, c->yoff, st[1], dn[1]);
  if (c->ndim >= 3)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) bc_line(f + sy * j + i, nz, sz, c->zoff, st[2], dn[2]);
}

void qor_bound(const qor_cfg *c, double *const val[8])
{
  int m;
  for (m = 0; m < 8; m++) {
    int st[3] = {c->st[0][m], c->st[1][m], c->st[2][m]};
    int dn[3] = {c->dn[0][m], c->dn[1][m], c->dn[2][m]};
    bc_array(c, val[m], st, dn);
  }
  if (c->phi && c->ndim == 2) {
    /* with gravity along y the energy's y ghosts hold the hydrostatic
     * extrapolation (2D/SERIAL/RTI/mhd2d_class.cpp:10-32) */
    const int nx = c->nx, ny = c->ny, yoff = c->yoff;
    const double fac = (2.0 - c->gam) / (c->gam - 1.0);
    const double *phi = c->phi, *ro = val[0];
    double *en = val[7];
    int i, j;
    for (j = 0; j < yoff; j++)
      for (i = 0; i < nx; i++) {
        long ss = (long)nx * j + i, sb = (long)nx * yoff + i;
        en[ss] = en[sb] - fac * ro[sb] * (phi[ss] - phi[sb]);
      }
    for (j = ny - yoff; j < ny; j++)
      for (i = 0; i < nx; i++) {
        long ss = (long)nx * j + i, sb = (long)nx * (ny - yoff - 1) + i;
        en[ss] = en[sb] - fac * ro[sb] * (phi[ss] - phi[sb]);
      }
  }
}

/* ---------------------------------------------------------------------- */
static const double *rk_coef(int r_k, int rk)
{
  static double tab[3][3][2];
  static int init = 0;
  if (!init) {
    int R;
    for (R = 1; R <= 3; R++) {
      tab[R - 1][0][0] = 0.0;
      tab[R - 1][0][1] = 1.0;
      tab[R - 1][1][0] = 0.5 + (R - 2) * 0.25;
      tab[R - 1][1][1] = 0.5 - (R - 2) * 0.25;
      tab[R - 1][2][0] = 1. / 3.;
      tab[R - 1][2][1] = 2. / 3.;
    }
    init = 1;
  }
  return tab[r_k - 1][rk];
}

/* primitive variables at every cell: v = m/ro; p = (gam-1)(E - (ro v^2 + B^2)/2) */
static void prim_all(long nd, const double *ro, const double *mx, const double *my, const double *mz,
                     const double *en, const double *cx, const double *cy, const double *cz, double gam,
                     double *vx, double *vy, double *vz, double *pr)
{
  long ss;
#pragma omp parallel for schedule(static)
  for (ss = 0; ss < nd; ss++) {
    double iro = 1.0 / ro[ss];
    vx[ss] = mx[ss] * iro;
    vy[ss] = my[ss] * iro;
    vz[ss] = mz[ss] * iro;
    double v2 = vx[ss] * vx[ss] + vy[ss] * vy[ss] + vz[ss] * vz[ss];
    double b2 = cx[ss] * cx[ss] + cy[ss] * cy[ss] + cz[ss] * cz[ss];
    pr[ss] = (gam - 1) * (en[ss] - 0.5 * (ro[ss] * v2 + b2));
  }
}

/* gather the lrstate stencil of cell ss along stride o */
static void gather7(const double *a0, const double *a1, const double *a2, const double *a3, const double *a4,
                    const double *a5, const double *a6, long ss, long o, double s[7][5])
{
  int m;
  for (m = 0; m < 5; m++) {
    long q = ss + (m - 2) * o;
    s[0][m] = a0[q];
    s[1][m] = a1[q];
    s[2][m] = a2[q];
    s[3][m] = a3[q];
    s[4][m] = a4[q];
    s[5][m] = a5[q];
    s[6][m] = a6[q];
  }
}

/* ---------------------------------------------------------------------- */
static void step1d(const qor_cfg *c, double *const val[8], double dt)
{
  const int nx = c->nx, nm = 8;
  const double dtdx = dt / c->dx;
  lr_fn flr = interp_tab[c->odr - 1];
  c_fn fdf = df_tab[c->odr - 1];
  double *ro = val[0], *mx = val[1], *my = val[2], *mz = val[3], *bx = val[4], *by = val[5], *bz = val[6],
         *en = val[7];
  double *ut = malloc(sizeof(double) * nm * nx), *ul = malloc(sizeof(double) * nm * nx),
         *ur = malloc(sizeof(double) * nm * nx), *fx = malloc(sizeof(double) * nm * nx),
         *vx = malloc(sizeof(double) * nx), *vy = malloc(sizeof(double) * nx), *vz = malloc(sizeof(double) * nx),
         *pr = malloc(sizeof(double) * nx);
  const double *cx = bx, *cy = by, *cz = bz;
  int i, m, rk;
  for (m = 0; m < nm; m++) memcpy(ut + (long)m * nx, val[m], sizeof(double) * nx);
  for (rk = 0; rk < c->r_k; rk++) {
    const double *rf = rk_coef(c->r_k, rk);
    prim_all(nx, ro, mx, my, mz, en, cx, cy, cz, c->gam, vx, vy, vz, pr);
    for (i = 2; i < nx - 2; i++) {
      double s[7][5], vl[7], vr[7];
      gather7(ro, vx, vy, vz, cy, cz, pr, i, 1, s);
      lrstate(s, cx[i], c->gam, flr, vl, vr);
      int sl = i + 1, sr = i;
      ul[nm * sl + 0] = vl[0]; ul[nm * sl + 1] = vl[1]; ul[nm * sl + 2] = vl[2]; ul[nm * sl + 3] = vl[3];
      ul[nm * sl + 4] = bx[sl]; ul[nm * sl + 5] = vl[4]; ul[nm * sl + 6] = vl[5]; ul[nm * sl + 7] = vl[6];
      ur[nm * sr + 0] = vr[0]; ur[nm * sr + 1] = vr[1]; ur[nm * sr + 2] = vr[2]; ur[nm * sr + 3] = vr[3];
      ur[nm * sr + 4] = bx[sr]; ur[nm * sr + 5] = vr[4]; ur[nm * sr + 6] = vr[5]; ur[nm * sr + 7] = vr[6];
    }
    for (i = 3; i < nx - 2; i++) {
      double L[7] = {ul[nm * i + 0], ul[nm * i + 1], ul[nm * i + 2], ul[nm * i + 3], ul[nm * i + 5], ul[nm * i + 6],
                     ul[nm * i + 7]};
      double R[7] = {ur[nm * i + 0], ur[nm * i + 1], ur[nm * i + 2], ur[nm * i + 3], ur[nm * i + 5], ur[nm * i + 6],
                     ur[nm * i + 7]};
      double bn = 0.5 * (ul[nm * i + 4] + ur[nm * i + 4]);
      double dv[2] = {(vx[i] - vx[i - 1]), 0.0};
      double F[7];
      riemann(c->rmn, L, R, bn, c->gam, dv, F);
      fx[nm * i + 0] = F[0]; fx[nm * i + 1] = F[1]; fx[nm * i + 2] = F[2]; fx[nm * i + 3] = F[3];
      fx[nm * i + 4] = 0; fx[nm * i + 5] = F[4]; fx[nm * i + 6] = F[5]; fx[nm * i + 7] = F[6];
    }
    for (i = c->xoff; i < nx - c->xoff; i++) {
      static const int upd[7] = {0, 1, 2, 3, 5, 6, 7};
      int q;
      for (q = 0; q < 7; q++) {
        m = upd[q];
        double f4[4] = {fx[nm * (i - 1) + m], fx[nm * i + m], fx[nm * (i + 1) + m], fx[nm * (i + 2) + m]};
        double du = fdf(&f4[1]);
        rk_upd(&val[m][i], ut[(long)m * nx + i], -dtdx * du, rf[0], rf[1]);
      }
    }
    qor_bound(c, val);
  }
  free(ut); free(ul); free(ur); free(fx); free(vx); free(vy); free(vz); free(pr);
}

/* ---------------------------------------------------------------------- */
static void step2d(const qor_cfg *c, double *const val[8], double dt)
{
  const int nx = c->nx, ny = c->ny, nm = 8;
  const long nxy = (long)nx * ny;
  const double dtdx = dt / c->dx, dtdy = dt / c->dy;
  const int riem = c->rmn;
  lr_fn lin = linear_tab[c->odr - 1], flr = interp_tab[c->odr - 1];
  c_fn fbc = fcen_tab[c->odr - 1], fdf = df_tab[c->odr - 1];
  double *ro = val[0], *mx = val[1], *my = val[2], *mz = val[3], *bx = val[4], *by = val[5], *bz = val[6],
         *en = val[7];
  double *ut = malloc(sizeof(double) * nm * nxy);
  double *ul = malloc(sizeof(double) * nm * nxy), *ur = malloc(sizeof(double) * nm * nxy);
  double *fx = malloc(sizeof(double) * nm * nxy), *fy = malloc(sizeof(double) * nm * nxy);
  double *ql = malloc(sizeof(double) * nxy), *qr = malloc(sizeof(double) * nxy), *fc = malloc(sizeof(double) * nxy);
  double *ez = malloc(sizeof(double) * nxy), *ct = malloc(sizeof(double) * nxy);
  double *dvx = malloc(sizeof(double) * nxy), *dvy = malloc(sizeof(double) * nxy);
  double *el = malloc(sizeof(double) * 2 * nxy), *er = malloc(sizeof(double) * 2 * nxy);
  double *vx = malloc(sizeof(double) * nxy), *vy = malloc(sizeof(double) * nxy), *vz = malloc(sizeof(double) * nxy),
         *pr = malloc(sizeof(double) * nxy), *cx = malloc(sizeof(double) * nxy), *cy = malloc(sizeof(double) * nxy);
  double *cz = bz;
  int m, rk;
  for (m = 0; m < nm; m++) memcpy(ut + m * nxy, val[m], sizeof(double) * nxy);
  for (rk = 0; rk < c->r_k; rk++) {
    const double *rf = rk_coef(c->r_k, rk);
    int i, j;
    /* cell-centre B */
#pragma omp parallel for private(i)
    for (j = 0; j < ny; j++)
      for (i = 1; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        double b4[4] = {bx[ss - 1], bx[ss], bx[ss + 1], bx[ss + 2]};
        cx[ss] = fbc(&b4[1]);
      }
#pragma omp parallel for private(i)
    for (j = 1; j < ny - 2; j++)
      for (i = 0; i < nx; i++) {
        long ss = (long)nx * j + i;
        double b4[4] = {by[ss - nx], by[ss], by[ss + nx], by[ss + 2 * nx]};
        cy[ss] = fbc(&b4[1]);
      }
    {
      int s0[3] = {0, 0, 0};
      int dcx[3] = {c->dn[0][4], c->dn[1][4], 0}, dcy[3] = {c->dn[0][5], c->dn[1][5], 0};
      bc_array(c, cx, s0, dcx);
      bc_array(c, cy, s0, dcy);
    }
    /* gravity: the potential energy is taken out of en for the pressure and
     * put back, in place (2D/SERIAL/RTI/mhd2d_solve.cpp:86-88) */
    if (c->phi)
      for (long ss = 0; ss < nxy; ss++) en[ss] -= ro[ss] * c->phi[ss];
    prim_all(nxy, ro, mx, my, mz, en, cx, cy, cz, c->gam, vx, vy, vz, pr);
    if (c->phi)
      for (long ss = 0; ss < nxy; ss++) en[ss] += ro[ss] * c->phi[ss];
    for (long ss = 0; ss < nxy; ss++) {
      ez[ss] = 0.0;
      ct[ss] = 0.5;
    }
    if (c->ctw) {
#pragma omp parallel for private(i)
      for (j = 1; j < ny; j++)
        for (i = 1; i < nx; i++) {
          long ss = (long)nx * j + i;
          ct[ss] = cuct_weight(&ro[ss], &vx[ss], &vy[ss], &bx[ss], &by[ss], 1, nx);
        }
    }
#pragma omp parallel for private(i)
    for (j = 1; j < ny - 1; j++)
      for (i = 1; i < nx - 1; i++) {
        long ss = (long)nx * j + i;
        dvx[ss] = omin((vx[ss + 1] - vx[ss]), (vx[ss] - vx[ss - 1]));
        dvy[ss] = omin((vy[ss + nx] - vy[ss]), (vy[ss] - vy[ss - nx]));
      }
    {
      int s0[3] = {0, 0, 0};
      int dnvx[3] = {-c->dn[0][1], c->dn[1][1], 0}, dnvy[3] = {c->dn[0][2], -c->dn[1][2], 0};
      bc_array(c, dvx, s0, dnvx);
      bc_array(c, dvy, s0, dnvy);
    }
    /* x faces */
#pragma omp parallel for private(i)
    for (j = 0; j < ny; j++) {
      for (i = 2; i < nx - 2; i++) {
        long ss = (long)nx * j + i, sl = ss + 1, sr = ss;
        double s[7][5], vl[7], vr[7];
        gather7(ro, vx, vy, vz, cy, cz, pr, ss, 1, s);
        lrstate(s, cx[ss], c->gam, flr, vl, vr);
        ul[nm * sl + 0] = vl[0]; ul[nm * sl + 1] = vl[1]; ul[nm * sl + 2] = vl[2]; ul[nm * sl + 3] = vl[3];
        ul[nm * sl + 4] = bx[sl]; ul[nm * sl + 5] = vl[4]; ul[nm * sl + 6] = vl[5]; ul[nm * sl + 7] = vl[6];
        ur[nm * sr + 0] = vr[0]; ur[nm * sr + 1] = vr[1]; ur[nm * sr + 2] = vr[2]; ur[nm * sr + 3] = vr[3];
        ur[nm * sr + 4] = bx[sr]; ur[nm * sr + 5] = vr[4]; ur[nm * sr + 6] = vr[5]; ur[nm * sr + 7] = vr[6];
      }
      for (i = 2; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        double a, b;
        lr_fb(&vx[ss], &vy[ss], &bx[ss], &cy[ss], 1, lin, &a, &b);
        ql[ss + 1] = a;
        qr[ss] = b;
      }
    }
#pragma omp parallel for private(i)
    for (j = 0; j < ny; j++)
      for (i = 3; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        double L[7] = {ul[nm * ss + 0], ul[nm * ss + 1], ul[nm * ss + 2], ul[nm * ss + 3], ul[nm * ss + 5],
                       ul[nm * ss + 6], ul[nm * ss + 7]};
        double R[7] = {ur[nm * ss + 0], ur[nm * ss + 1], ur[nm * ss + 2], ur[nm * ss + 3], ur[nm * ss + 5],
                       ur[nm * ss + 6], ur[nm * ss + 7]};
        double bn = 0.5 * (ul[nm * ss + 4] + ur[nm * ss + 4]);
        double dv[2] = {(vx[ss] - vx[ss - 1]), omin(dvy[ss - 1], dvy[ss])};
        double F[7];
        riemann(riem, L, R, bn, c->gam, dv, F);
        fx[nm * ss + 0] = F[0]; fx[nm * ss + 1] = F[1]; fx[nm * ss + 2] = F[2]; fx[nm * ss + 3] = F[3];
        fx[nm * ss + 4] = 0; fx[nm * ss + 5] = F[4]; fx[nm * ss + 6] = F[5]; fx[nm * ss + 7] = F[6];
        if (c->phi) fx[nm * ss + 7] += F[0] * 0.5 * (c->phi[ss - 1] + c->phi[ss]); /* work done by gravity */
        fc[ss] = 0.5 * (ql[ss] + qr[ss]);
        fx[nm * ss + 5] -= fc[ss];
      }
    /* Ez, part from the x fluxes (interpolated along y) */
#pragma omp parallel for private(i)
    for (j = 2; j < ny - 2; j++)
      for (i = 3; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        double a, b;
        lr_single(&fx[nm * ss + 5], nm * nx, lin, &a, &b);
        el[2 * (ss + nx) + 0] = a;
        er[2 * ss + 0] = b;
        lr_single(&fc[ss], nx, lin, &a, &b);
        el[2 * (ss + nx) + 1] = a;
        er[2 * ss + 1] = b;
      }
#pragma omp parallel for private(i)
    for (j = 3; j < ny - 2; j++)
      for (i = 3; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        ez[ss] += -0.5 * ((el[2 * ss + 0] + er[2 * ss + 0]) + (1.0 - ct[ss]) * (el[2 * ss + 1] + er[2 * ss + 1]));
      }
    /* y faces: cyclic (vy, vz, vx), (bz, bx) */
#pragma omp parallel for private(i)
    for (j = 2; j < ny - 2; j++) {
      for (i = 0; i < nx; i++) {
        long ss = (long)nx * j + i, sl = ss + nx, sr = ss;
        double s[7][5], vl[7], vr[7];
        gather7(ro, vy, vz, vx, cz, cx, pr, ss, nx, s);
        lrstate(s, cy[ss], c->gam, flr, vl, vr);
        ul[nm * sl + 0] = vl[0]; ul[nm * sl + 1] = vl[1]; ul[nm * sl + 2] = vl[2]; ul[nm * sl + 3] = vl[3];
        ul[nm * sl + 4] = by[sl]; ul[nm * sl + 5] = vl[4]; ul[nm * sl + 6] = vl[5]; ul[nm * sl + 7] = vl[6];
        ur[nm * sr + 0] = vr[0]; ur[nm * sr + 1] = vr[1]; ur[nm * sr + 2] = vr[2]; ur[nm * sr + 3] = vr[3];
        ur[nm * sr + 4] = by[sr]; ur[nm * sr + 5] = vr[4]; ur[nm * sr + 6] = vr[5]; ur[nm * sr + 7] = vr[6];
      }
      for (i = 0; i < nx; i++) {
        long ss = (long)nx * j + i;
        double a, b;
        lr_fb(&vy[ss], &vx[ss], &by[ss], &cx[ss], nx, lin, &a, &b);
        ql[ss + nx] = a;
        qr[ss] = b;
      }
    }
#pragma omp parallel for private(i)
    for (j = 3; j < ny - 2; j++)
      for (i = 0; i < nx; i++) {
        long ss = (long)nx * j + i;
        double L[7] = {ul[nm * ss + 0], ul[nm * ss + 1], ul[nm * ss + 2], ul[nm * ss + 3], ul[nm * ss + 5],
                       ul[nm * ss + 6], ul[nm * ss + 7]};
        double R[7] = {ur[nm * ss + 0], ur[nm * ss + 1], ur[nm * ss + 2], ur[nm * ss + 3], ur[nm * ss + 5],
                       ur[nm * ss + 6], ur[nm * ss + 7]};
        double bn = 0.5 * (ul[nm * ss + 4] + ur[nm * ss + 4]);
        double dv[2] = {(vy[ss] - vy[ss - nx]), omin(dvx[ss - nx], dvx[ss])};
        double F[7];
        riemann(riem, L, R, bn, c->gam, dv, F);
        fy[nm * ss + 0] = F[0]; fy[nm * ss + 2] = F[1]; fy[nm * ss + 3] = F[2]; fy[nm * ss + 1] = F[3];
        fy[nm * ss + 5] = 0; fy[nm * ss + 6] = F[4]; fy[nm * ss + 4] = F[5]; fy[nm * ss + 7] = F[6];
        if (c->phi) fy[nm * ss + 7] += F[0] * 0.5 * (c->phi[ss - nx] + c->phi[ss]);
        fc[ss] = 0.5 * (ql[ss] + qr[ss]);
        fy[nm * ss + 4] -= fc[ss];
      }
    /* Ez, part from the y fluxes (interpolated along x) */
#pragma omp parallel for private(i)
    for (j = 3; j < ny - 2; j++)
      for (i = 2; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        double a, b;
        lr_single(&fy[nm * ss + 4], nm, lin, &a, &b);
        el[2 * (ss + 1) + 0] = a;
        er[2 * ss + 0] = b;
        lr_single(&fc[ss], 1, lin, &a, &b);
        el[2 * (ss + 1) + 1] = a;
        er[2 * ss + 1] = b;
      }
#pragma omp parallel for private(i)
    for (j = 3; j < ny - 2; j++)
      for (i = 3; i < nx - 2; i++) {
        long ss = (long)nx * j + i;
        ez[ss] += +0.5 * ((el[2 * ss + 0] + er[2 * ss + 0]) + ct[ss] * (el[2 * ss + 1] + er[2 * ss + 1]));
      }
    /* cell-centred update: ro, mx, my, mz, bz, en */
#pragma omp parallel for private(i)
    for (j = c->yoff; j < ny - c->yoff; j++)
      for (i = c->xoff; i < nx - c->xoff; i++) {
        static const int upd[6] = {0, 1, 2, 3, 6, 7};
        long ss = (long)nx * j + i;
        int q;
        for (q = 0; q < 6; q++) {
          int v = upd[q];
          double fxs[4] = {fx[nm * (ss - 1) + v], fx[nm * ss + v], fx[nm * (ss + 1) + v], fx[nm * (ss + 2) + v]};
          double fys[4] = {fy[nm * (ss - nx) + v], fy[nm * ss + v], fy[nm * (ss + nx) + v], fy[nm * (ss + 2 * nx) + v]};
          double du0 = fdf(&fxs[1]);
          double du1 = fdf(&fys[1]);
          rk_upd(&val[v][ss], ut[v * nxy + ss], (-dtdx * du0 - dtdy * du1), rf[0], rf[1]);
        }
        if (c->phi) { /* gravity on the momentum, with the updated density (RTI/mhd2d_solve.cpp:361-362) */
          const double *phi = c->phi;
          mx[ss] += -rf[1] * ro[ss] * 0.5 * (phi[ss + 1] - phi[ss - 1]) * dtdx;
          my[ss] += -rf[1] * ro[ss] * 0.5 * (phi[ss + nx] - phi[ss - nx]) * dtdy;
        }
      }
    /* CT: bx (one extra face in x), by (one extra face in y) */
#pragma omp parallel for private(i)
    for (j = c->yoff; j < ny - c->yoff; j++)
      for (i = c->xoff; i < nx - c->xoff + 1; i++) {
        long ss = (long)nx * j + i;
        double e4[4] = {ez[ss - nx], ez[ss], ez[ss + nx], ez[ss + 2 * nx]};
        double du = fdf(&e4[1]);
        double dty = +dtdy;
        rk_upd(&bx[ss], ut[4 * nxy + ss], -dty * du, rf[0], rf[1]);
      }
#pragma omp parallel for private(i)
    for (j = c->yoff; j < ny - c->yoff + 1; j++)
      for (i = c->xoff; i < nx - c->xoff; i++) {
        long ss = (long)nx * j + i;
        double e4[4] = {ez[ss - 1], ez[ss], ez[ss + 1], ez[ss + 2]};
        double du = fdf(&e4[1]);
        double dty = -dtdx;
        rk_upd(&by[ss], ut[5 * nxy + ss], -dty * du, rf[0], rf[1]);
      }
    qor_bound(c, val);
  }
  free(ut); free(ul); free(ur); free(fx); free(fy); free(ql); free(qr); free(fc); free(ez); free(ct);
  free(dvx); free(dvy); free(el); free(er); free(vx); free(vy); free(vz); free(pr); free(cx); free(cy);
}

/* ---------------------------------------------------------------------- */
/* 3D.  Strides: x 1, y nx, z nxy.  Sweep d uses the cyclic permutation
 * (n, t, u) = (x,y,z), (y,z,x), (z,x,y).                                    */
static void step3d(const qor_cfg *c, double *const val[8], double dt)
{
  const int nx = c->nx, ny = c->ny, nz = c->nz, nm = 8;
  const long nxy = (long)nx * ny, nd = nxy * nz;
  const double dtdx = dt / c->dx, dtdy = dt / c->dy, dtdz = dt / c->dz;
  const int riem = c->rmn;
  lr_fn lin = linear_tab[c->odr - 1], flr = interp_tab[c->odr - 1];
  c_fn fbc = fcen_tab[c->odr - 1], fdf = df_tab[c->odr - 1];
  double *ro = val[0], *mx = val[1], *my = val[2], *mz = val[3], *bx = val[4], *by = val[5], *bz = val[6],
         *en = val[7];
  double *ut = malloc(sizeof(double) * nm * nd);
  double *fx = malloc(sizeof(double) * nm * nd), *fy = malloc(sizeof(double) * nm * nd),
         *fz = malloc(sizeof(double) * nm * nd);
  double *fc = malloc(sizeof(double) * 2 * nd);
  double *ex = malloc(sizeof(double) * nd), *ey = malloc(sizeof(double) * nd), *ez = malloc(sizeof(double) * nd);
  double *ct = malloc(sizeof(double) * 3 * nd);
  double *dvx = malloc(sizeof(double) * nd), *dvy = malloc(sizeof(double) * nd), *dvz = malloc(sizeof(double) * nd);
  double *vx = malloc(sizeof(double) * nd), *vy = malloc(sizeof(double) * nd), *vz = malloc(sizeof(double) * nd),
         *pr = malloc(sizeof(double) * nd), *cx = malloc(sizeof(double) * nd), *cy = malloc(sizeof(double) * nd),
         *cz = malloc(sizeof(double) * nd);
  int m, rk;
  for (m = 0; m < nm; m++) memcpy(ut + m * nd, val[m], sizeof(double) * nd);
  for (rk = 0; rk < c->r_k; rk++) {
    const double *rf = rk_coef(c->r_k, rk);
    int i, j, k;
#pragma omp parallel for private(i, j)
    for (k = 0; k < nz; k++) {
      for (j = 0; j < ny; j++)
        for (i = 1; i < nx - 2; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          double b4[4] = {bx[ss - 1], bx[ss], bx[ss + 1], bx[ss + 2]};
          cx[ss] = fbc(&b4[1]);
        }
      for (j = 1; j < ny - 2; j++)
        for (i = 0; i < nx; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          double b4[4] = {by[ss - nx], by[ss], by[ss + nx], by[ss + 2 * nx]};
          cy[ss] = fbc(&b4[1]);
        }
    }
#pragma omp parallel for private(i, j)
    for (k = 1; k < nz - 2; k++)
      for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          double b4[4] = {bz[ss - nxy], bz[ss], bz[ss + nxy], bz[ss + 2 * nxy]};
          cz[ss] = fbc(&b4[1]);
        }
    {
      int s0[3] = {0, 0, 0};
      int d1[3] = {c->dn[0][4], c->dn[1][4], c->dn[2][4]}, d2[3] = {c->dn[0][5], c->dn[1][5], c->dn[2][5]},
          d3[3] = {c->dn[0][6], c->dn[1][6], c->dn[2][6]};
      bc_array(c, cx, s0, d1);
      bc_array(c, cy, s0, d2);
      bc_array(c, cz, s0, d3);
    }
    prim_all(nd, ro, mx, my, mz, en, cx, cy, cz, c->gam, vx, vy, vz, pr);
#pragma omp parallel for
    for (long ss = 0; ss < nd; ss++) {
      ex[ss] = ey[ss] = ez[ss] = 0.0;
      ct[ss] = ct[nd + ss] = ct[2 * nd + ss] = 0.5;
    }
    if (c->ctw) {
#pragma omp parallel for private(i, j)
      for (k = 0; k < nz; k++)
        for (j = 1; j < ny; j++)
          for (i = 1; i < nx; i++) {
            long ss = nx * (ny * (long)k + j) + i;
            ct[ss] = cuct_weight(&ro[ss], &vx[ss], &vy[ss], &bx[ss], &by[ss], 1, nx);
          }
#pragma omp parallel for private(i, j)
      for (k = 1; k < nz; k++) {
        for (j = 1; j < ny; j++)
          for (i = 0; i < nx; i++) {
            long ss = nx * (ny * (long)k + j) + i;
            ct[nd + ss] = cuct_weight(&ro[ss], &vy[ss], &vz[ss], &by[ss], &bz[ss], nx, nxy);
          }
        for (j = 0; j < ny; j++)
          for (i = 1; i < nx; i++) {
            long ss = nx * (ny * (long)k + j) + i;
            ct[2 * nd + ss] = cuct_weight(&ro[ss], &vz[ss], &vx[ss], &bz[ss], &bx[ss], nxy, 1);
          }
      }
    }
#pragma omp parallel for private(i, j)
    for (k = 1; k < nz - 1; k++)
      for (j = 1; j < ny - 1; j++)
        for (i = 1; i < nx - 1; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          dvx[ss] = omin((vx[ss + 1] - vx[ss]), (vx[ss] - vx[ss - 1]));
          dvy[ss] = omin((vy[ss + nx] - vy[ss]), (vy[ss] - vy[ss - nx]));
          dvz[ss] = omin((vz[ss + nxy] - vz[ss]), (vz[ss] - vz[ss - nxy]));
        }
    {
      int s0[3] = {0, 0, 0};
      int d1[3] = {-c->dn[0][1], c->dn[1][1], c->dn[2][1]};
      int d2[3] = {c->dn[0][2], -c->dn[1][2], c->dn[2][2]};
      int d3[3] = {c->dn[0][3], c->dn[1][3], -c->dn[2][3]};
      bc_array(c, dvx, s0, d1);
      bc_array(c, dvy, s0, d2);
      bc_array(c, dvz, s0, d3);
    }

    /* the three sweeps share one routine; per-sweep tables below */
    for (int d = 0; d < 3; d++) {
      /* n = normal axis, t = d+1, u = d+2 */
      const long strd[3] = {1, nx, nxy};
      const int dims[3] = {nx, ny, nz};
      const double *vn = (d == 0) ? vx : (d == 1) ? vy : vz;
      const double *vt = (d == 0) ? vy : (d == 1) ? vz : vx;
      const double *vu = (d == 0) ? vz : (d == 1) ? vx : vy;
      const double *cn = (d == 0) ? cx : (d == 1) ? cy : cz;
      const double *ctn = (d == 0) ? cy : (d == 1) ? cz : cx;
      const double *cun = (d == 0) ? cz : (d == 1) ? cx : cy;
      const double *bnf = (d == 0) ? bx : (d == 1) ? by : bz;
      const double *dvt = (d == 0) ? dvy : (d == 1) ? dvz : dvx;
      const double *dvu = (d == 0) ? dvz : (d == 1) ? dvx : dvy;
      double *ff = (d == 0) ? fx : (d == 1) ? fy : fz;
      /* conserved index of (ro, mn, mt, mu, (bn), bt, bu, en) */
      const int mi[3][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {0, 2, 3, 1, 5, 6, 4, 7}, {0, 3, 1, 2, 6, 4, 5, 7}};
      const long o = strd[d];
      const int nn = dims[d];
      /* the two transverse axes */
      const int ta = (d + 1) % 3, tb = (d + 2) % 3;
      const int na = dims[ta], nb = dims[tb];
      const long sa = strd[ta], sb = strd[tb];
      int b;
#pragma omp parallel for private(b)
      for (int a = 0; a < na; a++)
        for (b = 0; b < nb; b++) {
          long base = sa * a + sb * b;
          double *ul = malloc(sizeof(double) * nm * nn), *ur = malloc(sizeof(double) * nm * nn);
          double *ql = malloc(sizeof(double) * 2 * nn), *qr = malloc(sizeof(double) * 2 * nn);
          int q;
          for (q = 2; q < nn - 2; q++) {
            long ss = base + o * q;
            double s[7][5], vl[7], vr[7];
            gather7(ro, vn, vt, vu, ctn, cun, pr, ss, o, s);
            lrstate(s, cn[ss], c->gam, flr, vl, vr);
            int l1 = q + 1, r1 = q;
            ul[nm * l1 + 0] = vl[0]; ul[nm * l1 + 1] = vl[1]; ul[nm * l1 + 2] = vl[2]; ul[nm * l1 + 3] = vl[3];
            ul[nm * l1 + 4] = bnf[ss + o]; ul[nm * l1 + 5] = vl[4]; ul[nm * l1 + 6] = vl[5]; ul[nm * l1 + 7] = vl[6];
            ur[nm * r1 + 0] = vr[0]; ur[nm * r1 + 1] = vr[1]; ur[nm * r1 + 2] = vr[2]; ur[nm * r1 + 3] = vr[3];
            ur[nm * r1 + 4] = bnf[ss]; ur[nm * r1 + 5] = vr[4]; ur[nm * r1 + 6] = vr[5]; ur[nm * r1 + 7] = vr[6];
          }
          for (q = 2; q < nn - 2; q++) {
            long ss = base + o * q;
            double a0, b0;
            lr_fb(&vn[ss], &vt[ss], &bnf[ss], &ctn[ss], o, lin, &a0, &b0);
            ql[2 * (q + 1) + 0] = a0;
            qr[2 * q + 0] = b0;
            lr_fb(&vn[ss], &vu[ss], &bnf[ss], &cun[ss], o, lin, &a0, &b0);
            ql[2 * (q + 1) + 1] = a0;
            qr[2 * q + 1] = b0;
          }
          for (q = 3; q < nn - 2; q++) {
            long ss = base + o * q;
            double L[7] = {ul[nm * q + 0], ul[nm * q + 1], ul[nm * q + 2], ul[nm * q + 3], ul[nm * q + 5],
                           ul[nm * q + 6], ul[nm * q + 7]};
            double R[7] = {ur[nm * q + 0], ur[nm * q + 1], ur[nm * q + 2], ur[nm * q + 3], ur[nm * q + 5],
                           ur[nm * q + 6], ur[nm * q + 7]};
            double bn = 0.5 * (ul[nm * q + 4] + ur[nm * q + 4]);
            double dvtm = omin(omin(dvt[ss - o], dvt[ss]), omin(dvu[ss - o], dvu[ss]));
            double dv[2] = {(vn[ss] - vn[ss - o]), dvtm};
            double F[7];
            riemann(riem, L, R, bn, c->gam, dv, F);
            ff[nm * ss + mi[d][0]] = F[0];
            ff[nm * ss + mi[d][1]] = F[1];
            ff[nm * ss + mi[d][2]] = F[2];
            ff[nm * ss + mi[d][3]] = F[3];
            ff[nm * ss + mi[d][4]] = 0;
            ff[nm * ss + mi[d][5]] = F[4];
            ff[nm * ss + mi[d][6]] = F[5];
            ff[nm * ss + mi[d][7]] = F[6];
            fc[2 * ss + 0] = 0.5 * (ql[2 * q + 0] + qr[2 * q + 0]);
            ff[nm * ss + mi[d][5]] -= fc[2 * ss + 0];
            fc[2 * ss + 1] = 0.5 * (ql[2 * q + 1] + qr[2 * q + 1]);
            ff[nm * ss + mi[d][6]] -= fc[2 * ss + 1];
          }
          free(ul); free(ur); free(ql); free(qr);
        }

      /* EMF contributions of this sweep:
       *   t-flux (B_t) -> E_u interpolated along t, sign -, weight 1-ct[E_u]
       *   u-flux (B_u) -> E_t interpolated along u, sign +, weight ct[E_t]
       * ct planes: ct[0] for Ez, ct[1] for Ex, ct[2] for Ey.                 */
      double *Eax[3] = {ex, ey, ez};
      const int ctix[3] = {1, 2, 0}; /* ct plane of E_x, E_y, E_z */
      for (int part = 0; part < 2; part++) {
        const int comp = (part == 0) ? 5 : 6;              /* ff slot in mi[d] */
        const int ax = (part == 0) ? ta : tb;              /* interpolation axis */
        const int eax = (part == 0) ? tb : ta;             /* E component */
        const double sgnE = (part == 0) ? -0.5 : +0.5;
        double *E = Eax[eax];
        const double *cw = ct + nd * ctix[eax];
        const long oi = strd[ax];
        const int ni = dims[ax];
        /* ranges of the reference (mhd3d_solve.cpp:259-328, 423-492,
         * 587-656): faces [3, nn-2) along the sweep axis, every cell along
         * the third axis, interpolation centres [2, ni-2) along ax */
        const int oth = 3 - d - ax;
        const int no = dims[oth];
        const long so = strd[oth];
        const int olo = 0, ohi = no, nlo = 3, nhi = nn - 2;
        int b2;
#pragma omp parallel for private(b2)
        for (int aN = nlo; aN < nhi; aN++)
          for (b2 = olo; b2 < ohi; b2++) {
            double *tl = malloc(sizeof(double) * 2 * ni), *tr = malloc(sizeof(double) * 2 * ni);
            long base = o * aN + so * b2;
            int q;
            for (q = 2; q < ni - 2; q++) {
              long ss = base + oi * q;
              double a0, b0;
              lr_single(&ff[nm * ss + mi[d][comp]], (int)(nm * oi), lin, &a0, &b0);
              tl[2 * (q + 1) + 0] = a0;
              tr[2 * q + 0] = b0;
              lr_single(&fc[2 * ss + (comp - 5)], (int)(2 * oi), lin, &a0, &b0);
              tl[2 * (q + 1) + 1] = a0;
              tr[2 * q + 1] = b0;
            }
            for (q = 3; q < ni - 2; q++) {
              long ss = base + oi * q;
              double w = (part == 0) ? (1.0 - cw[ss]) : cw[ss];
              E[ss] += sgnE * ((tl[2 * q + 0] + tr[2 * q + 0]) + w * (tl[2 * q + 1] + tr[2 * q + 1]));
            }
            free(tl); free(tr);
          }
      }
    }

    /* cell-centred update: ro, mx, my, mz, en */
#pragma omp parallel for private(i, j)
    for (k = c->zoff; k < nz - c->zoff; k++)
      for (j = c->yoff; j < ny - c->yoff; j++)
        for (i = c->xoff; i < nx - c->xoff; i++) {
          static const int upd[5] = {0, 1, 2, 3, 7};
          long ss = nx * (ny * (long)k + j) + i;
          int q;
          for (q = 0; q < 5; q++) {
            int v = upd[q];
            double fxs[4] = {fx[nm * (ss - 1) + v], fx[nm * ss + v], fx[nm * (ss + 1) + v], fx[nm * (ss + 2) + v]};
            double fys[4] = {fy[nm * (ss - nx) + v], fy[nm * ss + v], fy[nm * (ss + nx) + v], fy[nm * (ss + 2 * nx) + v]};
            double fzs[4] = {fz[nm * (ss - nxy) + v], fz[nm * ss + v], fz[nm * (ss + nxy) + v],
                             fz[nm * (ss + 2 * nxy) + v]};
            double du0 = fdf(&fxs[1]), du1 = fdf(&fys[1]), du2 = fdf(&fzs[1]);
            rk_upd(&val[v][ss], ut[v * nd + ss], (-dtdx * du0 - dtdy * du1 - dtdz * du2), rf[0], rf[1]);
          }
        }
    /* CT (mhd_updt3d_ctb): b_n += -dt_a * d_a(E_b) + ... in the reference's
     * argument convention */
#pragma omp parallel for private(i, j)
    for (k = c->zoff; k < nz - c->zoff; k++) {
      for (j = c->yoff; j < ny - c->yoff; j++)
        for (i = c->xoff; i < nx - c->xoff + 1; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          double e1[4] = {ez[ss - nx], ez[ss], ez[ss + nx], ez[ss + 2 * nx]};
          double e2[4] = {ey[ss - nxy], ey[ss], ey[ss + nxy], ey[ss + 2 * nxy]};
          double du0 = +fdf(&e1[1]);
          double du1 = -fdf(&e2[1]);
          rk_upd(&bx[ss], ut[4 * nd + ss], -dtdy * du0 - dtdz * du1, rf[0], rf[1]);
        }
      for (j = c->yoff; j < ny - c->yoff + 1; j++)
        for (i = c->xoff; i < nx - c->xoff; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          double e1[4] = {ex[ss - nxy], ex[ss], ex[ss + nxy], ex[ss + 2 * nxy]};
          double e2[4] = {ez[ss - 1], ez[ss], ez[ss + 1], ez[ss + 2]};
          double du0 = +fdf(&e1[1]);
          double du1 = -fdf(&e2[1]);
          rk_upd(&by[ss], ut[5 * nd + ss], -dtdz * du0 - dtdx * du1, rf[0], rf[1]);
        }
    }
#pragma omp parallel for private(i, j)
    for (k = c->zoff; k < nz - c->zoff + 1; k++)
      for (j = c->yoff; j < ny - c->yoff; j++)
        for (i = c->xoff; i < nx - c->xoff; i++) {
          long ss = nx * (ny * (long)k + j) + i;
          double e1[4] = {ey[ss - 1], ey[ss], ey[ss + 1], ey[ss + 2]};
          double e2[4] = {ex[ss - nx], ex[ss], ex[ss + nx], ex[ss + 2 * nx]};
          double du0 = +fdf(&e1[1]);
          double du1 = -fdf(&e2[1]);
          rk_upd(&bz[ss], ut[6 * nd + ss], -dtdx * du0 - dtdy * du1, rf[0], rf[1]);
        }
    qor_bound(c, val);
  }
  free(ut); free(fx); free(fy); free(fz); free(fc); free(ex); free(ey); free(ez); free(ct);
  free(dvx); free(dvy); free(dvz); free(vx); free(vy); free(vz); free(pr); free(cx); free(cy); free(cz);
}

void qor_step(const qor_cfg *c, double *const val[8], double dt)
{
  if (c->ndim == 1)
    step1d(c, val, dt);
  else if (c->ndim == 2)
    step2d(c, val, dt);
  else
    step3d(c, val, dt);
}

/* ---------------------------------------------------------------------- */
double qor_cfl_dt(const qor_cfg *c, double *const val[8])
{
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const long nxy = (long)nx * ny;
  const double *ro = val[0], *mx = val[1], *my = val[2], *mz = val[3], *bx = val[4], *by = val[5],
               *bz = val[6], *en = val[7];
  const double gam = c->gam;
  double vmax = 1.0;
  int i, j, k;
  if (c->ndim == 1) {
    for (i = c->xoff; i < nx - c->xoff; i++) {
      double iro = 1.0 / ro[i];
      double vx = mx[i] * iro, vy = my[i] * iro, vz = mz[i] * iro;
      double pr = (gam - 1) * (en[i] - 0.5 * (ro[i] * (vx * vx + vy * vy + vz * vz)
                                              + (bx[i] * bx[i] + by[i] * by[i] + bz[i] * bz[i])));
      double vf = sqrt((gam * pr + (bx[i] * bx[i] + by[i] * by[i] + bz[i] * bz[i])) / ro[i]);
      double v = fabs(vx) + vf;
      if (v > vmax) vmax = v;
    }
    return c->cfl * c->dx / vmax;
  }
  if (c->ndim == 2) {
    for (j = c->yoff; j < ny - c->yoff; j++)
      for (i = c->xoff; i < nx - c->xoff; i++) {
        long ss = (long)nx * j + i;
        double cx = 0.5 * (bx[ss] + bx[(long)nx * j + (i + c->st[0][4])]);
        double cy = 0.5 * (by[ss] + by[(long)nx * (j + c->st[1][5]) + i]);
        double cz = bz[ss];
        double iro = 1.0 / ro[ss];
        double vx = mx[ss] * iro, vy = my[ss] * iro, vz = mz[ss] * iro;
        double eng = en[ss];
        if (c->phi) { /* the potential round trip of setdt, in place (RTI/mhd2d_class.cpp:46-48) */
          eng = en[ss] - ro[ss] * c->phi[ss];
          val[7][ss] = eng + ro[ss] * c->phi[ss];
        }
        double pr = (gam - 1) * (eng - 0.5 * (ro[ss] * (vx * vx + vy * vy + vz * vz) + (cx * cx + cy * cy + cz * cz)));
        double vf = sqrt((gam * pr + (cx * cx + cy * cy + cz * cz)) / ro[ss]);
        double v = sqrt(vx * vx + vy * vy) + vf;
        if (v > vmax) vmax = v;
      }
    double dr = omin(c->dx, c->dy);
    return c->cfl * dr / vmax;
  }
  for (k = c->zoff; k < nz - c->zoff; k++)
    for (j = c->yoff; j < ny - c->yoff; j++)
      for (i = c->xoff; i < nx - c->xoff; i++) {
        long ss = nx * (ny * (long)k + j) + i;
        double cx = 0.5 * (bx[ss] + bx[nx * (ny * (long)k + j) + (i + c->st[0][4])]);
        double cy = 0.5 * (by[ss] + by[nx * (ny * (long)k + (j + c->st[1][5])) + i]);
        double cz = 0.5 * (bz[ss] + bz[nx * (ny * (long)(k + c->st[2][6]) + j) + i]);
        double iro = 1.0 / ro[ss];
        double vx = mx[ss] * iro, vy = my[ss] * iro, vz = mz[ss] * iro;
        double pr = (gam - 1) * (en[ss] - 0.5 * (ro[ss] * (vx * vx + vy * vy + vz * vz) + (cx * cx + cy * cy + cz * cz)));
        double vf = sqrt((gam * pr + (cx * cx + cy * cy + cz * cz)) / ro[ss]);
        double v = sqrt(vx * vx + vy * vy + vz * vz) + vf;
        if (v > vmax) vmax = v;
      }
  (void)nxy;
  double dr = omin(omin(c->dx, c->dy), c->dz);
  return c->cfl * dr / vmax;
}

void qor_run(const qor_cfg *c, double *const val[8], int nsteps, int flg, double *dt, double *dt_hist)
{
  int n;
  for (n = 1; n <= nsteps; n++) {
    if (dt_hist) dt_hist[n - 1] = *dt;
    qor_bound(c, val);
    qor_step(c, val, *dt);
    if (flg && (n % 2 == 0)) *dt = qor_cfl_dt(c, val);
  }
}

/* ---------------------------------------------------------------------- */
/* Dissipative step: DMHD2D::dsptv (2D/SERIAL/MRX/dmhd2d_solve.cpp:3-151).   */

/* calc_flux_viscous, common/mhd_flux.c:700-753: viscous fluxes of momentum
 * and energy at face i-1/2 of the cell at v*, offsets as the reference's. */
static void visc_flux(const double *vx, const double *vy, const double *vz, double rnu, double idx,
                      double idy, double idz, long xo, long yo, long zo, c_fn fc, c_fn df,
                      double *fmx, double *fmy, double *fmz, double *fen)
{
  const long ss[4] = {-2 * xo, -xo, 0, +xo};
  const double c2 = 2. / 3., c4 = 4. / 3.;
  double vxi[4], vyi[4], vzi[4], dv[4];
  int k;
  for (k = 0; k < 4; k++) {
    vxi[k] = vx[ss[k]];
    vyi[k] = vy[ss[k]];
    vzi[k] = vz[ss[k]];
  }
  const double vx0 = fc(&vxi[1]), vy0 = fc(&vyi[1]), vz0 = fc(&vzi[1]);
  const double dvxdx = df(&vxi[1]) * idx;
  const double dvydx = df(&vyi[1]) * idx;
  const double dvzdx = df(&vzi[1]) * idx;
  for (k = 0; k < 4; k++) dv[k] = 0.5 * (vx[ss[k] + yo] - vx[ss[k] - yo]);
  const double dvxdy = fc(&dv[1]) * idy;
  for (k = 0; k < 4; k++) dv[k] = 0.5 * (vy[ss[k] + yo] - vy[ss[k] - yo]);
  const double dvydy = fc(&dv[1]) * idy;
  for (k = 0; k < 4; k++) dv[k] = 0.5 * (vx[ss[k] + zo] - vx[ss[k] - zo]);
  const double dvxdz = fc(&dv[1]) * idz;
  for (k = 0; k < 4; k++) dv[k] = 0.5 * (vz[ss[k] + zo] - vz[ss[k] - zo]);
  const double dvzdz = fc(&dv[1]) * idz;
  *fmx = -rnu * (c4 * dvxdx - c2 * (dvydy + dvzdz));
  *fmy = -rnu * (dvydx + dvxdy);
  *fmz = -rnu * (dvzdx + dvxdz);
  *fen = (*fmx) * vx0 + (*fmy) * vy0 + (*fmz) * vz0;
}

/* calc_flux_resistive, common/mhd_flux.c:755-782 */
static void res_flux(const double *by, const double *bz, const double *ey, const double *ez, long xo, long yo,
                     long zo, c_fn fc, double *fby, double *fbz, double *fen)
{
  const long ss[4] = {-2 * xo, -xo, 0, +xo};
  const double ey0 = 0.5 * (ey[0] + ey[zo]);
  const double ez0 = 0.5 * (ez[0] + ez[yo]);
  double byi[4], bzi[4];
  int k;
  for (k = 0; k < 4; k++) {
    byi[k] = 0.5 * (by[ss[k]] + by[ss[k] + yo]);
    bzi[k] = 0.5 * (bz[ss[k]] + bz[ss[k] + zo]);
  }
  const double by0 = fc(&byi[1]);
  const double bz0 = fc(&bzi[1]);
  *fby = -ez0;
  *fbz = +ey0;
  *fen = ey0 * bz0 - ez0 * by0;
}

/* mhd_ct_eres, common/mhd_func.c:78-106: eta j at the CT edges of cell b*. */
static void ct_eres(const double *bx, const double *by, const double *bz, const double *eta, double idx,
                    double idy, double idz, long xo, long yo, long zo, c_fn df, double *ex, double *ey,
                    double *ez)
{
  const long si[4] = {-2 * xo, -xo, 0, +xo}, sj[4] = {-2 * yo, -yo, 0, +yo}, sk[4] = {-2 * zo, -zo, 0, +zo};
  double bxj[4], bxk[4], byk[4], byi[4], bzi[4], bzj[4];
  int k;
  for (k = 0; k < 4; k++) {
    bxj[k] = bx[sj[k]];
    bxk[k] = bx[sk[k]];
    byk[k] = by[sk[k]];
    byi[k] = by[si[k]];
    bzi[k] = bz[si[k]];
    bzj[k] = bz[sj[k]];
  }
  const double jx = df(&bzj[1]) * idy - df(&byk[1]) * idz;
  const double jy = df(&bxk[1]) * idz - df(&bzi[1]) * idx;
  const double jz = df(&byi[1]) * idx - df(&bxj[1]) * idy;
  *ex = 0.25 * (eta[-yo - zo] + eta[-zo] + eta[-yo] + eta[0]) * jx;
  *ey = 0.25 * (eta[-zo - xo] + eta[-xo] + eta[-zo] + eta[0]) * jy;
  *ez = 0.25 * (eta[-xo - yo] + eta[-yo] + eta[-xo] + eta[0]) * jz;
}

void qor_dsptv(const qor_cfg *c, double *const val[8], const double *nu, const double *eta, double dt)
{
  const int nx = c->nx, ny = c->ny;
  const long nxy = (long)nx * ny;
  const double idx = 1.0 / c->dx, idy = 1.0 / c->dy;
  const double dtdx = dt * idx, dtdy = dt * idy;
  c_fn fc = fcen_tab[c->odr - 1];
  c_fn df = df_tab[c->odr - 1];
  double *ut = (double *)malloc(sizeof(double) * 8 * nxy);
  double *fx = (double *)calloc(8 * nxy, sizeof(double));
  double *fy = (double *)calloc(8 * nxy, sizeof(double));
  double *ex = (double *)calloc(nxy, sizeof(double));
  double *ey = (double *)calloc(nxy, sizeof(double));
  double *ez = (double *)calloc(nxy, sizeof(double));
  double *vx = (double *)malloc(sizeof(double) * nxy);
  double *vy = (double *)malloc(sizeof(double) * nxy);
  double *vz = (double *)malloc(sizeof(double) * nxy);
  const double *ro = val[0];
  double *bx = val[4], *by = val[5], *bz = val[6];
  int m, rk;
  long ss;
  for (m = 0; m < 8; m++) memcpy(ut + m * nxy, val[m], sizeof(double) * nxy);
  for (rk = 0; rk < c->r_k; rk++) {
    const double *rf = rk_coef(c->r_k, rk);
    int i, j;
    for (ss = 0; ss < nxy; ss++) { /* vlcty, common/mhd_class.hpp:59-66 */
      double iro = 1.0 / ro[ss];
      vx[ss] = val[1][ss] * iro;
      vy[ss] = val[2][ss] * iro;
      vz[ss] = val[3][ss] * iro;
    }
    for (j = 2; j < ny - 1; j++)
      for (i = 2; i < nx - 1; i++) {
        ss = (long)nx * j + i;
        ct_eres(&bx[ss], &by[ss], &bz[ss], &eta[ss], idx, idy, 0.0, 1, nx, 0, df, &ex[ss], &ey[ss], &ez[ss]);
      }
    /* faces whose fluxes the updates read (the reference computes j, i in
     * [2, n-2]; the rest are never used) */
    for (j = 2; j < ny - 1; j++)
      for (i = 2; i < nx - 1; i++) {
        double rnu, flux[2][8] = {{0.0}, {0.0}};
        ss = (long)nx * j + i;
        if (j >= c->yoff && j < ny - c->yoff && i >= 3 && i <= nx - 3) {
          rnu = 0.5 * (ro[ss - 1] * nu[ss - 1] + ro[ss] * nu[ss]);
          visc_flux(&vx[ss], &vy[ss], &vz[ss], rnu, idx, idy, 0.0, 1, nx, 0, fc, df, &flux[0][1], &flux[0][2],
                    &flux[0][3], &flux[0][7]);
          res_flux(&by[ss], &bz[ss], &ey[ss], &ez[ss], 1, nx, 0, fc, &flux[1][5], &flux[1][6], &flux[1][7]);
          for (m = 0; m < 8; m++) fx[8 * ss + m] = (m == 4) ? 0 : flux[0][m] + flux[1][m];
        }
        if (i >= c->xoff && i < nx - c->xoff && j >= 3 && j <= ny - 3) {
          double g[2][8] = {{0.0}, {0.0}};
          rnu = 0.5 * (ro[ss - nx] * nu[ss - nx] + ro[ss] * nu[ss]);
          visc_flux(&vy[ss], &vz[ss], &vx[ss], rnu, idy, 0.0, idx, nx, 0, 1, fc, df, &g[0][1], &g[0][2], &g[0][3],
                    &g[0][7]);
          res_flux(&bz[ss], &bx[ss], &ez[ss], &ex[ss], nx, 0, 1, fc, &g[1][5], &g[1][6], &g[1][7]);
          fy[8 * ss + 0] = g[0][0] + g[1][0];
          fy[8 * ss + 2] = g[0][1] + g[1][1];
          fy[8 * ss + 3] = g[0][2] + g[1][2];
          fy[8 * ss + 1] = g[0][3] + g[1][3];
          fy[8 * ss + 5] = 0;
          fy[8 * ss + 6] = g[0][5] + g[1][5];
          fy[8 * ss + 4] = g[0][6] + g[1][6];
          fy[8 * ss + 7] = g[0][7] + g[1][7];
        }
      }
    /* cell update (mhd_updt2d, mhd_func.c:222-248) of ro mx my mz bz en */
    for (j = c->yoff; j < ny - c->yoff; j++)
      for (i = c->xoff; i < nx - c->xoff; i++) {
        static const int vm[6] = {0, 1, 2, 3, 6, 7};
        int v;
        ss = (long)nx * j + i;
        for (v = 0; v < 6; v++) {
          const int q = vm[v];
          double a4[4], b4[4];
          int k;
          for (k = 0; k < 4; k++) {
            a4[k] = fx[8 * (ss + (k - 1)) + q];
            b4[k] = fy[8 * (ss + (long)(k - 1) * nx) + q];
          }
          const double du0 = df(&a4[1]), du1 = df(&b4[1]);
          rk_upd(&val[q][ss], ut[q * nxy + ss], (-dtdx * du0 - dtdy * du1), rf[0], rf[1]);
        }
      }
    /* CT (mhd_updt2d_ctb, mhd_func.c:286-300) */
    for (j = c->yoff; j < ny - c->yoff; j++)
      for (i = c->xoff; i < nx - c->xoff + 1; i++) {
        double e4[4];
        int k;
        ss = (long)nx * j + i;
        for (k = 0; k < 4; k++) e4[k] = ez[ss + (long)(k - 1) * nx];
        const double dtdy_ct = +dtdy;
        rk_upd(&bx[ss], ut[4 * nxy + ss], -dtdy_ct * df(&e4[1]), rf[0], rf[1]);
      }
    for (j = c->yoff; j < ny - c->yoff + 1; j++)
      for (i = c->xoff; i < nx - c->xoff; i++) {
        double e4[4];
        int k;
        ss = (long)nx * j + i;
        for (k = 0; k < 4; k++) e4[k] = ez[ss + (k - 1)];
        const double dtdy_ct = -dtdx;
        rk_upd(&by[ss], ut[5 * nxy + ss], -dtdy_ct * df(&e4[1]), rf[0], rf[1]);
      }
    qor_bound(c, val);
  }
  free(ut);
  free(fx);
  free(fy);
  free(ex);
  free(ey);
  free(ez);
  free(vx);
  free(vy);
  free(vz);
}

/* DMHD2D::exec_ step body (dmhd2d_class.cpp:15-33) without the output:
 * bound; ideal(dt); if dcmax > 1e-15: ncmax = 1 + (int)(2 (dcmax dt) / dr^2)
 * times {bound; dsptv(dt / ncmax)}; then setdt when flg and n even.       */
void qor_run_dissipative(const qor_cfg *c, double *const val[8], const double *nu, const double *eta,
                         double dcmax, int nsteps, int flg, double *dt, double *dt_hist)
{
  const double dr = omin(c->dx, c->dy);
  int n, nc;
  for (n = 1; n <= nsteps; n++) {
    if (dt_hist) dt_hist[n - 1] = *dt;
    qor_bound(c, val);
    qor_step(c, val, *dt);
    if (dcmax > 1e-15) {
      const int ncmax = 1 + (int)(2 * (dcmax * (*dt)) / (dr * dr));
      for (nc = 0; nc < ncmax; nc++) {
        qor_bound(c, val);
        qor_dsptv(c, val, nu, eta, (*dt) / ncmax);
      }
    }
    if (flg && (n % 2 == 0)) *dt = qor_cfl_dt(c, val);
  }
}

/* ---------------------------------------------------------------------- */
void qor_interp_batch(int kind, int n, const double *f5, double *fl, double *fr)
{
  lr_fn f = (kind >= 10) ? linear_tab[kind - 10] : interp_tab[kind];
  for (int q = 0; q < n; q++) f(&f5[5 * q + 2], &fl[q], &fr[q]);
}

void qor_fcen_batch(int odr, int n, const double *f4, double *out)
{
  for (int q = 0; q < n; q++) out[q] = fcen_tab[odr - 1](&f4[4 * q + 1]);
}

void qor_df_batch(int odr, int n, const double *f4, double *out)
{
  for (int q = 0; q < n; q++) out[q] = df_tab[odr - 1](&f4[4 * q + 1]);
}

void qor_lrstate_batch(int odr, int n, const double *st, const double *bn, double gam, double *vl, double *vr)
{
  for (int q = 0; q < n; q++) {
    double s[7][5];
    memcpy(s, st + 35 * (long)q, sizeof(s));
    lrstate(s, bn[q], gam, interp_tab[odr - 1], &vl[7 * (long)q], &vr[7 * (long)q]);
  }
}

void qor_flux_batch(int rmn, int n, const double *ql, const double *qr, const double *bn, double gam,
                    const double *dv, double *f)
{
  for (int q = 0; q < n; q++)
    riemann(rmn, &ql[7 * (long)q], &qr[7 * (long)q], bn[q], gam, &dv[2 * (long)q], &f[7 * (long)q]);
}

void qor_cuct_batch(int n, const double *c, const double *b, double *out)
{
  for (int q = 0; q < n; q++) {
    /* lay the 2x2 cells out as a 2x2 grid with strides (1, 2) */
    double ro[4], va[4], vb[4], ba[4], bb[4];
    for (int t = 0; t < 4; t++) {
      ro[t] = c[12 * q + t];
      va[t] = c[12 * q + 4 + t];
      vb[t] = c[12 * q + 8 + t];
      ba[t] = 0.0;
      bb[t] = 0.0;
    }
    ba[1] = b[4 * q + 0]; /* bx(i, j-1) */
    ba[3] = b[4 * q + 1]; /* bx(i, j)   */
    bb[2] = b[4 * q + 2]; /* by(i-1, j) */
    bb[3] = b[4 * q + 3]; /* by(i, j)   */
    out[q] = cuct_weight(&ro[3], &va[3], &vb[3], &ba[3], &bb[3], 1, 2);
  }
}
