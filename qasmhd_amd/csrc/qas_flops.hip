// qas_flops.hip -- algorithmic work model of the stage kernels (host code).
//
// The per-point functions of qas_device.hpp are instantiated on the host
// with an op-counting scalar: + - * / and sqrt count 1 each; negation,
// fabs, comparisons, min/max and selects count 0.  Structural zeros and
// unit entries of the eigenvector matrices are never evaluated by those
// functions, so they are not counted.  The characteristic reconstruction is
// counted for each projection separately (fast/Alfven and Alfven/entropy)
// and the cheaper one is used, so the model is a lower bound of the
// algorithmic work.  Bytes are the compulsory HBM traffic of one launch:
// every array element a kernel must read or write once.
#include <cmath>
#include <cstring>

#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

static thread_local long long g_ops = 0;

struct CT {
  double v;
  CT() : v(0) {}
  CT(double x) : v(x) {}
};
static inline CT operator+(CT a, CT b) { g_ops++; return CT(a.v + b.v); }
static inline CT operator-(CT a, CT b) { g_ops++; return CT(a.v - b.v); }
static inline CT operator*(CT a, CT b) { g_ops++; return CT(a.v * b.v); }
static inline CT operator/(CT a, CT b) { g_ops++; return CT(a.v / b.v); }
static inline CT operator-(CT a) { return CT(-a.v); }
static inline CT operator+(CT a) { return a; }
[[maybe_unused]] static inline CT& operator+=(CT& a, CT b) { a = a + b; return a; }
static inline CT& operator-=(CT& a, CT b) { a = a - b; return a; }
static inline CT& operator*=(CT& a, CT b) { a = a * b; return a; }
static inline CT& operator/=(CT& a, CT b) { a = a / b; return a; }
static inline bool operator>(CT a, CT b) { return a.v > b.v; }
[[maybe_unused]] static inline bool operator<(CT a, CT b) { return a.v < b.v; }
static inline bool operator>=(CT a, CT b) { return a.v >= b.v; }
static inline bool operator<=(CT a, CT b) { return a.v <= b.v; }
static inline bool operator==(CT a, CT b) { return a.v == b.v; }
static inline bool operator!=(CT a, CT b) { return a.v != b.v; }
static inline CT sqrt(CT a) { g_ops++; return CT(std::sqrt(a.v)); }
static inline CT fabs(CT a) { return CT(std::fabs(a.v)); }

// a smooth, moderately magnetised stencil (rho = gam^2, p = gam scale)
static void sample_stencil(CT (&s)[7][5])
{
  for (int m = 0; m < 5; m++) {
    const double x = 0.1 * (m - 2);
    s[0][m] = CT(2.78 + 0.2 * x);
    s[1][m] = CT(0.6 + 0.3 * x);
    s[2][m] = CT(-0.4 + 0.1 * x);
    s[3][m] = CT(0.2 - 0.2 * x);
    s[4][m] = CT(0.7 + 0.25 * x);
    s[5][m] = CT(-0.3 + 0.15 * x);
    s[6][m] = CT(1.67 - 0.1 * x);
  }
}

template <int ODR>
static void count_lrstate(double* fast, double* ae)
{
  CT s[7][5], vl[7], vr[7];
  sample_stencil(s);
  g_ops = 0;
  lrstate<ODR, CT, 1>(s, CT(0.5), CT(5.0 / 3.0), vl, vr);
  *fast = (double)g_ops;
  g_ops = 0;
  lrstate<ODR, CT, 2>(s, CT(0.5), CT(5.0 / 3.0), vl, vr);
  *ae = (double)g_ops;
}

template <int RMN>
static double count_riemann()
{
  CT L[7] = {CT(2.9), CT(0.62), CT(-0.41), CT(0.19), CT(0.72), CT(-0.31), CT(1.66)};
  CT R[7] = {CT(2.7), CT(0.58), CT(-0.38), CT(0.22), CT(0.69), CT(-0.28), CT(1.69)};
  CT F[7];
  g_ops = 0;
  riemann<RMN>(L, R, CT(0.5), CT(5.0 / 3.0), CT(-0.01), CT(-0.02), F);
  return (double)g_ops;
}

template <int ODR>
static double count_lrfb()
{
  CT p[5] = {CT(1), CT(1.1), CT(1.2), CT(1.25), CT(1.3)}, q[5] = {CT(0.1), CT(0.2), CT(0.25), CT(0.3), CT(0.2)};
  CT vl, vr;
  g_ops = 0;
  lr_fb<ODR>(p, q, CT(0.4), CT(0.45), vl, vr);
  return (double)g_ops;
}

template <int ODR>
static double count_lin()  // one side of lin5
{
  CT f[5] = {CT(1), CT(1.1), CT(1.2), CT(1.25), CT(1.3)}, a, b;
  g_ops = 0;
  lin5<ODR>(f, a, b);
  return 0.5 * (double)g_ops;
}

template <int ODR>
static double count_df()
{
  g_ops = 0;
  (void)dfc<ODR>(CT(1), CT(2), CT(3), CT(5));
  return (double)g_ops;
}

template <int ODR>
static double count_fcen()
{
  g_ops = 0;
  (void)fcen<ODR>(CT(1), CT(2), CT(3), CT(5));
  return (double)g_ops;
}

static double count_cuct()
{
  g_ops = 0;
  (void)cuct_weight(CT(1), CT(1.1), CT(1.2), CT(1.3), CT(0.1), CT(0.2), CT(0.3), CT(0.4), CT(0.5), CT(0.6),
                    CT(0.7), CT(0.8), CT(0.9), CT(1.0), CT(1.1), CT(1.2));
  return (double)g_ops;
}

struct PointCounts {
  double lr_fast, lr_ae, riem, lrfb, lin, df, fc;
};

static PointCounts counts(int rmn, int odr)
{
  PointCounts c;
  switch (odr) {
    case 1: count_lrstate<1>(&c.lr_fast, &c.lr_ae); c.lrfb = count_lrfb<1>(); c.lin = count_lin<1>();
            c.df = count_df<1>(); c.fc = count_fcen<1>(); break;
    case 2: count_lrstate<2>(&c.lr_fast, &c.lr_ae); c.lrfb = count_lrfb<2>(); c.lin = count_lin<2>();
            c.df = count_df<2>(); c.fc = count_fcen<2>(); break;
    case 3: count_lrstate<3>(&c.lr_fast, &c.lr_ae); c.lrfb = count_lrfb<3>(); c.lin = count_lin<3>();
            c.df = count_df<3>(); c.fc = count_fcen<3>(); break;
    default: count_lrstate<4>(&c.lr_fast, &c.lr_ae); c.lrfb = count_lrfb<4>(); c.lin = count_lin<4>();
             c.df = count_df<4>(); c.fc = count_fcen<4>(); break;
  }
  switch (rmn) {
    case 0: c.riem = count_riemann<0>(); break;
    case 1: c.riem = count_riemann<1>(); break;
    case 2: c.riem = count_riemann<2>(); break;
    default: c.riem = count_riemann<3>(); break;
  }
  return c;
}

}  // namespace qas

using namespace qas;

// Work of one launch of kernel `kid` (QAS_K_*) for this context.
extern "C" int qas_kernel_model(const qas_ctx* c, int kid, double* flops, double* bytes)
{
  if (!c || !flops || !bytes) return QAS_E_ARG;
  const PointCounts pc = counts(c->s.rmn, c->s.odr);
  const double lr = pc.lr_fast < pc.lr_ae ? pc.lr_fast : pc.lr_ae;
  const int nd3 = c->g.ndim;
  const double n[3] = {(double)c->g.nx, (double)c->g.ny, (double)c->g.nz};
  const double cells = n[0] * n[1] * n[2];
  const double inter = (n[0] - 8) * (nd3 >= 2 ? n[1] - 8 : 1) * (nd3 >= 3 ? n[2] - 8 : 1);
  const double dvf = c->s.rmn >= 2 ? (nd3 == 3 ? 9.0 : 5.0) : 0.0;  // shock-sensor work per face
  const int nupd = nd3 == 3 ? 5 : (nd3 == 2 ? 6 : 7);               // cell-updated variables
  const int ntan = nd3 == 3 ? 2 : (nd3 == 2 ? 1 : 0);                // B-flux splits per face
  *flops = 0;
  *bytes = 0;
  if (kid == QAS_K_SWEEP_X || kid == QAS_K_SWEEP_Y || kid == QAS_K_SWEEP_Z) {
    const int d = kid - QAS_K_SWEEP_X;
    if (d >= nd3) return QAS_E_ARG;
    const double lines = cells / n[d];
    const double lrc = (n[d] - 4) * lines;   // reconstructed cells
    const double fac = (n[d] - 5) * lines;   // solved faces
    *flops = lrc * (lr + ntan * pc.lrfb) + fac * (pc.riem + dvf + 3.0 * ntan);
    // The transverse sweeps run first and store dt/dy Dy (and dt/dz Dz):
    // difference + scale.  The x sweep scales its own difference, subtracts
    // the stored ones and applies the RK stage sum (2 mul + 2 add).
    const double per_var = d == 0 ? pc.df + 1.0 + (nd3 - 1) + 4.0 : pc.df + 1.0;
    *flops += inter * nupd * per_var;
    // bytes: 7 stencil arrays + cell-centred normal B + staggered B (read),
    // 2*ntan face EMF parts (write), accumulators / state
    double b = cells * (9.0 + 2.0 * ntan);
    if (nd3 == 1) b = cells * 8.0;
    if (nd3 == 2 && d == 1) {
      // the 2D y sweep computes the primitives of every cell itself from the
      // 8 state arrays and stores 6 of them for the x sweep
      *flops += cells * (1 + 3 + 5 + 5 + 4 + 2 * pc.fc);
      b = cells * (8.0 + 6.0 + 2.0 * ntan);
    }
    if (d == 0) b += inter * nupd * ((nd3 - 1) + 3.0);  // stored divergences + in, u0 read; out write
    else b += inter * nupd * 1.0;                        // this sweep's divergence written
    *bytes = b * 8.0;
    return QAS_OK;
  }
  if (kid == QAS_K_PRIM) {
    if (nd3 == 2) return QAS_OK;  // folded into the y sweep
    *flops = cells * (1 + 3 + 5 + 5 + 4 + (nd3 >= 2 ? nd3 * (pc.fc + 0) : 0));
    *bytes = cells * 8.0 * (8 + (nd3 >= 2 ? nd3 : 0) + 4);
    return QAS_OK;
  }
  if (kid == QAS_K_EMF) {
    const double ne = (n[0] - 5) * (nd3 >= 2 ? n[1] - 5 : 1) * (nd3 >= 3 ? n[2] - 5 : 1);
    const double ncomp = nd3 == 3 ? 3 : 1;
    const double part = 4 * pc.lin + 2 + 1 + 2 + 1;  // 4 one-sided interps, (uL+uR)+w(cL+cR), scale
    *flops = ne * ncomp * (2 * part + 2 + (c->s.ctw ? count_cuct() : 0));
    *bytes = ne * ncomp * 8.0 * (4 + 1) + (c->s.ctw ? ne * 8.0 * 5 : 0);
    return QAS_OK;
  }
  if (kid == QAS_K_CT) {
    const double nf = (n[0] - 7) * (nd3 >= 2 ? n[1] - 7 : 1) * (nd3 >= 3 ? n[2] - 7 : 1);
    const double ncomp = nd3 == 3 ? 3 : 2;
    *flops = nf * ncomp * ((nd3 == 3 ? 2 : 1) * (pc.df + 1) + 1 + 4);
    *bytes = nf * ncomp * 8.0 * (3 + (nd3 == 3 ? 2 : 1));
    return QAS_OK;
  }
  return QAS_E_ARG;
}

// Per-point counts (for documentation and tests): out[0] lrstate fast,
// out[1] lrstate Alfven/entropy, out[2] Riemann solver, out[3] lr_fb,
// out[4] one-sided linear interpolation, out[5] 4th-order difference,
// out[6] face-to-centre, out[7] CUCT weight.
extern "C" int qas_point_flops(int rmn, int odr, double* out)
{
  if (!out || rmn < 0 || rmn > 3 || odr < 1 || odr > 4) return QAS_E_ARG;
  const PointCounts pc = counts(rmn, odr);
  out[0] = pc.lr_fast;
  out[1] = pc.lr_ae;
  out[2] = pc.riem;
  out[3] = pc.lrfb;
  out[4] = pc.lin;
  out[5] = pc.df;
  out[6] = pc.fc;
  out[7] = count_cuct();
  return QAS_OK;
}
