// qas_adjoint_check.hip -- host-side dot-product checks of the per-point
// adjoints of qas_adjoint.hpp against central finite differences of the
// forward functions of qas_device.hpp.  Runs on the CPU (the functions are
// __host__ __device__); no device call is made.
//
// For random x, direction d and output weights g:
//   D(h) = g . (f(x + h d) - f(x - h d)) / (2 h),   fd = (4 D(h/2) - D(h)) / 3,
//   ad = (J^T g) . d
// and relerr = |fd - ad| / max(|fd|, |ad|).  A branch that flips inside
// [x - h d, x + h d] (a min/max, an HLLD region flag) gives a large error in
// that trial only; callers test a quantile, not every trial.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/qasmhd.h"
#include "qas_adjoint.hpp"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
  double uni()  // [0, 1)
  {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  double in(double a, double b) { return a + (b - a) * uni(); }
};

// primitive order of the reconstruction: ro, vn, vt, vu, bt, bu, pr
double prim_sample(Rng& r, int k)
{
  if (k == 0) return r.in(0.5, 2.0);
  if (k == 6) return r.in(0.3, 2.0);
  return r.in(-1.0, 1.0);
}

struct Wcns {
  static const int NI = 5, NO = 2;
  static void sample(Rng& r, double* x) { for (int i = 0; i < 5; i++) x[i] = r.in(-1.0, 1.0); }
  static void fwd(const double* x, double* y) { qas::interp5<4>(x, y[0], y[1]); }
  static void adj(const double* x, const double* g, double* gx) { qas::wcns4_adj(x, g[0], g[1], gx); }
};

struct Lin4 {
  static const int NI = 5, NO = 2;
  static void sample(Rng& r, double* x) { for (int i = 0; i < 5; i++) x[i] = r.in(-1.0, 1.0); }
  static void fwd(const double* x, double* y) { qas::lin5<4>(x, y[0], y[1]); }
  static void adj(const double*, const double* g, double* gx) { qas::lin4_adj(g[0], g[1], gx); }
};

// x = s[7][5] (row-major) then bn
struct Lrstate {
  static const int NI = 36, NO = 14;
  static constexpr double gam = 5.0 / 3.0;
  static void sample(Rng& r, double* x)
  {
    // a smooth-ish profile plus noise, so both projection branches occur
    const double mach = r.in(0.0, 2.0);
    for (int k = 0; k < 7; k++) {
      const double base = prim_sample(r, k) * ((k >= 1 && k <= 3) ? mach : 1.0);
      for (int m = 0; m < 5; m++) x[5 * k + m] = base * (1.0 + 0.2 * r.in(-1.0, 1.0));
    }
    x[35] = r.in(-1.0, 1.0);
  }
  static void unpack(const double* x, double (&s)[7][5])
  {
    for (int k = 0; k < 7; k++)
      for (int m = 0; m < 5; m++) s[k][m] = x[5 * k + m];
  }
  static void fwd(const double* x, double* y)
  {
    double s[7][5], vl[7], vr[7];
    unpack(x, s);
    qas::lrstate<4, double, 0>(s, x[35], gam, vl, vr);
    for (int k = 0; k < 7; k++) {
      y[k] = vl[k];
      y[7 + k] = vr[k];
    }
  }
  static void adj(const double* x, const double* g, double* gx)
  {
    double s[7][5], gs[7][5] = {}, gvl[7], gvr[7], gbn = 0.0;
    unpack(x, s);
    for (int k = 0; k < 7; k++) {
      gvl[k] = g[k];
      gvr[k] = g[7 + k];
    }
    qas::lrstate_adj(s, x[35], gam, gvl, gvr, gs, gbn);
    for (int k = 0; k < 7; k++)
      for (int m = 0; m < 5; m++) gx[5 * k + m] += gs[k][m];
    gx[35] += gbn;
  }
};

// x = p[5], q[5], bnm, bnp
struct LrFb {
  static const int NI = 12, NO = 2;
  static void sample(Rng& r, double* x) { for (int i = 0; i < 12; i++) x[i] = r.in(-1.0, 1.0); }
  static void fwd(const double* x, double* y) { qas::lr_fb<4>(x, x + 5, x[10], x[11], y[0], y[1]); }
  static void adj(const double* x, const double* g, double* gx)
  {
    qas::lr_fb_adj(x, x + 5, x[10], x[11], g[0], g[1], gx, gx + 5, gx[10], gx[11]);
  }
};

struct Cuct {
  static const int NI = 16, NO = 1;
  static void sample(Rng& r, double* x)
  {
    for (int i = 0; i < 4; i++) x[i] = r.in(0.5, 2.0);
    for (int i = 4; i < 16; i++) x[i] = r.in(-1.0, 1.0);
  }
  static void fwd(const double* x, double* y)
  {
    y[0] = qas::cuct_weight(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12], x[13],
                            x[14], x[15]);
  }
  static void adj(const double* x, const double* g, double* gx) { qas::cuct_weight_adj(x, g[0], gx); }
};

// x = L[7], R[7], bn
struct Hlld {
  static const int NI = 15, NO = 7;
  static constexpr double gam = 5.0 / 3.0;
  static void sample(Rng& r, double* x)
  {
    for (int k = 0; k < 7; k++) {
      const double base = prim_sample(r, k);
      x[k] = base;
      x[7 + k] = base * (1.0 + 0.5 * r.in(-1.0, 1.0)) + ((k >= 1 && k <= 5) ? 0.3 * r.in(-1.0, 1.0) : 0.0);
    }
    x[14] = r.in(-1.0, 1.0);
  }
  static void fwd(const double* x, double* y) { qas::flux_hlld_family<double>(false, x, x + 7, x[14], gam, 0.0, 0.0, y); }
  static void adj(const double* x, const double* g, double* gx)
  {
    qas::hlld_adj(x, x + 7, x[14], gam, g, gx, gx + 7, gx[14]);
  }
};

// x = L[7], R[7], bn, dv0, dv1
struct Mlau {
  static const int NI = 17, NO = 7;
  static constexpr double gam = 5.0 / 3.0;
  static void sample(Rng& r, double* x)
  {
    // Mach up to about 2 so that the sub- and supersonic polynomial branches
    // both occur; dv of either sign so that the pressure-diffusion weight
    // theta varies
    const double mach = r.in(0.0, 2.5);
    for (int k = 0; k < 7; k++) {
      const double base = prim_sample(r, k) * ((k >= 1 && k <= 3) ? mach : 1.0);
      x[k] = base;
      x[7 + k] = base * (1.0 + 0.5 * r.in(-1.0, 1.0)) + ((k >= 1 && k <= 5) ? 0.3 * r.in(-1.0, 1.0) : 0.0);
    }
    x[14] = r.in(-1.0, 1.0);
    x[15] = r.in(-1.0, 1.0);
    x[16] = r.in(-1.0, 1.0);
  }
  static void fwd(const double* x, double* y) { qas::flux_mlau<double>(x, x + 7, x[14], gam, x[15], x[16], y); }
  static void adj(const double* x, const double* g, double* gx)
  {
    qas::mlau_adj(x, x + 7, x[14], gam, x[15], x[16], g, gx, gx + 7, gx[14], gx[15], gx[16]);
  }
};

// x = L[7], R[7], bn, dv0, dv1 (sampled as for MLAU)
struct Lhlld {
  static const int NI = 17, NO = 7;
  static constexpr double gam = 5.0 / 3.0;
  static void sample(Rng& r, double* x) { Mlau::sample(r, x); }
  static void fwd(const double* x, double* y)
  {
    qas::flux_hlld_family<double>(true, x, x + 7, x[14], gam, x[15], x[16], y);
  }
  static void adj(const double* x, const double* g, double* gx)
  {
    qas::hlld_adj_g<true>(x, x + 7, x[14], gam, x[15], x[16], g, gx, gx + 7, gx[14], gx[15], gx[16]);
  }
};

template <class F>
void run_check(int ntrials, uint64_t seed, double* relerr)
{
  Rng r(seed);
  double x[F::NI], d[F::NI], xp[F::NI], xm[F::NI], gx[F::NI], g[F::NO], yp[F::NO], ym[F::NO];
  for (int t = 0; t < ntrials; t++) {
    F::sample(r, x);
    for (int i = 0; i < F::NI; i++) d[i] = r.in(-1.0, 1.0);
    for (int o = 0; o < F::NO; o++) g[o] = r.in(-1.0, 1.0);
    double xs = 0.0;
    for (int i = 0; i < F::NI; i++) xs = std::fmax(xs, std::fabs(x[i]));
    double fdh[2];
    for (int k = 0; k < 2; k++) {
      const double h = (k ? 0.5e-5 : 1e-5) * (xs > 0.0 ? xs : 1.0);
      for (int i = 0; i < F::NI; i++) {
        xp[i] = x[i] + h * d[i];
        xm[i] = x[i] - h * d[i];
      }
      F::fwd(xp, yp);
      F::fwd(xm, ym);
      fdh[k] = 0.0;
      for (int o = 0; o < F::NO; o++) fdh[k] += g[o] * (yp[o] - ym[o]) / (2.0 * h);
    }
    // Richardson: cancels the h^2 term of the central difference
    const double fd = (4.0 * fdh[1] - fdh[0]) / 3.0;
    std::memset(gx, 0, sizeof(gx));
    F::adj(x, g, gx);
    double ad = 0.0;
    for (int i = 0; i < F::NI; i++) ad += gx[i] * d[i];
    const double sc = std::fmax(std::fabs(fd), std::fabs(ad));
    relerr[t] = sc > 0.0 ? std::fabs(fd - ad) / sc : 0.0;
  }
}

}  // namespace

extern "C" int qas_adjoint_selfcheck(int which, int ntrials, unsigned long long seed, double* relerr)
{
  if (ntrials <= 0 || !relerr) return QAS_E_ARG;
  switch (which) {
    case QAS_ADJ_WCNS4: run_check<Wcns>(ntrials, seed, relerr); break;
    case QAS_ADJ_LIN4: run_check<Lin4>(ntrials, seed, relerr); break;
    case QAS_ADJ_LRSTATE: run_check<Lrstate>(ntrials, seed, relerr); break;
    case QAS_ADJ_LRFB: run_check<LrFb>(ntrials, seed, relerr); break;
    case QAS_ADJ_CUCT: run_check<Cuct>(ntrials, seed, relerr); break;
    case QAS_ADJ_HLLD: run_check<Hlld>(ntrials, seed, relerr); break;
    case QAS_ADJ_MLAU: run_check<Mlau>(ntrials, seed, relerr); break;
    case QAS_ADJ_LHLLD: run_check<Lhlld>(ntrials, seed, relerr); break;
    default: return QAS_E_ARG;
  }
  return QAS_OK;
}
