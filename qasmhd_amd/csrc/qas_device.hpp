// qas_device.hpp -- per-point MHD arithmetic for the HIP kernels (gfx950).
//
// Every function keeps the floating-point operation order of the reference
// routine it replaces, so that with -ffp-contract=off the FP64
// instantiation is bit-identical to the reference:
//   interp5 / lin5 ...... common/common_interp.c:7-150 (interpol / l_interp
//                         tables, common/mhd_class.hpp:107-109)
//   fcen4 / df4 ......... common/common_interp.h:6-24
//   lrstate ............. common/mhd_func.c:108-161 with
//                         mhd_eigen.c:175-434 and 513-665
//   flux_* .............. common/mhd_flux.c:7-698
//   cuct_weight ......... common/mhd_func.c:56-76
// The characteristic projection is applied with the structural zeros and
// unit entries of the eigenvector matrices skipped.  Without FMA this is
// value-identical to the reference's dense 7x7 products (x + 0 = x and
// 1 * x = x exactly; only the sign of an exact zero can differ, which no
// later operation can observe since no denominator is an exact zero).  The
// fast/Alfven (supersonic) and Alfven/entropy (subsonic) projections are
// both formed and selected per lane, so a wavefront never diverges on the
// branch of mhd_func.c:143.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

// host+device: the host instantiation with an op-counting scalar gives the
// algorithmic flop model (qas_flops.hip)
#define QAS_HD __host__ __device__ __forceinline__

namespace qas {

// ternary max/min of common/common_func.h:4-14 (not fmax/fmin)
template <typename T> QAS_HD T tmax(T a, T b) { return (a >= b) ? a : b; }
template <typename T> QAS_HD T tmin(T a, T b) { return (a >= b) ? b : a; }

template <typename T> QAS_HD T dsqrt(T x) { return sqrt(x); }
template <typename T> QAS_HD T dabs(T x) { return fabs(x); }
template <> QAS_HD float dsqrt<float>(float x) { return sqrtf(x); }
template <> QAS_HD float dabs<float>(float x) { return fabsf(x); }

// 1.0 / d for 2^-767 <= d <= 2^1022, bit-identical to the IEEE division.
// On gfx950 an FP64 division is: operand scaling (v_div_scale, twice), a
// reciprocal estimate, two Newton steps, the quotient and its residual,
// the corrected quotient (v_div_fmas) and the special-case fix-up
// (v_div_fixup).  Inside that range neither operand of 1.0 / d is scaled,
// v_div_fmas is a plain FMA and the fix-up returns its input, so the same
// estimate, Newton steps and correction give the same bits in 7
// instructions instead of 11.  Callers check the range.
QAS_HD double rcp_inrange(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(e, r, r);
#else
  return 1.0 / d;
#endif
}

// ------------------------------------------------------------------ 1-D interpolation
// f[0..4] = stencil i-2..i+2; fl = left state at i+1/2, fr = right state at i-1/2.
template <int ODR, typename T>
QAS_HD void lin5(const T* f, T& fl, T& fr)
{
  if constexpr (ODR == 1) {
    fl = f[2];
    fr = f[2];
  } else if constexpr (ODR == 2) {
    fl = (-f[1] + T(5.0) * f[2] + T(2.0) * f[3]) / T(6.0);
    fr = (-f[3] + T(5.0) * f[2] + T(2.0) * f[1]) / T(6.0);
  } else if constexpr (ODR == 3) {
    fl = T(0.125) * (-f[1] + T(6.0) * f[2] + T(3.0) * f[3]);
    fr = T(0.125) * (-f[3] + T(6.0) * f[2] + T(3.0) * f[1]);
  } else {
    fl = T(0.003125) * (T(9.0) * f[0] - T(56.0) * f[1] + T(234.0) * f[2] + T(144.0) * f[3] - T(11.0) * f[4]);
    fr = T(0.003125) * (T(9.0) * f[4] - T(56.0) * f[3] + T(234.0) * f[2] + T(144.0) * f[1] - T(11.0) * f[0]);
  }
}

// eps of the WCNS smoothness weights 1/(is^2 + eps): the reference's 1e-40
// in FP64.  1e-40 is not a normal FP32 number and 1/1e-40 overflows FP32,
// so the FP32 (mixed) instantiation uses 1e-30, still far below is^2 of any
// non-uniform stencil.
template <typename T> QAS_HD T wcns_eps() { return T(1e-40); }
template <> QAS_HD float wcns_eps<float>() { return 1e-30f; }

template <typename T>
QAS_HD T minmod(T a, T b)
{
  return (+tmin(tmax(a, b), T(0.0)) + tmax(tmin(a, b), T(0.0)));
}

// The weights and face values of WCNS-4 (common_interp.c:102-150) from
// the three denominators d_k = IS_k^2 + eps; IN_RANGE takes the reciprocals
// with rcp_inrange (FP64, the caller checked the range), else 1.0 / x.
template <bool IN_RANGE, typename T>
QAS_HD void wcns4_weights(const T* f, T d0, T d1, T d2, T& fl, T& fr)
{
  auto inv = [](T x) -> T {
    if constexpr (IN_RANGE) return rcp_inrange(x);
    else return T(1.0) / x;
  };
  const T s0 = inv(d0);
  const T s1 = inv(d1);
  const T s2 = inv(d2);
  // sub-stencil values (shared by both sides)
  const T qa = T(3.0) * f[0] - T(10.0) * f[1] + T(15.0) * f[2];
  const T qb = -f[1] + T(6.0) * f[2] + T(3.0) * f[3];
  const T qc = +T(3.0) * f[2] + T(6.0) * f[3] - f[4];
  const T ra = T(3.0) * f[4] - T(10.0) * f[3] + T(15.0) * f[2];
  const T rb = -f[3] + T(6.0) * f[2] + T(3.0) * f[1];
  const T rc = +T(3.0) * f[2] + T(6.0) * f[1] - f[0];
  T w0 = T(0.075) * s0;
  T w1 = T(0.650) * s1;
  T w2 = T(0.275) * s2;
  T den = inv(w0 + w1 + w2);
  w0 *= den;
  w1 *= den;
  w2 = T(1.0) - w0 - w1;
  fl = T(0.125) * (+w0 * qa + w1 * qb + w2 * qc);
  w0 = T(0.275) * s0;
  w1 = T(0.650) * s1;
  w2 = T(0.075) * s2;
  den = inv(w0 + w1 + w2);
  w0 *= den;
  w1 *= den;
  w2 = T(1.0) - w0 - w1;
  fr = T(0.125) * (+w2 * ra + w1 * rb + w0 * rc);
}

// INRANGE: the WCNS-4 reciprocals take rcp_inrange when in range (FP64);
// false keeps the plain division (a smaller body, for the adjoint kernels)
template <int ODR, typename T, bool INRANGE = true>
QAS_HD void interp5(const T* f, T& fl, T& fr)
{
  if constexpr (ODR == 1) {
    fl = f[2];
    fr = f[2];
  } else if constexpr (ODR == 2) {  // MUSCL minmod
    T d = T(0.5) * minmod(f[2] - f[1], f[3] - f[2]);
    fl = f[2] + d;
    fr = f[2] - d;
  } else if constexpr (ODR == 3) {  // WCNS3
    const T eps = wcns_eps<T>();
    T a = f[1] - f[2];
    T s0 = a * a;
    a = f[3] - f[2];
    T s1 = a * a;
    T t = f[3] - T(2.0) * f[2] + f[1];
    t = t * t;
    s0 = (T(1.0) + t / (s0 + eps));
    s1 = (T(1.0) + t / (s1 + eps));
    T w0 = T(0.25) * s0;
    T w1 = T(0.75) * s1;
    w0 /= (w0 + w1);
    w1 = T(1.0) - w0;
    fl = T(0.5) * (w0 * (-f[1] + T(3.0) * f[2]) + w1 * (f[2] + f[3]));
    w0 = T(0.75) * s0;
    w1 = T(0.25) * s1;
    w0 /= (w0 + w1);
    w1 = T(1.0) - w0;
    fr = T(0.5) * (w1 * (-f[3] + T(3.0) * f[2]) + w0 * (f[2] + f[1]));
  } else {  // WCNS4
    const T eps = wcns_eps<T>();
    const T c13 = T(13. / 12.);
    T a = f[0] - T(4.0) * f[1] + T(3.0) * f[2];
    T b = f[0] - T(2.0) * f[1] + f[2];
    const T s0 = c13 * b * b + T(0.25) * a * a;
    a = f[1] - f[3];
    b = f[1] - T(2.0) * f[2] + f[3];
    const T s1 = c13 * b * b + T(0.25) * a * a;
    a = f[4] - T(4.0) * f[3] + T(3.0) * f[2];
    b = f[4] - T(2.0) * f[3] + f[2];
    const T s2 = c13 * b * b + T(0.25) * a * a;
    const T d0 = s0 * s0 + eps, d1 = s1 * s1 + eps, d2 = s2 * s2 + eps;
    if constexpr (INRANGE && std::is_same<T, double>::value) {
      // d >= eps = 1e-40 always; d <= 2^700 keeps every 1/d in
      // [2^-700, 1e40] and both weight sums in [0.65 * 2^-700, 1e40], so
      // all five reciprocals are inside rcp_inrange's range (NaN fails
      // the test and takes the division)
      if ((d0 <= 0x1p700) & (d1 <= 0x1p700) & (d2 <= 0x1p700)) {
        wcns4_weights<true>(f, d0, d1, d2, fl, fr);
        return;
      }
    }
    wcns4_weights<false>(f, d0, d1, d2, fl, fr);
  }
}

// face -> centre (f0 = value at i-1/2) and 4th-order difference
template <int ODR, typename T>
QAS_HD T fcen(T fm1, T f0, T f1, T f2)
{
  if constexpr (ODR <= 2) return (T(0.5) * (f0 + f1));
  else return (T(0.0625) * (T(9.0) * (f0 + f1) - (fm1 + f2)));
}
template <int ODR, typename T>
QAS_HD T dfc(T fm1, T f0, T f1, T f2)
{
  if constexpr (ODR <= 2) return (f1 - f0);
  else return ((T(27.0) * (f1 - f0) - (f2 - fm1)) / T(24.0));
}

// ------------------------------------------------------------------ characteristic reconstruction
// s[k][m]: k = (ro, vn, vt, vu, bt, bu, pr), m = stencil i-2..i+2.
// bn: cell-centred normal B.  vl: face i+1/2 (left state), vr: face i-1/2.
// MODE 0: both projections formed, selected per lane (the kernels); in the
// WCNS-4 instances with WAVEB (FP64, and the FP32 LHLLD / MLAU y sweeps)
// the fast-only terms are skipped by a wavefront none of whose lanes takes
// the fast branch (the selects would discard them: same results; in the
// other FP32 sweeps the extra branch cost a wave per SIMD);
// MODE 1 / 2: only the fast/Alfven resp. Alfven/entropy projection is
// evaluated (used by the flop model to count one branch).  fbc, when given,
// receives the positivity-fallback flags (1: left state, 2: right state).
template <int ODR, typename T, int MODE = 0, bool INRANGE = true, bool WAVEB = std::is_same<T, double>::value>
QAS_HD void lrstate(const T (&s)[7][5], T bn, T gam, T (&vl)[7], T (&vr)[7], int* fbc = nullptr)
{
  const T ro0 = s[0][2], bt0 = s[4][2], bu0 = s[5][2], pr0 = s[6][2];
  // exact-zero flags of mhd_eigen.c:575-587
  bool zf[5];
#pragma unroll
  for (int n = 1; n <= 5; n++)
    zf[n - 1] = (s[n][0] == T(0.0)) & (s[n][1] == T(0.0)) & (s[n][2] == T(0.0)) & (s[n][3] == T(0.0)) &
                (s[n][4] == T(0.0));
  // branch of mhd_func.c:143 (strict >)
  const bool fast = MODE == 1 ? true : (MODE == 2 ? false : (s[0][2] * (s[1][2] * s[1][2] + s[2][2] * s[2][2] + s[3][2] * s[3][2]) > gam * s[6][2]));
  // whether the fast-only terms are formed: in WCNS-4 MODE 0 with WAVEB on
  // the device, when any lane of the wavefront takes the fast branch (a
  // uniform value)
#if defined(__HIP_DEVICE_COMPILE__)
  constexpr bool WAVE_BRANCH = MODE == 0 && ODR == 4 && WAVEB;
  const bool anyf = WAVE_BRANCH ? (__any(fast) != 0) : MODE != 2;
#else
  const bool anyf = MODE != 2;
#endif
  // shared eigen quantities (mhd_eigen_mi / mhd_eigen_ae)
  const T iro = T(1.0) / ro0;
  const T cs2 = gam * pr0 * iro;
  const T ca2 = bn * bn * iro;
  const T ct2 = (bt0 * bt0 + bu0 * bu0) * iro;
  const T cb2 = cs2 + ca2 + ct2;
  const T cf2 = T(0.5) * (cb2 + dsqrt(dabs(cb2 * cb2 - T(4.0) * cs2 * ca2)));
  T rcf = T(0.0), icf2h = T(0.0), ircfh = T(0.0);
  if (MODE != 2 && anyf) {
    const T cf = dsqrt(cf2);
    rcf = ro0 * cf;
  }
  const T sro = dsqrt(ro0);
  const T isroh = T(0.5) / sro;
  const T icf2 = T(1.0) / cf2;
  if (MODE != 2 && anyf) {
    icf2h = T(0.5) * icf2;
    ircfh = T(0.5) / rcf;
  }
  const T byh = T(0.5) * bt0;
  const T bzh = T(0.5) * bu0;

  T w[5], wl[7], wr[7];
  // k = 0: fast -> -rcf vn + bt0 bt + bu0 bu + pr ; Alfven/entropy -> vn
#pragma unroll
  for (int m = 0; m < 5; m++) {
    if (MODE == 1) w[m] = -rcf * s[1][m] + bt0 * s[4][m] + bu0 * s[5][m] + s[6][m];
    else w[m] = s[1][m];
  }
  if (MODE == 0 && anyf) {
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const T wf = -rcf * s[1][m] + bt0 * s[4][m] + bu0 * s[5][m] + s[6][m];
      w[m] = fast ? wf : w[m];
    }
  }
  interp5<ODR, T, INRANGE>(w, wl[0], wr[0]);
  // k = 1: cf2 ro - bt0 bt - bu0 bu - pr (both)
#pragma unroll
  for (int m = 0; m < 5; m++) w[m] = cf2 * s[0][m] - bt0 * s[4][m] - bu0 * s[5][m] - s[6][m];
  interp5<ODR, T, INRANGE>(w, wl[1], wr[1]);
  // k = 2: fast -> rcf vn + bt0 bt + bu0 bu + pr ; A/E -> bt0 bt + bu0 bu + pr
#pragma unroll
  for (int m = 0; m < 5; m++) {
    if (MODE == 1) w[m] = rcf * s[1][m] + bt0 * s[4][m] + bu0 * s[5][m] + s[6][m];
    else w[m] = bt0 * s[4][m] + bu0 * s[5][m] + s[6][m];
  }
  if (MODE == 0 && anyf) {
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const T wf = rcf * s[1][m] + bt0 * s[4][m] + bu0 * s[5][m] + s[6][m];
      w[m] = fast ? wf : w[m];
    }
  }
  interp5<ODR, T, INRANGE>(w, wl[2], wr[2]);
  // k = 3..6: Alfven pairs (both)
#pragma unroll
  for (int m = 0; m < 5; m++) w[m] = sro * s[2][m] + s[4][m];
  interp5<ODR, T, INRANGE>(w, wl[3], wr[3]);
#pragma unroll
  for (int m = 0; m < 5; m++) w[m] = sro * s[2][m] - s[4][m];
  interp5<ODR, T, INRANGE>(w, wl[4], wr[4]);
#pragma unroll
  for (int m = 0; m < 5; m++) w[m] = sro * s[3][m] + s[5][m];
  interp5<ODR, T, INRANGE>(w, wl[5], wr[5]);
#pragma unroll
  for (int m = 0; m < 5; m++) w[m] = sro * s[3][m] - s[5][m];
  interp5<ODR, T, INRANGE>(w, wl[6], wr[6]);

  // back projection
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const T* q = side ? wr : wl;
    T* u = side ? vr : vl;
    if (MODE == 1) {
      u[0] = icf2h * q[0] + icf2 * q[1] + icf2h * q[2];
      u[1] = -ircfh * q[0] + ircfh * q[2];
      u[6] = T(0.5) * q[0] + T(0.5) * q[2] - byh * q[3] + byh * q[4] - bzh * q[5] + bzh * q[6];
    } else if (MODE == 2) {
      u[0] = icf2 * q[1] + icf2 * q[2];
      u[1] = q[0];
      u[6] = q[2] - byh * q[3] + byh * q[4] - bzh * q[5] + bzh * q[6];
    } else {
      u[0] = icf2 * q[1] + icf2 * q[2];
      u[1] = q[0];
      u[6] = q[2] - byh * q[3] + byh * q[4] - bzh * q[5] + bzh * q[6];
      if (anyf) {
        const T rof = icf2h * q[0] + icf2 * q[1] + icf2h * q[2];
        const T vnf = -ircfh * q[0] + ircfh * q[2];
        const T prf = T(0.5) * q[0] + T(0.5) * q[2] - byh * q[3] + byh * q[4] - bzh * q[5] + bzh * q[6];
        u[0] = fast ? rof : u[0];
        u[1] = fast ? vnf : u[1];
        u[6] = fast ? prf : u[6];
      }
    }
    u[2] = isroh * q[3] + isroh * q[4];
    u[3] = isroh * q[5] + isroh * q[6];
    u[4] = T(0.5) * q[3] - T(0.5) * q[4];
    u[5] = T(0.5) * q[5] - T(0.5) * q[6];
  }
#pragma unroll
  for (int n = 0; n < 5; n++) {
    if (zf[n]) {
      vl[n + 1] = T(0.0);
      vr[n + 1] = T(0.0);
    }
  }
  // positivity fallback (mhd_func.c:150-160)
  const T eps = T(1e-12);
  const bool fbl = vl[0] <= eps || vl[6] <= eps;
  const bool fbr = vr[0] <= eps || vr[6] <= eps;
  if (fbl) {
    vl[0] = s[0][2];
    vl[1] = s[1][2];
    vl[6] = s[6][2];
  }
  if (fbr) {
    vr[0] = s[0][2];
    vr[1] = s[1][2];
    vr[6] = s[6][2];
  }
  if (fbc) *fbc = (fbl ? 1 : 0) | (fbr ? 2 : 0);
}

// face states of the B-flux b_t v_n - b_n v_t (mhd_func.c:163-188); the
// caller passes the 5-point products bt*vn, the vt stencil and the two
// staggered b_n values at i-1/2 (bnm) and i+1/2 (bnp).
template <int ODR, typename T>
QAS_HD void lr_fb(const T* p, const T* q, T bnm, T bnp, T& vl, T& vr)
{
  T fl0, fr0, fl1, fr1;
  lin5<ODR>(p, fl0, fr0);
  lin5<ODR>(q, fl1, fr1);
  vl = fl0 - bnp * fl1;
  vr = fr0 - bnm * fr1;
}

// ------------------------------------------------------------------ Riemann solvers
// q = (ro, vn, vt, vu, bt, bu, pr); F = (ro, mn, mt, mu, bt, bu, en)
#define QAS_FL_EPS 1e-8

template <typename T>
QAS_HD void flux_hlld_family(const bool lowdiss, const T* L, const T* R, T bnc, T gamma,
                                                 T dv0, T dv1, T* F)
{
  const T FL_EPS = T(QAS_FL_EPS);
  const T rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const T ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  const T gammam1i = T(1.0) / (gamma - T(1.0));
  const T bnc2 = bnc * bnc;
  const T sgn = (bnc > T(0)) ? T(1.0) : T(-1.0);
  const T roli = T(1.0) / rol;
  const T pml = T(0.5) * (btl * btl + bul * bul);
  const T ptl = prl + pml;
  const T vl2 = vnl * vnl + vtl * vtl + vul * vul;
  const T enl = gammam1i * prl + pml + T(0.5) * rol * vl2;
  const T vbl = vtl * btl + vul * bul;
  const T rori = T(1.0) / ror;
  const T pmr = T(0.5) * (btr * btr + bur * bur);
  const T ptr = prr + pmr;
  const T vr2 = vnr * vnr + vtr * vtr + vur * vur;
  const T enr = gammam1i * prr + pmr + T(0.5) * ror * vr2;
  const T vbr = vtr * btr + vur * bur;
  const T cl2 = gamma * prl * roli;
  const T cr2 = gamma * prr * rori;
  const T cal2 = bnc2 * roli;
  const T car2 = bnc2 * rori;
  const T cbl2 = cl2 + cal2 + T(2.0) * pml * roli;
  const T cbr2 = cr2 + car2 + T(2.0) * pmr * rori;
  const T cfl2 = T(0.5) * (cbl2 + dsqrt(dabs(cbl2 * cbl2 - T(4.0) * cl2 * cal2)));
  const T cfr2 = T(0.5) * (cbr2 + dsqrt(dabs(cbr2 * cbr2 - T(4.0) * cr2 * car2)));
  T cmax, ccl2 = T(0), ccr2 = T(0);
  if (lowdiss) {
    const T cbl2c = vl2 + cal2 + T(2.0) * pml * roli;
    const T cbr2c = vr2 + car2 + T(2.0) * pmr * rori;
    ccl2 = T(0.5) * (cbl2c + dsqrt(dabs(cbl2c * cbl2c - T(4.0) * vl2 * cal2)));
    ccr2 = T(0.5) * (cbr2c + dsqrt(dabs(cbr2c * cbr2c - T(4.0) * vr2 * car2)));
    cmax = dsqrt(tmax(cfl2, cfr2));
  } else {
    cmax = tmax(dsqrt(cfl2), dsqrt(cfr2));
  }
  const T sl = tmin(T(0.0), tmin(vnl, vnr) - cmax);
  const T sr = tmax(T(0.0), tmax(vnl, vnr) + cmax);
  const T slvl = sl - vnl;
  const T srvr = sr - vnr;
  const T rslvl = rol * slvl;
  const T rsrvr = ror * srvr;
  const T drsvi = T(1.0) / (rsrvr - rslvl);
  T vnc, ptc;
  if (lowdiss) {
    T theta = tmin(T(1.0), (cmax - tmin(dv0, T(0.0))) / (cmax - tmin(dv1, T(0.0))));
    theta *= theta;
    theta *= theta;
    vnc = (rsrvr * vnr - rslvl * vnl - theta * (ptr - ptl)) * drsvi;
    const T mcc = tmin(T(1.0), dsqrt(tmax(ccl2, ccr2)) / cmax);
    ptc = (rsrvr * ptl - rslvl * ptr + (mcc * (T(2.0) - mcc)) * rsrvr * rslvl * (vnr - vnl)) * drsvi;
  } else {
    vnc = (rsrvr * vnr - rslvl * vnl - ptr + ptl) * drsvi;
    ptc = (rsrvr * ptl - rslvl * ptr + rsrvr * rslvl * (vnr - vnl)) * drsvi;
  }
  const T slvc = sl - vnc;
  const T srvc = sr - vnc;
  const T ro2l = rslvl / slvc;
  const T ro2r = rsrvr / srvc;
  const T rhdl = rslvl * slvc - bnc2;
  const T rhdr = rsrvr * srvc - bnc2;
  T vt2l = vtl, vu2l = vul, bt2l = btl, bu2l = bul;
  T vt2r = vtr, vu2r = vur, bt2r = btr, bu2r = bur;
  if (dabs(rhdl) > FL_EPS) {
    const T rhdli = T(1.0) / rhdl;
    const T rhnvl = (vnl - vnc) * bnc;
    const T rhnbl = rslvl * slvl - bnc2;
    vt2l = vtl + rhnvl * rhdli * btl;
    vu2l = vul + rhnvl * rhdli * bul;
    bt2l = rhnbl * rhdli * btl;
    bu2l = rhnbl * rhdli * bul;
  }
  if (dabs(rhdr) > FL_EPS) {
    const T rhdri = T(1.0) / rhdr;
    const T rhnvr = (vnr - vnc) * bnc;
    const T rhnbr = rsrvr * srvr - bnc2;
    vt2r = vtr + rhnvr * rhdri * btr;
    vu2r = vur + rhnvr * rhdri * bur;
    bt2r = rhnbr * rhdri * btr;
    bu2r = rhnbr * rhdri * bur;
  }
  const T vb2l = vt2l * bt2l + vu2l * bu2l;
  const T vb2r = vt2r * bt2r + vu2r * bu2r;
  const T en2l = (slvl * enl - ptl * vnl + ptc * vnc + bnc * (vbl - vb2l)) / slvc;
  const T en2r = (srvr * enr - ptr * vnr + ptc * vnc + bnc * (vbr - vb2r)) / srvc;
  const T rro2l = dsqrt(ro2l);
  const T rro2r = dsqrt(ro2r);
  const T rro2i = T(1.0) / (rro2r + rro2l);
  const T vt3m = (rro2r * vt2r + rro2l * vt2l + (bt2r - bt2l) * sgn) * rro2i;
  const T vu3m = (rro2r * vu2r + rro2l * vu2l + (bu2r - bu2l) * sgn) * rro2i;
  const T bt3m = (rro2l * bt2r + rro2r * bt2l + rro2r * rro2l * (vt2r - vt2l) * sgn) * rro2i;
  const T bu3m = (rro2l * bu2r + rro2r * bu2l + rro2r * rro2l * (vu2r - vu2l) * sgn) * rro2i;
  const T vb3m = vt3m * bt3m + vu3m * bu3m;
  const T en3l = en2l - rro2l * (vb2l - vb3m) * sgn;
  const T en3r = en2r + rro2r * (vb2r - vb3m) * sgn;
  const T hl = (vnc > T(0)) ? T(1.0) : T(0.0);
  const T hr = T(1.0) - hl;
  const T h2l = (vnc - dabs(bnc) / rro2l > T(0)) ? T(1.0) : T(0.0);
  const T h3l = (T(1.0) - h2l) * hl;
  const T h2r = (vnc + dabs(bnc) / rro2r > T(0)) ? T(0.0) : T(1.0);
  const T h3r = (T(1.0) - h2r) * hr;
  const T rou = ro2l * hl + ro2r * hr;
  const T vtu = vt2l * h2l + vt3m * (h3l + h3r) + vt2r * h2r;
  const T vuu = vu2l * h2l + vu3m * (h3l + h3r) + vu2r * h2r;
  const T btu = bt2l * h2l + bt3m * (h3l + h3r) + bt2r * h2r;
  const T buu = bu2l * h2l + bu3m * (h3l + h3r) + bu2r * h2r;
  const T enu = en2l * h2l + en3l * h3l + en3r * h3r + en2r * h2r;
  F[0] = rou * vnc;
  F[1] = rou * vnc * vnc + ptc - bnc2 * T(0.5);
  F[2] = rou * vtu * vnc - bnc * btu;
  F[3] = rou * vuu * vnc - bnc * buu;
  F[4] = btu * vnc - bnc * vtu;
  F[5] = buu * vnc - bnc * vuu;
  F[6] = (enu + ptc) * vnc - bnc * (vtu * btu + vuu * buu);
}

template <typename T>
QAS_HD void flux_mlau(const T* L, const T* R, T bnc, T gamma, T dv0, T dv1, T* F)
{
  const T FL_EPS = T(QAS_FL_EPS);
  const T rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const T ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  const T gammam1i = T(1.0) / (gamma - T(1.0));
  const T bnc2 = bnc * bnc;
  const T sgn = (bnc > T(0)) ? T(1.0) : T(-1.0);
  const T bnca = sgn * bnc;
  const T roli = T(1.0) / rol;
  const T vl2 = vnl * vnl + vtl * vtl + vul * vul;
  const T pml = T(0.5) * (btl * btl + bul * bul);
  const T ptl = prl + pml;
  const T enl = gammam1i * prl + pml + T(0.5) * rol * vl2;
  const T hnl = (enl + ptl) * roli;
  const T rori = T(1.0) / ror;
  const T vr2 = vnr * vnr + vtr * vtr + vur * vur;
  const T pmr = T(0.5) * (btr * btr + bur * bur);
  const T ptr = prr + pmr;
  const T enr = gammam1i * prr + pmr + T(0.5) * ror * vr2;
  const T hnr = (enr + ptr) * rori;
  const T cl2 = gamma * prl * roli;
  const T cr2 = gamma * prr * rori;
  const T cal2 = bnc2 * roli;
  const T car2 = bnc2 * rori;
  T cbl2 = cl2 + cal2 + T(2.0) * pml * roli;
  T cbr2 = cr2 + car2 + T(2.0) * pmr * rori;
  const T cfl2 = T(0.5) * (cbl2 + dsqrt(dabs(cbl2 * cbl2 - T(4.0) * cl2 * cal2)));
  const T cfr2 = T(0.5) * (cbr2 + dsqrt(dabs(cbr2 * cbr2 - T(4.0) * cr2 * car2)));
  cbl2 = vl2 + cal2 + T(2.0) * pml * roli;
  cbr2 = vr2 + car2 + T(2.0) * pmr * rori;
  const T ccl2 = T(0.5) * (cbl2 + dsqrt(dabs(cbl2 * cbl2 - T(4.0) * vl2 * cal2)));
  const T ccr2 = T(0.5) * (cbr2 + dsqrt(dabs(cbr2 * cbr2 - T(4.0) * vr2 * car2)));
  const T cmax = dsqrt(tmax(cfl2, cfr2));
  const T cfn = cmax;
  const T cfni = T(1.0) / cfn;
  const T sl = tmin(T(0.0), tmin(vnl, vnr) - cmax);
  const T sr = tmax(T(0.0), tmax(vnl, vnr) + cmax);
  const T mal = vnl * cfni;
  const T mar = vnr * cfni;
  const T snl = (mal > T(0)) ? T(1.0) : T(-1.0);
  const T snr = (mar > T(0)) ? T(1.0) : T(-1.0);
  T theta = tmin(T(1.0), (cfn - tmin(dv0, T(0.0))) / (cfn - tmin(dv1, T(0.0))));
  theta *= theta;
  theta *= theta;
  const T alpha = T(0.1875);
  const T beta = T(0.125);
  const T malp1 = mal + T(1.0);
  const T marm1 = mar - T(1.0);
  const T mal2m1 = mal * mal - T(1.0);
  const T mar2m1 = mar * mar - T(1.0);
  T bfl, ffl, bfr, ffr;
  if (snl * mal > T(1.0)) {
    bfl = T(0.5) * (T(1.0) + snl);
    ffl = bfl * mal;
  } else {
    bfl = +T(0.25) * malp1 * malp1 * (T(2.0) - mal) + alpha * mal * mal2m1 * mal2m1;
    ffl = +T(0.25) * malp1 * malp1 + beta * mal2m1 * mal2m1;
  }
  if (snr * mar > T(1.0)) {
    bfr = T(0.5) * (T(1.0) - snr);
    ffr = bfr * mar;
  } else {
    bfr = +T(0.25) * marm1 * marm1 * (T(2.0) + mar) - alpha * mar * mar2m1 * mar2m1;
    ffr = T(-0.25) * marm1 * marm1 - beta * mar2m1 * mar2m1;
  }
  const T mnc = ffl + ffr;
  T mflux = mnc * cfn;
  mflux -= tmax(T(1.0) - dabs(mnc), T(0.0)) * theta * cfni * (ptr - ptl) / (rol + ror);
  const T dir = (mflux > T(0)) ? T(1.0) : T(-1.0);
  const T dirl = T(0.5) * (T(1.0) + dir);
  const T dirr = T(1.0) - dirl;
  mflux *= (dirl * rol + dirr * ror);
  const T mfl = mflux * dirl;
  const T mfr = mflux - mfl;
  T ptot = bfl * ptl + bfr * ptr;
  const T cfmod = dsqrt(tmax(ccl2, ccr2));
  ptot -= T(0.25) * bfl * bfr * (rol + ror) * cfmod * (vnr - vnl);
  ptot -= T(0.5) * (T(1.0) - cfmod * cfni) * (bfl + bfr - T(1.0)) * (ptl + ptr);
  const T slvl = sl - vnl;
  const T srvr = sr - vnr;
  const T snc = dirl * sl + dirr * sr;
  const T vndnm = mflux + dirl * rol * slvl + dirr * ror * srvr;
  const T vnc = ((dabs(vndnm) > FL_EPS) && (snc != T(0))) ? (mflux * snc / vndnm) : (dirl * vnl + dirr * vnr);
  const T slvc = sl - vnc;
  const T srvc = sr - vnc;
  const T slvci = T(1.0) / slvc;
  const T srvci = T(1.0) / srvc;
  const T rhdl = rol * slvl * slvc - bnc2;
  const T rhdr = ror * srvr * srvc - bnc2;
  T dvtl = T(0.0), dvul = T(0.0), dbtl = T(0.0), dbul = T(0.0);
  T dvtr = T(0.0), dvur = T(0.0), dbtr = T(0.0), dbur = T(0.0);
  if (dabs(rhdl) > FL_EPS) {
    const T q = (vnc - vnl) / rhdl;
    dvtl = -bnc * btl * q;
    dvul = -bnc * bul * q;
    dbtl = bnc2 * btl * q * slvci;
    dbul = bnc2 * bul * q * slvci;
  }
  if (dabs(rhdr) > FL_EPS) {
    const T q = (vnc - vnr) / rhdr;
    dvtr = -bnc * btr * q;
    dvur = -bnc * bur * q;
    dbtr = bnc2 * btr * q * srvci;
    dbur = bnc2 * bur * q * srvci;
  }
  const T cmpl = slvl * slvci;
  const T cmpr = srvr * srvci;
  const T ro2l = rol * cmpl;
  const T srol = dsqrt(ro2l);
  const T vt2l = dvtl + vtl;
  const T vu2l = dvul + vul;
  const T bt2l = dbtl + btl * cmpl;
  const T bu2l = dbul + bul * cmpl;
  const T ro2r = ror * cmpr;
  const T sror = dsqrt(ro2r);
  const T vt2r = dvtr + vtr;
  const T vu2r = dvur + vur;
  const T bt2r = dbtr + btr * cmpr;
  const T bu2r = dbur + bur * cmpr;
  const T sroi = T(1.0) / (srol + sror);
  const T srollri = srol * sroi;
  const T srorlri = sror * sroi;
  const T bnflb = sgn * tmin(tmax(T(0.0), srollri * (bnca + sror * vnc)), bnca);
  const T bnfrb = sgn * tmin(tmax(T(0.0), srorlri * (bnca - srol * vnc)), bnca);
  const T bnflv = sgn * tmin(tmax(T(0.0), srorlri * (bnca + mflux / sror)), bnca);
  const T bnfrv = sgn * tmin(tmax(T(0.0), srollri * (bnca - mflux / srol)), bnca);
  const T bnsrv = tmax(bnca - dir * vnc * (dirl * srol + dirr * sror), T(0.0));
  const T ccb = sroi * bnsrv;
  const T ccv = srol * sror * ccb;
  T bnvt = bnflb * vt2l + bnfrb * vt2r + ccb * (bt2r - bt2l);
  T bnvu = bnflb * vu2l + bnfrb * vu2r + ccb * (bu2r - bu2l);
  T bnbt = bnflv * bt2l + bnfrv * bt2r + ccv * (vt2r - vt2l);
  T bnbu = bnflv * bu2l + bnfrv * bu2r + ccv * (vu2r - vu2l);
  bnvt -= vnc * (dirl * dbtl + dirr * dbtr);
  bnvu -= vnc * (dirl * dbul + dirr * dbur);
  bnbt -= mflux * (dirl * dvtl + dirr * dvtr);
  bnbu -= mflux * (dirl * dvul + dirr * dvur);
  const T fmn = mfl * vnl + mfr * vnr + ptot - T(0.5) * bnc2;
  const T fmt = mfl * vtl + mfr * vtr - bnbt;
  const T fmu = mfl * vul + mfr * vur - bnbu;
  const T fbt = mfl * btl * roli + mfr * btr * rori - bnvt;
  const T fbu = mfl * bul * roli + mfr * bur * rori - bnvu;
  const T vb1 = dirl * (vtl * btl + vul * bul) + dirr * (vtr * btr + vur * bur);
  const T vb2 = dirl * (vt2l * bt2l + vu2l * bu2l) + dirr * (vt2r * bt2r + vu2r * bu2r);
  T rv2b2 = (mflux * vnc - bnc2);
  rv2b2 *= rv2b2;
  T vb3 = T(0.0);
  if (rv2b2 > FL_EPS * FL_EPS) {
    vb3 = (+(fmt * vnc + fbt * bnc) * (fmt * bnc + fbt * mflux) + (fmu * vnc + fbu * bnc) * (fmu * bnc + fbu * mflux)) /
          rv2b2;
  }
  const T bnvb = sgn * (bnca * (snc * vb2 - vnc * vb1) * (dirl * slvci + dirr * srvci) + bnsrv * (vb3 - vb2));
  F[0] = mflux;
  F[1] = fmn;
  F[2] = fmt;
  F[3] = fmu;
  F[4] = fbt;
  F[5] = fbu;
  F[6] = mfl * hnl + mfr * hnr - bnvb;
}

template <typename T>
QAS_HD void flux_roe(const T* L, const T* R, T bnc, T gamma, T* F)
{
  const T FL_EPS = T(QAS_FL_EPS);
  const T rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const T ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  const T gammam1 = gamma - T(1.0);
  const T gammam2 = gamma - T(2.0);
  const T gammam1i = T(1.0) / gammam1;
  const T roli = T(1.0) / rol;
  const T rori = T(1.0) / ror;
  const T rrol = dsqrt(rol);
  const T rror = dsqrt(ror);
  const T rroi = T(1.0) / (rrol + rror);
  const T robar = rrol * rror;
  const T robari = T(1.0) / robar;
  const T rrobar = dsqrt(robar);
  const T rrobai = dsqrt(robari);
  const T vnbar = (rrol * vnl + rror * vnr) * rroi;
  const T vtbar = (rrol * vtl + rror * vtr) * rroi;
  const T vubar = (rrol * vul + rror * vur) * rroi;
  const T bnbar = bnc;
  const T btbar = (rrol * btr + rror * btl) * rroi;
  const T bubar = (rrol * bur + rror * bul) * rroi;
  const T hl = (+gamma * gammam1i * prl * roli + T(0.5) * (vnl * vnl + vtl * vtl + vul * vul) +
                (bnbar * bnbar + btl * btl + bul * bul) * roli);
  const T hr = (+gamma * gammam1i * prr * rori + T(0.5) * (vnr * vnr + vtr * vtr + vur * vur) +
                (bnbar * bnbar + btr * btr + bur * bur) * rori);
  const T hbar = (rrol * hl + rror * hr) * rroi;
  const T btave = T(0.5) * (btl + btr);
  const T buave = T(0.5) * (bul + bur);
  const T sgn = (bnbar > T(0)) ? T(1.0) : T(-1.0);
  const T db2 = T(0.5) * gammam2 * gammam1i * ((btr - btl) * (btr - btl) + (bur - bul) * (bur - bul)) * rroi * rroi;
  const T v2b = T(0.5) * (vnbar * vnbar + vtbar * vtbar + vubar * vubar);
  const T bb2 = (bnbar * bnbar + btbar * btbar + bubar * bubar);
  const T c2 = gammam1 * (+hbar - v2b - db2 - bb2 * robari);
  const T a2 = (gammam1 * (+hbar - v2b - db2) - gammam2 * bb2 * robari);
  const T ca2 = bnbar * bnbar * robari;
  const T cb2 = (btbar * btbar + bubar * bubar) * robari;
  const T cf2 = T(0.5) * (a2 + dsqrt(cb2 * (a2 + c2 + ca2) + (c2 - ca2) * (c2 - ca2)));
  const T cs2 = c2 * ca2 / cf2;
  const T cf = dsqrt(cf2);
  const T cs = dsqrt(cs2);
  const T ca = dsqrt(ca2);
  const T c = dsqrt(c2);
  const T cfi = T(1.0) / cf;
  const T sgr1 = btbar * btbar + bubar * bubar - FL_EPS;
  const T sf11 = (sgr1 > T(0)) ? T(1.0) : T(0.0);
  const T sf21 = T(1.0) - sf11;
  const T beti = T(1.0) / dsqrt(btbar * btbar + bubar * bubar + sf21);
  const T bett = sf11 * btbar * beti + sf21 * dsqrt(T(0.5));
  const T betu = sf11 * bubar * beti + sf21 * dsqrt(T(0.5));
  const T sgr2 = (btbar * btbar + bubar * bubar) * robari + dabs(ca2 - c2) - FL_EPS;
  const T sf12 = (sgr2 > T(0)) ? T(1.0) : T(0.0);
  const T sf22 = T(1.0) - sf12;
  const T cfca = tmax(T(0.0), cf2 - ca2);
  const T cfcs = tmax(T(0.0), cf2 - cs2);
  const T cfa = tmax(T(0.0), cf2 - c2);
  const T alpi = T(1.0) / dsqrt(cfcs + sf22);
  const T alpf = sf12 * dsqrt(cfca) * alpi + sf22;
  const T alps = sf12 * dsqrt(cfa) * alpi;
  const T eeps = T(0.0);
  const T evpf = tmax(dabs(vnbar + cf), eeps);
  const T evmf = tmax(dabs(vnbar - cf), eeps);
  const T evps = tmax(dabs(vnbar + cs), eeps);
  const T evms = tmax(dabs(vnbar - cs), eeps);
  const T evpa = tmax(dabs(vnbar + ca), eeps);
  const T evma = tmax(dabs(vnbar - ca), eeps);
  const T ev00 = tmax(dabs(vnbar), eeps);
  const T dltro = ror - rol;
  const T dltmn = robar * (vnr - vnl);
  const T dltmt = robar * (vtr - vtl);
  const T dltmu = robar * (vur - vul);
  const T dltbt = btr - btl;
  const T dltbu = bur - bul;
  const T t1 = bett * dltbt + betu * dltbu;
  const T t2 = (prr - prl + (btave * dltbt + buave * dltbu) + gammam2 * (btbar * dltbt + bubar * dltbu)) * gammam1i;
  const T t3 = betu * dltbt - bett * dltbu;
  const T s1 = dltmn;
  const T s2 = bett * dltmt + betu * dltmu;
  const T s3 = betu * dltmt - bett * dltmu;
  const T p11 = alps * cf * rrobai;
  const T p12 = -alpf * c2 * cfi * rrobai;
  const T p21 = alpf * (cf2 - c2 * gammam2 * gammam1i);
  const T p22 = alps * (cs2 - c2 * gammam2 * gammam1i);
  const T q11 = alpf * cf;
  const T q12 = alps * cs;
  const T q21 = -alps * ca * sgn;
  const T q22 = alpf * c * sgn;
  const T detpi = T(1.0) / (p11 * p22 - p12 * p21);
  const T detqi = T(1.0) / (q11 * q22 - q12 * q21);
  const T wvpf = T(0.5) * ((p22 * t1 - p12 * t2) * detpi + (q22 * s1 - q12 * s2) * detqi);
  const T wvmf = T(0.5) * ((p22 * t1 - p12 * t2) * detpi - (q22 * s1 - q12 * s2) * detqi);
  const T wvps = T(0.5) * ((-p21 * t1 + p11 * t2) * detpi + (-q21 * s1 + q11 * s2) * detqi);
  const T wvms = T(0.5) * ((-p21 * t1 + p11 * t2) * detpi - (-q21 * s1 + q11 * s2) * detqi);
  const T wvpa = T(0.5) * (rrobar * t3 - sgn * s3);
  const T wvma = T(0.5) * (rrobar * t3 + sgn * s3);
  const T wv00 = dltro - alpf * (wvpf + wvmf) - alps * (wvps + wvms);
  const T frol = rol * vnl;
  const T fmnl = rol * vnl * vnl + prl + T(0.5) * (-bnbar * bnbar + btl * btl + bul * bul);
  const T fmtl = rol * vnl * vtl - bnbar * btl;
  const T fmul = rol * vnl * vul - bnbar * bul;
  const T fbtl = vnl * btl - vtl * bnbar;
  const T fbul = vnl * bul - vul * bnbar;
  const T fenl = rol * vnl * hl - bnbar * (bnbar * vnl + btl * vtl + bul * vul);
  const T fror = ror * vnr;
  const T fmnr = ror * vnr * vnr + prr + T(0.5) * (-bnbar * bnbar + btr * btr + bur * bur);
  const T fmtr = ror * vnr * vtr - bnbar * btr;
  const T fmur = ror * vnr * vur - bnbar * bur;
  const T fbtr = vnr * btr - vtr * bnbar;
  const T fbur = vnr * bur - vur * bnbar;
  const T fenr = ror * vnr * hr - bnbar * (bnbar * vnr + btr * vtr + bur * vur);
  const T bv = (bett * vtbar + betu * vubar);
  const T rvpfro = alpf;
  const T rvpfmn = alpf * (vnbar + cf);
  const T rvpfmt = alpf * vtbar - alps * bett * ca * sgn;
  const T rvpfmu = alpf * vubar - alps * betu * ca * sgn;
  const T rvpfbt = alps * bett * cf * rrobai;
  const T rvpfbu = alps * betu * cf * rrobai;
  const T rvpfen = alpf * (v2b + db2 + cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i) - alps * ca * bv * sgn;
  const T rvmfro = alpf;
  const T rvmfmn = alpf * (vnbar - cf);
  const T rvmfmt = alpf * vtbar + alps * bett * ca * sgn;
  const T rvmfmu = alpf * vubar + alps * betu * ca * sgn;
  const T rvmfbt = rvpfbt;
  const T rvmfbu = rvpfbu;
  const T rvmfen = alpf * (v2b + db2 - cf * vnbar + cf2 * gammam1i + (cf2 - c2) * gammam2 * gammam1i) + alps * ca * bv * sgn;
  const T rvpsro = alps;
  const T rvpsmn = alps * (vnbar + cs);
  const T rvpsmt = alps * vtbar + c * sgn * alpf * bett;
  const T rvpsmu = alps * vubar + c * sgn * alpf * betu;
  const T rvpsbt = -rrobai * c2 * alpf * bett * cfi;
  const T rvpsbu = -rrobai * c2 * alpf * betu * cfi;
  const T rvpsen = alps * (v2b + db2 + cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i) + alpf * c * bv * sgn;
  const T rvmsro = alps;
  const T rvmsmn = alps * (vnbar - cs);
  const T rvmsmt = alps * vtbar - c * sgn * alpf * bett;
  const T rvmsmu = alps * vubar - c * sgn * alpf * betu;
  const T rvmsbt = rvpsbt;
  const T rvmsbu = rvpsbu;
  const T rvmsen = alps * (v2b + db2 - cs * vnbar + cs2 * gammam1i + (cs2 - c2) * gammam2 * gammam1i) - alpf * c * bv * sgn;
  const T rvparo = T(0.0);
  const T rvpamn = T(0.0);
  const T rvpamt = -sgn * betu;
  const T rvpamu = sgn * bett;
  const T rvpabt = rrobai * betu;
  const T rvpabu = -rrobai * bett;
  const T rvpaen = -(betu * vtbar - bett * vubar) * sgn;
  const T rvmaro = T(0.0);
  const T rvmamn = T(0.0);
  const T rvmamt = -rvpamt;
  const T rvmamu = -rvpamu;
  const T rvmabt = rvpabt;
  const T rvmabu = rvpabu;
  const T rvmaen = -rvpaen;
  const T rv00ro = T(1.0);
  const T rv00mn = vnbar;
  const T rv00mt = vtbar;
  const T rv00mu = vubar;
  const T rv00bt = T(0.0);
  const T rv00bu = T(0.0);
  const T rv00en = v2b + db2;
  const T apf = evpf * wvpf, amf = evmf * wvmf, aps = evps * wvps, ams = evms * wvms, apa = evpa * wvpa,
          ama = evma * wvma, a00 = ev00 * wv00;
  F[0] = T(0.5) * (+frol + fror - apf * rvpfro - amf * rvmfro - aps * rvpsro - ams * rvmsro - apa * rvparo -
                   ama * rvmaro - a00 * rv00ro);
  F[1] = T(0.5) * (+fmnl + fmnr - apf * rvpfmn - amf * rvmfmn - aps * rvpsmn - ams * rvmsmn - apa * rvpamn -
                   ama * rvmamn - a00 * rv00mn);
  F[2] = T(0.5) * (fmtl + fmtr - apf * rvpfmt - amf * rvmfmt - aps * rvpsmt - ams * rvmsmt - apa * rvpamt -
                   ama * rvmamt - a00 * rv00mt);
  F[3] = T(0.5) * (fmul + fmur - apf * rvpfmu - amf * rvmfmu - aps * rvpsmu - ams * rvmsmu - apa * rvpamu -
                   ama * rvmamu - a00 * rv00mu);
  F[4] = T(0.5) * (fbtl + fbtr - apf * rvpfbt - amf * rvmfbt - aps * rvpsbt - ams * rvmsbt - apa * rvpabt -
                   ama * rvmabt - a00 * rv00bt);
  F[5] = T(0.5) * (fbul + fbur - apf * rvpfbu - amf * rvmfbu - aps * rvpsbu - ams * rvmsbu - apa * rvpabu -
                   ama * rvmabu - a00 * rv00bu);
  F[6] = T(0.5) * (fenl + fenr - apf * rvpfen - amf * rvmfen - aps * rvpsen - ams * rvmsen - apa * rvpaen -
                   ama * rvmaen - a00 * rv00en);
}

// riemann[RMN] (common/mhd_class.hpp:102-105)
template <int RMN, typename T>
QAS_HD void riemann(const T* L, const T* R, T bn, T gam, T dv0, T dv1, T* F)
{
  if constexpr (RMN == 0) flux_roe(L, R, bn, gam, F);
  else if constexpr (RMN == 1) flux_hlld_family(false, L, R, bn, gam, dv0, dv1, F);
  else if constexpr (RMN == 2) flux_hlld_family(true, L, R, bn, gam, dv0, dv1, F);
  else flux_mlau(L, R, bn, gam, dv0, dv1, F);
}

// mhd_cuct_weight from the corner's four cell means and two face means
template <typename T>
QAS_HD T cuct_weight(T ro0, T ro1, T ro2, T ro3, T va0, T va1, T va2, T va3, T vb0, T vb1,
                                         T vb2, T vb3, T ba1, T ba3, T bb2, T bb3)
{
  const T eps = T(1e-12);
  T roc = T(0.25) * (ro0 + ro1 + ro2 + ro3);
  roc = T(1.0) / dsqrt(roc);
  T vac = T(0.25) * (va0 + va1 + va2 + va3);
  T vbc = T(0.25) * (vb0 + vb1 + vb2 + vb3);
  const T bac = T(0.5) * (ba1 + ba3);
  const T bbc = T(0.5) * (bb2 + bb3);
  vac = dabs(vac) + dabs(bac * roc);
  vbc = dabs(vbc) + dabs(bbc * roc);
  return (vac + T(0.5) * eps) / (vac + vbc + eps);
}

// rk_updt (common/common_func.h:28-32)
QAS_HD double rk_upd(double f1, double f0, double df, double a, double b)
{
  return a * f0 + b * (f1 + df);
}

}  // namespace qas
