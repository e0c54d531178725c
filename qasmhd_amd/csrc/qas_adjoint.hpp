// qas_adjoint.hpp -- reverse-mode (adjoint) versions of the per-point
// functions of qas_device.hpp, written by hand.
//
// Each *_adj function recomputes the forward intermediates of its primal
// (same formulas as qas_device.hpp) and then propagates output adjoints to
// input adjoints, accumulating (+=) into the caller's input-adjoint arrays.
// Non-smooth operations are differentiated at the branch the forward pass
// takes: the WCNS nonlinear weights are differentiated exactly; max/min and
// the HLLD region flags route the derivative to the selected operand / fix
// the region; the positivity fallback and the exact-zero flags of the
// characteristic reconstruction keep the branch the forward takes; |x| has
// derivative sign(x).
#pragma once
#include "qas_device.hpp"

namespace qas {

template <typename T> QAS_HD T sgnd(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

// ---------------------------------------------------------------- WCNS-4 / lin4
// gf[0..4] += d(fl, fr)/df ^T (gl, gr); fl, fr themselves (the forward
// face values) go to *flo, *fro when given
template <typename T>
QAS_HD void wcns4_adj(const T* f, T gl, T gr, T* gf, T* flo = nullptr, T* fro = nullptr)
{
  const T eps = T(1e-40);
  const T c13 = T(13. / 12.);
  const T a0 = f[0] - T(4.0) * f[1] + T(3.0) * f[2];
  const T b0 = f[0] - T(2.0) * f[1] + f[2];
  const T s0 = c13 * b0 * b0 + T(0.25) * a0 * a0;
  const T a1 = f[1] - f[3];
  const T b1 = f[1] - T(2.0) * f[2] + f[3];
  const T s1 = c13 * b1 * b1 + T(0.25) * a1 * a1;
  const T a2 = f[4] - T(4.0) * f[3] + T(3.0) * f[2];
  const T b2 = f[4] - T(2.0) * f[3] + f[2];
  const T s2 = c13 * b2 * b2 + T(0.25) * a2 * a2;
  const T S0 = T(1.0) / (s0 * s0 + eps), S1 = T(1.0) / (s1 * s1 + eps), S2 = T(1.0) / (s2 * s2 + eps);
  const T qa = T(3.0) * f[0] - T(10.0) * f[1] + T(15.0) * f[2];
  const T qb = -f[1] + T(6.0) * f[2] + T(3.0) * f[3];
  const T qc = T(3.0) * f[2] + T(6.0) * f[3] - f[4];
  const T ra = T(3.0) * f[4] - T(10.0) * f[3] + T(15.0) * f[2];
  const T rb = -f[3] + T(6.0) * f[2] + T(3.0) * f[1];
  const T rc = T(3.0) * f[2] + T(6.0) * f[1] - f[0];
  // left weights
  const T u0 = T(0.075) * S0, u1 = T(0.650) * S1, u2 = T(0.275) * S2;
  const T D = T(1.0) / (u0 + u1 + u2);
  const T w0 = u0 * D, w1 = u1 * D, w2 = T(1.0) - w0 - w1;
  // right weights
  const T v0 = T(0.275) * S0, v1 = T(0.650) * S1, v2 = T(0.075) * S2;
  const T E = T(1.0) / (v0 + v1 + v2);
  const T z0 = v0 * E, z1 = v1 * E, z2 = T(1.0) - z0 - z1;
  if (flo) *flo = T(0.125) * (w0 * qa + w1 * qb + w2 * qc);
  if (fro) *fro = T(0.125) * (z2 * ra + z1 * rb + z0 * rc);
  // reverse: fl = 0.125 (w0 qa + w1 qb + w2 qc)
  T gw0 = T(0.125) * qa * gl, gw1 = T(0.125) * qb * gl, gw2 = T(0.125) * qc * gl;
  const T gqa = T(0.125) * w0 * gl, gqb = T(0.125) * w1 * gl, gqc = T(0.125) * w2 * gl;
  gw0 -= gw2;
  gw1 -= gw2;
  T gu0 = D * gw0, gu1 = D * gw1;
  const T gD = u0 * gw0 + u1 * gw1;
  const T gsu = -D * D * gD;
  gu0 += gsu;
  gu1 += gsu;
  const T gu2 = gsu;
  // fr = 0.125 (z2 ra + z1 rb + z0 rc)
  T gz2 = T(0.125) * ra * gr, gz1 = T(0.125) * rb * gr, gz0 = T(0.125) * rc * gr;
  const T gra = T(0.125) * z2 * gr, grb = T(0.125) * z1 * gr, grc = T(0.125) * z0 * gr;
  gz0 -= gz2;
  gz1 -= gz2;
  T gv0 = E * gz0, gv1 = E * gz1;
  const T gE = v0 * gz0 + v1 * gz1;
  const T gsv = -E * E * gE;
  gv0 += gsv;
  gv1 += gsv;
  const T gv2 = gsv;
  const T gS0 = T(0.075) * gu0 + T(0.275) * gv0;
  const T gS1 = T(0.650) * gu1 + T(0.650) * gv1;
  const T gS2 = T(0.275) * gu2 + T(0.075) * gv2;
  const T gs0 = -S0 * S0 * T(2.0) * s0 * gS0;
  const T gs1 = -S1 * S1 * T(2.0) * s1 * gS1;
  const T gs2 = -S2 * S2 * T(2.0) * s2 * gS2;
  const T ga0 = T(0.5) * a0 * gs0, gb0 = T(2.0) * c13 * b0 * gs0;
  const T ga1 = T(0.5) * a1 * gs1, gb1 = T(2.0) * c13 * b1 * gs1;
  const T ga2 = T(0.5) * a2 * gs2, gb2 = T(2.0) * c13 * b2 * gs2;
  gf[0] += ga0 + gb0 + T(3.0) * gqa - grc;
  gf[1] += T(-4.0) * ga0 - T(2.0) * gb0 + ga1 + gb1 - T(10.0) * gqa - gqb + T(3.0) * grb + T(6.0) * grc;
  gf[2] += T(3.0) * ga0 + gb0 - T(2.0) * gb1 + T(3.0) * ga2 + gb2 + T(15.0) * gqa + T(6.0) * gqb + T(3.0) * gqc +
           T(15.0) * gra + T(6.0) * grb + T(3.0) * grc;
  gf[3] += -ga1 + gb1 - T(4.0) * ga2 - T(2.0) * gb2 + T(3.0) * gqb + T(6.0) * gqc - T(10.0) * gra - grb;
  gf[4] += ga2 + gb2 - gqc + T(3.0) * gra;
}

// linear 4th-order face interpolation (cal_flr_4th), adjoint
template <typename T>
QAS_HD void lin4_adj(T gl, T gr, T* gf)
{
  const T c = T(0.003125);
  gf[0] += c * (T(9.0) * gl - T(11.0) * gr);
  gf[1] += c * (T(-56.0) * gl + T(144.0) * gr);
  gf[2] += c * (T(234.0) * gl + T(234.0) * gr);
  gf[3] += c * (T(144.0) * gl - T(56.0) * gr);
  gf[4] += c * (T(-11.0) * gl + T(9.0) * gr);
}

// ---------------------------------------------------------------- lrstate
// Adjoint of lrstate<4> (characteristic WCNS-4 reconstruction of one cell,
// mhd_func.c:108-161 with mhd_eigen.c's projections).  The stencil comes
// through a loader, ld(j, m) = s[j][m] (ld.fence() marks where values read
// so far need not be kept), and its adjoint leaves through an
// accumulator, acc(j, m, v) adds v to the adjoint of s[j][m]; *gbn gets the
// adjoint of the cell-centred bn.  The seven characteristic fields are
// reversed one group at a time (fields 0-2, 3-4, 5-6), so no 7 x 5 array
// of characteristic values or of their adjoints is ever held at once; the
// face values of a group's fields (which the adjoint of the back projection
// needs) come out of the WCNS adjoints of that group.  FB_IN: fbc holds
// the forward's positivity-fallback flags (lrstate's fbc), kept by the
// caller; otherwise they are recomputed by a forward reconstruction of the
// seven fields.
template <bool FB_IN = false, typename T, typename Ld, typename Acc>
QAS_HD void lrstate_adj_g(const Ld& ld, T bn, T gam, const T (&gvl_in)[7], const T (&gvr_in)[7], Acc& acc,
                          T& gbn, int fbc = 0)
{
  T gvl[7], gvr[7];
  for (int k = 0; k < 7; k++) {
    gvl[k] = gvl_in[k];
    gvr[k] = gvr_in[k];
  }
  // forward recompute (the projection branch of mhd_func.c:143)
  const T ro0 = ld(0, 2), bt0 = ld(4, 2), bu0 = ld(5, 2), pr0 = ld(6, 2);
  const bool fast = (ld(0, 2) * (ld(1, 2) * ld(1, 2) + ld(2, 2) * ld(2, 2) + ld(3, 2) * ld(3, 2)) > gam * ld(6, 2));
  const T eps = T(1e-12);
  // exact-zero flags
  bool zf[5];
  for (int n = 1; n <= 5; n++)
    zf[n - 1] = (ld(n, 0) == T(0.0)) && (ld(n, 1) == T(0.0)) && (ld(n, 2) == T(0.0)) && (ld(n, 3) == T(0.0)) &&
                (ld(n, 4) == T(0.0));
  // eigen quantities
  const T iro = T(1.0) / ro0;
  const T cs2 = gam * pr0 * iro;
  const T ca2 = bn * bn * iro;
  const T ct2 = (bt0 * bt0 + bu0 * bu0) * iro;
  const T cb2 = cs2 + ca2 + ct2;
  const T disc = cb2 * cb2 - T(4.0) * cs2 * ca2;
  const T sq = dsqrt(dabs(disc));
  const T cf2 = T(0.5) * (cb2 + sq);
  const T cf = dsqrt(cf2);
  const T rcf = ro0 * cf;
  const T sro = dsqrt(ro0);
  const T isroh = T(0.5) / sro;
  const T icf2 = T(1.0) / cf2;
  const T icf2h = T(0.5) * icf2;
  const T ircfh = T(0.5) / rcf;
  const T byh = T(0.5) * bt0;
  const T bzh = T(0.5) * bu0;
  // characteristic value of field k at stencil point m (w = L s)
  auto proj = [&](int k, int m) -> T {
    switch (k) {
      case 0: return fast ? (-rcf * ld(1, m) + bt0 * ld(4, m) + bu0 * ld(5, m) + ld(6, m)) : ld(1, m);
      case 1: return cf2 * ld(0, m) - bt0 * ld(4, m) - bu0 * ld(5, m) - ld(6, m);
      case 2:
        return fast ? (rcf * ld(1, m) + bt0 * ld(4, m) + bu0 * ld(5, m) + ld(6, m))
                    : (bt0 * ld(4, m) + bu0 * ld(5, m) + ld(6, m));
      case 3: return sro * ld(2, m) + ld(4, m);
      case 4: return sro * ld(2, m) - ld(4, m);
      case 5: return sro * ld(3, m) + ld(5, m);
      default: return sro * ld(3, m) - ld(5, m);
    }
  };
  bool fb[2];
  if constexpr (!FB_IN) {
    // face values of the characteristic fields (pre-fallback), back
    // projection only as far as the positivity test needs it
    T wl[7], wr[7];
#pragma unroll
    for (int k = 0; k < 7; k++) {
      T wk[5];
#pragma unroll
      for (int m = 0; m < 5; m++) wk[m] = proj(k, m);
      interp5<4, T, false>(wk, wl[k], wr[k]);
    }
    ld.fence();
    for (int side = 0; side < 2; side++) {
      const T* q = side ? wr : wl;
      const T u0 = fast ? (icf2h * q[0] + icf2 * q[1] + icf2h * q[2]) : (icf2 * q[1] + icf2 * q[2]);
      const T u6 = fast ? (T(0.5) * q[0] + T(0.5) * q[2] - byh * q[3] + byh * q[4] - bzh * q[5] + bzh * q[6])
                        : (q[2] - byh * q[3] + byh * q[4] - bzh * q[5] + bzh * q[6]);
      fb[side] = (u0 <= eps || u6 <= eps);
    }
  } else {
    fb[0] = (fbc & 1) != 0;
    fb[1] = (fbc & 2) != 0;
  }
  // reverse: positivity fallback (first order for ro, vn, p)
  T gc0 = T(0), gc1 = T(0), gc6 = T(0);  // adjoints of the centre values
  if (fb[0]) {
    gc0 += gvl[0];
    gc1 += gvl[1];
    gc6 += gvl[6];
    gvl[0] = gvl[1] = gvl[6] = T(0.0);
  }
  if (fb[1]) {
    gc0 += gvr[0];
    gc1 += gvr[1];
    gc6 += gvr[6];
    gvr[0] = gvr[1] = gvr[6] = T(0.0);
  }
  for (int n = 0; n < 5; n++)
    if (zf[n]) gvl[n + 1] = gvr[n + 1] = T(0.0);
  // reverse: back projection (both sides), the part that does not need the
  // face values: the adjoints of the characteristic face values
  T gwl[7] = {}, gwr[7] = {};
  for (int side = 0; side < 2; side++) {
    const T* g = side ? gvr : gvl;
    T* gq = side ? gwr : gwl;
    if (fast) {
      gq[0] += icf2h * g[0] - ircfh * g[1] + T(0.5) * g[6];
      gq[1] += icf2 * g[0];
      gq[2] += icf2h * g[0] + ircfh * g[1] + T(0.5) * g[6];
    } else {
      gq[0] += g[1];
      gq[1] += icf2 * g[0];
      gq[2] += icf2 * g[0] + g[6];
    }
    gq[3] += isroh * g[2] + T(0.5) * g[4] - byh * g[6];
    gq[4] += isroh * g[2] - T(0.5) * g[4] + byh * g[6];
    gq[5] += isroh * g[3] + T(0.5) * g[5] - bzh * g[6];
    gq[6] += isroh * g[3] - T(0.5) * g[5] + bzh * g[6];
  }
  // reverse: WCNS of field k, then the projection w = L s, by groups; the
  // rest of the back-projection adjoint (that of its coefficients) per group
  T gicf2 = T(0), gicf2h = T(0), gircfh = T(0), gisroh = T(0), gbyh = T(0), gbzh = T(0);
  T grcf = T(0), gcf2 = T(0), gsro = T(0), gbt0 = T(0), gbu0 = T(0);
  auto field_adj = [&](int k, T (&gk)[5], T& ql, T& qr) {
    T wk[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      wk[m] = proj(k, m);
      gk[m] = T(0);
    }
    wcns4_adj(wk, gwl[k], gwr[k], gk, &ql, &qr);
  };
  {  // fields 0-2 (fast pair / slow and entropy)
    ld.fence();
    T g0[5], g1[5], g2[5];
    T q0l, q0r, q1l, q1r, q2l, q2r;
    field_adj(0, g0, q0l, q0r);
    field_adj(1, g1, q1l, q1r);
    field_adj(2, g2, q2l, q2r);
    if (fast) {
      gicf2h += (q0l + q2l) * gvl[0] + (q0r + q2r) * gvr[0];
      gicf2 += q1l * gvl[0] + q1r * gvr[0];
      gircfh += (q2l - q0l) * gvl[1] + (q2r - q0r) * gvr[1];
    } else {
      gicf2 += (q1l + q2l) * gvl[0] + (q1r + q2r) * gvr[0];
    }
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const T s0 = ld(0, m), s1 = ld(1, m), s4 = ld(4, m), s5 = ld(5, m);
      if (fast) {
        const T t = g0[m] + g2[m];
        acc(1, m, -rcf * g0[m] + rcf * g2[m]);
        acc(4, m, bt0 * t - bt0 * g1[m]);
        acc(5, m, bu0 * t - bu0 * g1[m]);
        acc(6, m, t - g1[m]);
        grcf += s1 * (g2[m] - g0[m]);
        gbt0 += s4 * (t - g1[m]);
        gbu0 += s5 * (t - g1[m]);
      } else {
        acc(1, m, g0[m]);
        acc(4, m, bt0 * g2[m] - bt0 * g1[m]);
        acc(5, m, bu0 * g2[m] - bu0 * g1[m]);
        acc(6, m, g2[m] - g1[m]);
        gbt0 += s4 * (g2[m] - g1[m]);
        gbu0 += s5 * (g2[m] - g1[m]);
      }
      acc(0, m, cf2 * g1[m]);
      gcf2 += s0 * g1[m];
    }
  }
  {  // fields 3-4 (Alfven, t component)
    ld.fence();
    T g3[5], g4[5];
    T q3l, q3r, q4l, q4r;
    field_adj(3, g3, q3l, q3r);
    field_adj(4, g4, q4l, q4r);
    gisroh += (q3l + q4l) * gvl[2] + (q3r + q4r) * gvr[2];
    gbyh += (q4l - q3l) * gvl[6] + (q4r - q3r) * gvr[6];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      acc(2, m, sro * (g3[m] + g4[m]));
      acc(4, m, g3[m] - g4[m]);
      gsro += ld(2, m) * (g3[m] + g4[m]);
    }
  }
  {  // fields 5-6 (Alfven, u component)
    ld.fence();
    T g5[5], g6[5];
    T q5l, q5r, q6l, q6r;
    field_adj(5, g5, q5l, q5r);
    field_adj(6, g6, q6l, q6r);
    gisroh += (q5l + q6l) * gvl[3] + (q5r + q6r) * gvr[3];
    gbzh += (q6l - q5l) * gvl[6] + (q6r - q5r) * gvr[6];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      acc(3, m, sro * (g5[m] + g6[m]));
      acc(5, m, g5[m] - g6[m]);
      gsro += ld(3, m) * (g5[m] + g6[m]);
    }
  }
  ld.fence();
  // reverse: eigen quantities
  gbt0 += T(0.5) * gbyh;
  gbu0 += T(0.5) * gbzh;
  // ircfh = 0.5 / rcf ; icf2h = 0.5 icf2 ; icf2 = 1/cf2 ; isroh = 0.5/sro
  grcf += -ircfh / rcf * gircfh;
  gicf2 += T(0.5) * gicf2h;
  gcf2 += -icf2 * icf2 * gicf2;
  gsro += -isroh / sro * gisroh;
  // rcf = ro0 cf ; cf = sqrt(cf2) ; sro = sqrt(ro0)
  T gro0 = cf * grcf;
  const T gcf = ro0 * grcf;
  gcf2 += (cf > T(0) ? T(0.5) / cf * gcf : T(0));
  gro0 += (sro > T(0) ? T(0.5) / sro * gsro : T(0));
  // cf2 = 0.5 (cb2 + sq) ; sq = sqrt(|disc|)
  T gcb2 = T(0.5) * gcf2;
  const T gsq = T(0.5) * gcf2;
  const T gdisc = (sq > T(0) ? T(0.5) / sq * gsq : T(0)) * sgnd(disc);
  // disc = cb2^2 - 4 cs2 ca2
  gcb2 += T(2.0) * cb2 * gdisc;
  T gcs2 = gcb2 - T(4.0) * ca2 * gdisc;
  T gca2 = gcb2 - T(4.0) * cs2 * gdisc;
  const T gct2 = gcb2;
  // cs2 = gam pr0 iro ; ca2 = bn^2 iro ; ct2 = (bt0^2 + bu0^2) iro
  T giro = gam * pr0 * gcs2 + bn * bn * gca2 + (bt0 * bt0 + bu0 * bu0) * gct2;
  T gpr0 = gam * iro * gcs2;
  gbn += T(2.0) * bn * iro * gca2;
  gbt0 += T(2.0) * bt0 * iro * gct2;
  gbu0 += T(2.0) * bu0 * iro * gct2;
  gro0 += -iro * iro * giro;
  acc(0, 2, gro0 + gc0);
  acc(1, 2, gc1);
  acc(4, 2, gbt0);
  acc(5, 2, gbu0);
  acc(6, 2, gpr0 + gc6);
}

// the same with the stencil and its adjoint as 7 x 5 arrays (gs += ...)
template <typename T>
struct StencilArr {
  const T (&s)[7][5];
  QAS_HD T operator()(int j, int m) const { return s[j][m]; }
  QAS_HD void fence() const {}
};
template <typename T>
struct StencilAcc {
  T (&gs)[7][5];
  QAS_HD void operator()(int j, int m, T v) { gs[j][m] += v; }
};
template <typename T>
QAS_HD void lrstate_adj(const T (&s)[7][5], T bn, T gam, const T (&gvl)[7], const T (&gvr)[7], T (&gs)[7][5],
                        T& gbn)
{
  StencilArr<T> ld{s};
  StencilAcc<T> acc{gs};
  lrstate_adj_g(ld, bn, gam, gvl, gvr, acc, gbn);
}

// ---------------------------------------------------------------- lr_fb
// vl = lin_l(p) - bnp lin_l(q) ; vr = lin_r(p) - bnm lin_r(q)  (ODR 4)
template <typename T>
QAS_HD void lr_fb_adj(const T* p, const T* q, T bnm, T bnp, T gvl, T gvr, T* gp, T* gq, T& gbnm, T& gbnp)
{
  (void)p;  // the face states are linear in p
  T fl1, fr1;
  lin5<4>(q, fl1, fr1);
  lin4_adj(gvl, gvr, gp);
  lin4_adj(-bnp * gvl, -bnm * gvr, gq);
  gbnp += -fl1 * gvl;
  gbnm += -fr1 * gvr;
}

// ---------------------------------------------------------------- CUCT weight
// g[0..3] ro, g[4..7] va, g[8..11] vb, g[12] ba1, g[13] ba3, g[14] bb2, g[15] bb3
template <typename T>
QAS_HD void cuct_weight_adj(const T* x, T gw, T* g)
{
  const T eps = T(1e-12);
  const T rs = T(0.25) * (x[0] + x[1] + x[2] + x[3]);
  const T srs = dsqrt(rs);
  const T roc = T(1.0) / srs;
  const T vac0 = T(0.25) * (x[4] + x[5] + x[6] + x[7]);
  const T vbc0 = T(0.25) * (x[8] + x[9] + x[10] + x[11]);
  const T bac = T(0.5) * (x[12] + x[13]);
  const T bbc = T(0.5) * (x[14] + x[15]);
  const T va = dabs(vac0) + dabs(bac * roc);
  const T vb = dabs(vbc0) + dabs(bbc * roc);
  const T num = va + T(0.5) * eps, den = va + vb + eps;
  // w = num / den
  const T gnum = gw / den;
  const T gden = -num / (den * den) * gw;
  const T gva = gnum + gden;
  const T gvb = gden;
  const T gvac0 = sgnd(vac0) * gva, gvbc0 = sgnd(vbc0) * gvb;
  const T ga = sgnd(bac * roc) * gva, gb = sgnd(bbc * roc) * gvb;
  const T gbac = roc * ga, gbbc = roc * gb;
  const T groc = bac * ga + bbc * gb;
  const T grs = -roc / srs * T(0.5) / srs * groc;  // d(1/sqrt(rs)) = -0.5 rs^-3/2
  for (int k = 0; k < 4; k++) {
    g[k] += T(0.25) * grs;
    g[4 + k] += T(0.25) * gvac0;
    g[8 + k] += T(0.25) * gvbc0;
  }
  g[12] += T(0.5) * gbac;
  g[13] += T(0.5) * gbac;
  g[14] += T(0.5) * gbbc;
  g[15] += T(0.5) * gbbc;
}

// ---------------------------------------------------------------- HLLD
// gL, gR (7 each) and gbn += adjoint of calc_flux_hlld; gm1i = 1/(gam-1)
// LOW: the low-dissipation LHLLD of the same family (calc_flux_lhlld,
// mhd_flux.c:563-698): the pressure jump in v* scaled by the shock-sensor
// weight theta(dv0, dv1), the velocity jump in p* by mcc (2 - mcc); gdv0,
// gdv1 get the adjoints of the sensor inputs.
template <bool LOW, typename T>
QAS_HD void hlld_adj_g(const T* L, const T* R, T bn, T gam, T dv0, T dv1, const T* gF, T* gL, T* gR, T& gbn,
                       T& gdv0, T& gdv1)
{
  const T FL_EPS = T(QAS_FL_EPS);
  const T rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const T ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  // ---- forward
  const T gm1i = T(1.0) / (gam - T(1.0));
  const T bnc2 = bn * bn;
  const T sgn = (bn > T(0)) ? T(1.0) : T(-1.0);
  const T roli = T(1.0) / rol, rori = T(1.0) / ror;
  const T pml = T(0.5) * (btl * btl + bul * bul), pmr = T(0.5) * (btr * btr + bur * bur);
  const T ptl = prl + pml, ptr = prr + pmr;
  const T vl2 = vnl * vnl + vtl * vtl + vul * vul, vr2 = vnr * vnr + vtr * vtr + vur * vur;
  const T enl = gm1i * prl + pml + T(0.5) * rol * vl2, enr = gm1i * prr + pmr + T(0.5) * ror * vr2;
  const T vbl = vtl * btl + vul * bul, vbr = vtr * btr + vur * bur;
  const T cl2 = gam * prl * roli, cr2 = gam * prr * rori;
  const T cal2 = bnc2 * roli, car2 = bnc2 * rori;
  const T cbl2 = cl2 + cal2 + T(2.0) * pml * roli, cbr2 = cr2 + car2 + T(2.0) * pmr * rori;
  const T dl = cbl2 * cbl2 - T(4.0) * cl2 * cal2, dr = cbr2 * cbr2 - T(4.0) * cr2 * car2;
  const T sql = dsqrt(dabs(dl)), sqr = dsqrt(dabs(dr));
  const T cfl2 = T(0.5) * (cbl2 + sql), cfr2 = T(0.5) * (cbr2 + sqr);
  const T cfl = dsqrt(cfl2), cfr = dsqrt(cfr2);
  // HLLD: max(sqrt(cfl2), sqrt(cfr2)); LHLLD: sqrt(max(cfl2, cfr2)) and the
  // second speed pair cc*2 (as in MLAU)
  const bool selL = LOW ? (cfl2 >= cfr2) : (cfl >= cfr);
  const T cmax = LOW ? dsqrt(selL ? cfl2 : cfr2) : (selL ? cfl : cfr);
  T cbl2c = T(0), cbr2c = T(0), elc = T(0), erc = T(0), selc = T(0), serc = T(0), ccl2 = T(0), ccr2 = T(0);
  if (LOW) {
    cbl2c = vl2 + cal2 + T(2.0) * pml * roli;
    cbr2c = vr2 + car2 + T(2.0) * pmr * rori;
    elc = cbl2c * cbl2c - T(4.0) * vl2 * cal2;
    erc = cbr2c * cbr2c - T(4.0) * vr2 * car2;
    selc = dsqrt(dabs(elc));
    serc = dsqrt(dabs(erc));
    ccl2 = T(0.5) * (cbl2c + selc);
    ccr2 = T(0.5) * (cbr2c + serc);
  }
  const bool nr_min = vnl >= vnr;  // min(vnl,vnr) = vnr, max = vnl
  const T mnv = nr_min ? vnr : vnl, mxv = nr_min ? vnl : vnr;
  const T xs = mnv - cmax, ys = mxv + cmax;
  const bool slx = (T(0.0) >= xs), sry = !(T(0.0) >= ys);
  const T sl = slx ? xs : T(0.0), sr = sry ? ys : T(0.0);
  const T slvl = sl - vnl, srvr = sr - vnr;
  const T rslvl = rol * slvl, rsrvr = ror * srvr;
  const T drsvi = T(1.0) / (rsrvr - rslvl);
  T tA = T(0), tB = T(1), tr = T(0), theta0 = T(1), theta = T(1), mccu = T(0), mcc = T(1), mccx = T(0);
  if (LOW) {
    tA = cmax - tmin(dv0, T(0.0));
    tB = cmax - tmin(dv1, T(0.0));
    tr = tA / tB;
    theta0 = tmin(T(1.0), tr);
    theta = theta0;
    theta *= theta;
    theta *= theta;
    mccx = tmax(ccl2, ccr2);
    mccu = dsqrt(mccx) / cmax;
    mcc = tmin(T(1.0), mccu);
  }
  const T Nvnc = LOW ? (rsrvr * vnr - rslvl * vnl - theta * (ptr - ptl)) : (rsrvr * vnr - rslvl * vnl - ptr + ptl);
  const T vnc = Nvnc * drsvi;
  const T kp = LOW ? (mcc * (T(2.0) - mcc)) : T(1.0);
  const T Nptc = LOW ? (rsrvr * ptl - rslvl * ptr + kp * rsrvr * rslvl * (vnr - vnl))
                     : (rsrvr * ptl - rslvl * ptr + rsrvr * rslvl * (vnr - vnl));
  const T ptc = Nptc * drsvi;
  const T slvc = sl - vnc, srvc = sr - vnc;
  const T ro2l = rslvl / slvc, ro2r = rsrvr / srvc;
  const T rhdl = rslvl * slvc - bnc2, rhdr = rsrvr * srvc - bnc2;
  const bool gl_ = dabs(rhdl) > FL_EPS, gr_ = dabs(rhdr) > FL_EPS;
  T rhdli = T(0), rhnvl = T(0), rhnbl = T(0), Al = T(0), Bl = T(0);
  T rhdri = T(0), rhnvr = T(0), rhnbr = T(0), Ar = T(0), Br = T(0);
  T vt2l = vtl, vu2l = vul, bt2l = btl, bu2l = bul, vt2r = vtr, vu2r = vur, bt2r = btr, bu2r = bur;
  if (gl_) {
    rhdli = T(1.0) / rhdl;
    rhnvl = (vnl - vnc) * bn;
    rhnbl = rslvl * slvl - bnc2;
    Al = rhnvl * rhdli;
    Bl = rhnbl * rhdli;
    vt2l = vtl + Al * btl;
    vu2l = vul + Al * bul;
    bt2l = Bl * btl;
    bu2l = Bl * bul;
  }
  if (gr_) {
    rhdri = T(1.0) / rhdr;
    rhnvr = (vnr - vnc) * bn;
    rhnbr = rsrvr * srvr - bnc2;
    Ar = rhnvr * rhdri;
    Br = rhnbr * rhdri;
    vt2r = vtr + Ar * btr;
    vu2r = vur + Ar * bur;
    bt2r = Br * btr;
    bu2r = Br * bur;
  }
  const T vb2l = vt2l * bt2l + vu2l * bu2l, vb2r = vt2r * bt2r + vu2r * bu2r;
  const T Nen2l = slvl * enl - ptl * vnl + ptc * vnc + bn * (vbl - vb2l);
  const T Nen2r = srvr * enr - ptr * vnr + ptc * vnc + bn * (vbr - vb2r);
  const T en2l = Nen2l / slvc, en2r = Nen2r / srvc;
  const T rro2l = dsqrt(ro2l), rro2r = dsqrt(ro2r);
  const T rro2i = T(1.0) / (rro2r + rro2l);
  const T Nvt = rro2r * vt2r + rro2l * vt2l + (bt2r - bt2l) * sgn;
  const T Nvu = rro2r * vu2r + rro2l * vu2l + (bu2r - bu2l) * sgn;
  const T Nbt = rro2l * bt2r + rro2r * bt2l + rro2r * rro2l * (vt2r - vt2l) * sgn;
  const T Nbu = rro2l * bu2r + rro2r * bu2l + rro2r * rro2l * (vu2r - vu2l) * sgn;
  const T vt3m = Nvt * rro2i, vu3m = Nvu * rro2i, bt3m = Nbt * rro2i, bu3m = Nbu * rro2i;
  const T vb3m = vt3m * bt3m + vu3m * bu3m;
  const T en3l = en2l - rro2l * (vb2l - vb3m) * sgn;
  const T en3r = en2r + rro2r * (vb2r - vb3m) * sgn;
  (void)en3l; (void)en3r;
  const T hl = (vnc > T(0)) ? T(1.0) : T(0.0);
  const T hr = T(1.0) - hl;
  const T h2l = (vnc - dabs(bn) / rro2l > T(0)) ? T(1.0) : T(0.0);
  const T h3l = (T(1.0) - h2l) * hl;
  const T h2r = (vnc + dabs(bn) / rro2r > T(0)) ? T(0.0) : T(1.0);
  const T h3r = (T(1.0) - h2r) * hr;
  const T rou = ro2l * hl + ro2r * hr;
  const T vtu = vt2l * h2l + vt3m * (h3l + h3r) + vt2r * h2r;
  const T vuu = vu2l * h2l + vu3m * (h3l + h3r) + vu2r * h2r;
  const T btu = bt2l * h2l + bt3m * (h3l + h3r) + bt2r * h2r;
  const T buu = bu2l * h2l + bu3m * (h3l + h3r) + bu2r * h2r;
  const T enu = en2l * h2l + en3l * h3l + en3r * h3r + en2r * h2r;
  // ---- reverse
  const T g0 = gF[0], g1 = gF[1], g2 = gF[2], g3 = gF[3], g4 = gF[4], g5 = gF[5], g6 = gF[6];
  T grou = vnc * g0 + vnc * vnc * g1 + vtu * vnc * g2 + vuu * vnc * g3;
  T gvnc = rou * g0 + T(2.0) * rou * vnc * g1 + rou * vtu * g2 + rou * vuu * g3 + btu * g4 + buu * g5 + (enu + ptc) * g6;
  T gptc = g1 + vnc * g6;
  T gbnc2 = T(-0.5) * g1;
  T gbn_ = -btu * g2 - buu * g3 - vtu * g4 - vuu * g5 - (vtu * btu + vuu * buu) * g6;
  const T gvtu = rou * vnc * g2 - bn * g4 - bn * btu * g6;
  const T gbtu = -bn * g2 + vnc * g4 - bn * vtu * g6;
  const T gvuu = rou * vnc * g3 - bn * g5 - bn * buu * g6;
  const T gbuu = -bn * g3 + vnc * g5 - bn * vuu * g6;
  const T genu = vnc * g6;
  T gro2l = hl * grou, gro2r = hr * grou;
  const T h33 = h3l + h3r;
  T gvt2l = h2l * gvtu, gvt2r = h2r * gvtu, gvt3m = h33 * gvtu;
  T gvu2l = h2l * gvuu, gvu2r = h2r * gvuu, gvu3m = h33 * gvuu;
  T gbt2l = h2l * gbtu, gbt2r = h2r * gbtu, gbt3m = h33 * gbtu;
  T gbu2l = h2l * gbuu, gbu2r = h2r * gbuu, gbu3m = h33 * gbuu;
  T gen2l = h2l * genu, gen2r = h2r * genu;
  const T gen3l = h3l * genu, gen3r = h3r * genu;
  T grro2l = T(0), grro2r = T(0), gvb2l = T(0), gvb2r = T(0), gvb3m = T(0);
  gen2l += gen3l;
  grro2l += -(vb2l - vb3m) * sgn * gen3l;
  gvb2l += -rro2l * sgn * gen3l;
  gvb3m += rro2l * sgn * gen3l;
  gen2r += gen3r;
  grro2r += (vb2r - vb3m) * sgn * gen3r;
  gvb2r += rro2r * sgn * gen3r;
  gvb3m += -rro2r * sgn * gen3r;
  gvt3m += bt3m * gvb3m;
  gbt3m += vt3m * gvb3m;
  gvu3m += bu3m * gvb3m;
  gbu3m += vu3m * gvb3m;
  T grro2i = Nvt * gvt3m + Nvu * gvu3m + Nbt * gbt3m + Nbu * gbu3m;
  {  // Nbt, Nbu
    const T gN = rro2i * gbt3m;
    grro2l += bt2r * gN + rro2r * (vt2r - vt2l) * sgn * gN;
    grro2r += bt2l * gN + rro2l * (vt2r - vt2l) * sgn * gN;
    gbt2r += rro2l * gN;
    gbt2l += rro2r * gN;
    gvt2r += rro2r * rro2l * sgn * gN;
    gvt2l -= rro2r * rro2l * sgn * gN;
    const T gM = rro2i * gbu3m;
    grro2l += bu2r * gM + rro2r * (vu2r - vu2l) * sgn * gM;
    grro2r += bu2l * gM + rro2l * (vu2r - vu2l) * sgn * gM;
    gbu2r += rro2l * gM;
    gbu2l += rro2r * gM;
    gvu2r += rro2r * rro2l * sgn * gM;
    gvu2l -= rro2r * rro2l * sgn * gM;
  }
  {  // Nvt, Nvu
    const T gN = rro2i * gvt3m;
    grro2r += vt2r * gN;
    gvt2r += rro2r * gN;
    grro2l += vt2l * gN;
    gvt2l += rro2l * gN;
    gbt2r += sgn * gN;
    gbt2l -= sgn * gN;
    const T gM = rro2i * gvu3m;
    grro2r += vu2r * gM;
    gvu2r += rro2r * gM;
    grro2l += vu2l * gM;
    gvu2l += rro2l * gM;
    gbu2r += sgn * gM;
    gbu2l -= sgn * gM;
  }
  {
    const T gs_ = -rro2i * rro2i * grro2i;
    grro2r += gs_;
    grro2l += gs_;
  }
  gro2l += (rro2l > T(0) ? T(0.5) / rro2l : T(0)) * grro2l;
  gro2r += (rro2r > T(0) ? T(0.5) / rro2r : T(0)) * grro2r;
  // en2l = Nen2l / slvc
  T gslvc = -en2l / slvc * gen2l, gsrvc = -en2r / srvc * gen2r;
  T gslvl = T(0), gsrvr = T(0), genl = T(0), genr = T(0), gptl = T(0), gptr = T(0), gvnl = T(0), gvnr = T(0);
  T gvbl = T(0), gvbr = T(0);
  {
    const T gN = gen2l / slvc;
    gslvl += enl * gN;
    genl += slvl * gN;
    gptl += -vnl * gN;
    gvnl += -ptl * gN;
    gptc += vnc * gN;
    gvnc += ptc * gN;
    gbn_ += (vbl - vb2l) * gN;
    gvbl += bn * gN;
    gvb2l += -bn * gN;
    const T gM = gen2r / srvc;
    gsrvr += enr * gM;
    genr += srvr * gM;
    gptr += -vnr * gM;
    gvnr += -ptr * gM;
    gptc += vnc * gM;
    gvnc += ptc * gM;
    gbn_ += (vbr - vb2r) * gM;
    gvbr += bn * gM;
    gvb2r += -bn * gM;
  }
  gvt2l += bt2l * gvb2l;
  gbt2l += vt2l * gvb2l;
  gvu2l += bu2l * gvb2l;
  gbu2l += vu2l * gvb2l;
  gvt2r += bt2r * gvb2r;
  gbt2r += vt2r * gvb2r;
  gvu2r += bu2r * gvb2r;
  gbu2r += vu2r * gvb2r;
  T grslvl = T(0), grsrvr = T(0), gvtl = T(0), gvul = T(0), gbtl = T(0), gbul = T(0);
  T gvtr = T(0), gvur = T(0), gbtr = T(0), gbur = T(0);
  if (gl_) {
    gvtl += gvt2l;
    gvul += gvu2l;
    const T gA = btl * gvt2l + bul * gvu2l;
    const T gB = btl * gbt2l + bul * gbu2l;
    gbtl += Al * gvt2l + Bl * gbt2l;
    gbul += Al * gvu2l + Bl * gbu2l;
    const T grhnvl = rhdli * gA, grhnbl = rhdli * gB;
    const T grhdli = rhnvl * gA + rhnbl * gB;
    const T grhdl = -rhdli * rhdli * grhdli;
    gvnl += bn * grhnvl;
    gvnc -= bn * grhnvl;
    gbn_ += (vnl - vnc) * grhnvl;
    grslvl += slvl * grhnbl + slvc * grhdl;
    gslvl += rslvl * grhnbl;
    gbnc2 -= grhnbl + grhdl;
    gslvc += rslvl * grhdl;
  } else {
    gvtl += gvt2l;
    gvul += gvu2l;
    gbtl += gbt2l;
    gbul += gbu2l;
  }
  if (gr_) {
    gvtr += gvt2r;
    gvur += gvu2r;
    const T gA = btr * gvt2r + bur * gvu2r;
    const T gB = btr * gbt2r + bur * gbu2r;
    gbtr += Ar * gvt2r + Br * gbt2r;
    gbur += Ar * gvu2r + Br * gbu2r;
    const T grhnvr = rhdri * gA, grhnbr = rhdri * gB;
    const T grhdri = rhnvr * gA + rhnbr * gB;
    const T grhdr = -rhdri * rhdri * grhdri;
    gvnr += bn * grhnvr;
    gvnc -= bn * grhnvr;
    gbn_ += (vnr - vnc) * grhnvr;
    grsrvr += srvr * grhnbr + srvc * grhdr;
    gsrvr += rsrvr * grhnbr;
    gbnc2 -= grhnbr + grhdr;
    gsrvc += rsrvr * grhdr;
  } else {
    gvtr += gvt2r;
    gvur += gvu2r;
    gbtr += gbt2r;
    gbur += gbu2r;
  }
  // ro2l = rslvl / slvc
  grslvl += gro2l / slvc;
  gslvc += -ro2l / slvc * gro2l;
  grsrvr += gro2r / srvc;
  gsrvc += -ro2r / srvc * gro2r;
  T gsl = gslvc, gsr = gsrvc;
  gvnc -= gslvc + gsrvc;
  // vnc, ptc
  T gdrsvi = Nvnc * gvnc + Nptc * gptc;
  T gcmax_low = T(0), gccl2 = T(0), gccr2 = T(0);
  {
    const T gN = drsvi * gvnc;
    grsrvr += vnr * gN;
    gvnr += rsrvr * gN;
    grslvl -= vnl * gN;
    gvnl -= rslvl * gN;
    gptr -= theta * gN;
    gptl += theta * gN;
    const T gM = drsvi * gptc;
    grsrvr += ptl * gM + kp * rslvl * (vnr - vnl) * gM;
    gptl += rsrvr * gM;
    grslvl += -ptr * gM + kp * rsrvr * (vnr - vnl) * gM;
    gptr -= rslvl * gM;
    gvnr += kp * rsrvr * rslvl * gM;
    gvnl -= kp * rsrvr * rslvl * gM;
    if (LOW) {
      // theta = min(1, tA / tB)^4, tA = cmax - min(dv0, 0), tB = cmax - min(dv1, 0)
      const T gtheta = -(ptr - ptl) * gN;
      if (T(1.0) >= tr) {
        const T gtr = gtheta * T(4.0) * theta0 * theta0 * theta0;
        const T gtA = gtr / tB, gtB = -gtr * tr / tB;
        gcmax_low += gtA + gtB;
        if (!(dv0 >= T(0.0))) gdv0 += -gtA;
        if (!(dv1 >= T(0.0))) gdv1 += -gtB;
      }
      // kp = mcc (2 - mcc), mcc = min(1, sqrt(max(ccl2, ccr2)) / cmax)
      const T gmcc = rsrvr * rslvl * (vnr - vnl) * gM * (T(2.0) - T(2.0) * mcc);
      if (T(1.0) >= mccu) {
        const T sq = dsqrt(mccx);
        gcmax_low += -gmcc * mccu / cmax;
        const T gx = gmcc / cmax * (sq > T(0) ? T(0.5) / sq : T(0));
        if (ccl2 >= ccr2) gccl2 += gx;
        else gccr2 += gx;
      }
    }
  }
  {
    const T gd = -drsvi * drsvi * gdrsvi;
    grsrvr += gd;
    grslvl -= gd;
  }
  T grol = slvl * grslvl, gror = srvr * grsrvr;
  gslvl += rol * grslvl;
  gsrvr += ror * grsrvr;
  gsl += gslvl;
  gvnl -= gslvl;
  gsr += gsrvr;
  gvnr -= gsrvr;
  // sl, sr, cmax
  T gcmax = gcmax_low, gmnv = T(0), gmxv = T(0);
  if (slx) {
    gmnv += gsl;
    gcmax -= gsl;
  }
  if (sry) {
    gmxv += gsr;
    gcmax += gsr;
  }
  if (nr_min) {
    gvnr += gmnv;
    gvnl += gmxv;
  } else {
    gvnl += gmnv;
    gvnr += gmxv;
  }
  T gcfl2 = T(0), gcfr2 = T(0);
  if (selL) gcfl2 += (cfl > T(0) ? T(0.5) / cfl : T(0)) * gcmax;  // (LHLLD: cmax = cfl when selL)
  else gcfr2 += (cfr > T(0) ? T(0.5) / cfr : T(0)) * gcmax;
  T groli = T(0), grori = T(0), gpml = T(0), gpmr = T(0), gprl = T(0), gprr = T(0);
  T gvl2x = T(0), gvr2x = T(0);
  if (LOW) {  // cc*2 = 0.5 (cb*2c + sqrt|e*c|), cb*2c = v*2 + ca*2 + 2 pm* ro*i
    T gcb = T(0.5) * gccl2;
    const T ge = T(0.5) * gccl2 * (selc > T(0) ? T(0.5) / selc : T(0)) * sgnd(elc);
    gcb += ge * T(2.0) * cbl2c;
    gvl2x += -T(4.0) * cal2 * ge + gcb;
    const T gcal2 = -T(4.0) * vl2 * ge + gcb;
    gpml += T(2.0) * roli * gcb;
    groli += T(2.0) * pml * gcb + bnc2 * gcal2;
    gbnc2 += roli * gcal2;
    gcb = T(0.5) * gccr2;
    const T ge2 = T(0.5) * gccr2 * (serc > T(0) ? T(0.5) / serc : T(0)) * sgnd(erc);
    gcb += ge2 * T(2.0) * cbr2c;
    gvr2x += -T(4.0) * car2 * ge2 + gcb;
    const T gcar2 = -T(4.0) * vr2 * ge2 + gcb;
    gpmr += T(2.0) * rori * gcb;
    grori += T(2.0) * pmr * gcb + bnc2 * gcar2;
    gbnc2 += rori * gcar2;
  }
  {  // left fast speed
    T gcb = T(0.5) * gcfl2;
    const T gsq = T(0.5) * gcfl2;
    const T gd = (sql > T(0) ? T(0.5) / sql : T(0)) * sgnd(dl) * gsq;
    gcb += T(2.0) * cbl2 * gd;
    const T gcl2 = gcb - T(4.0) * cal2 * gd;
    const T gcal2 = gcb - T(4.0) * cl2 * gd;
    gpml += T(2.0) * roli * gcb;
    groli += T(2.0) * pml * gcb + bnc2 * gcal2 + gam * prl * gcl2;
    gbnc2 += roli * gcal2;
    gprl += gam * roli * gcl2;
  }
  {  // right fast speed
    T gcb = T(0.5) * gcfr2;
    const T gsq = T(0.5) * gcfr2;
    const T gd = (sqr > T(0) ? T(0.5) / sqr : T(0)) * sgnd(dr) * gsq;
    gcb += T(2.0) * cbr2 * gd;
    const T gcr2 = gcb - T(4.0) * car2 * gd;
    const T gcar2 = gcb - T(4.0) * cr2 * gd;
    gpmr += T(2.0) * rori * gcb;
    grori += T(2.0) * pmr * gcb + bnc2 * gcar2 + gam * prr * gcr2;
    gbnc2 += rori * gcar2;
    gprr += gam * rori * gcr2;
  }
  // vbl, enl, ptl, pml
  gvtl += btl * gvbl;
  gbtl += vtl * gvbl;
  gvul += bul * gvbl;
  gbul += vul * gvbl;
  gvtr += btr * gvbr;
  gbtr += vtr * gvbr;
  gvur += bur * gvbr;
  gbur += vur * gvbr;
  gprl += gm1i * genl + gptl;
  gpml += genl + gptl;
  grol += T(0.5) * vl2 * genl;
  const T gvl2 = T(0.5) * rol * genl + gvl2x;
  gprr += gm1i * genr + gptr;
  gpmr += genr + gptr;
  gror += T(0.5) * vr2 * genr;
  const T gvr2 = T(0.5) * ror * genr + gvr2x;
  gvnl += T(2.0) * vnl * gvl2;
  gvtl += T(2.0) * vtl * gvl2;
  gvul += T(2.0) * vul * gvl2;
  gvnr += T(2.0) * vnr * gvr2;
  gvtr += T(2.0) * vtr * gvr2;
  gvur += T(2.0) * vur * gvr2;
  gbtl += btl * gpml;
  gbul += bul * gpml;
  gbtr += btr * gpmr;
  gbur += bur * gpmr;
  grol += -roli * roli * groli;
  gror += -rori * rori * grori;
  gbn_ += T(2.0) * bn * gbnc2;
  gL[0] += grol; gL[1] += gvnl; gL[2] += gvtl; gL[3] += gvul; gL[4] += gbtl; gL[5] += gbul; gL[6] += gprl;
  gR[0] += gror; gR[1] += gvnr; gR[2] += gvtr; gR[3] += gvur; gR[4] += gbtr; gR[5] += gbur; gR[6] += gprr;
  gbn += gbn_;
}

// HLLD (calc_flux_hlld): no shock-sensor inputs
template <typename T>
QAS_HD void hlld_adj(const T* L, const T* R, T bn, T gam, const T* gF, T* gL, T* gR, T& gbn)
{
  T gd0 = T(0), gd1 = T(0);
  hlld_adj_g<false>(L, R, bn, gam, T(0), T(0), gF, gL, gR, gbn, gd0, gd1);
}

// ---------------------------------------------------------------- MLAU
// gL, gR (7 each), gbn, gdv0, gdv1 += adjoint of calc_flux_mlau
// (mhd_flux.c:7-192, flux_mlau of qas_device.hpp): the forward is recomputed
// with the same expressions, then reversed statement by statement; min/max
// route the derivative to the operand taken, the subsonic/supersonic
// polynomial branches, the flux direction and the FL_EPS guards stay as
// the forward takes them.
template <typename T>
QAS_HD void mlau_adj(const T* L, const T* R, T bnc, T gamma, T dv0, T dv1, const T* gF, T* gL, T* gR, T& gbn,
                     T& gdv0, T& gdv1)
{
  const T FL_EPS = T(QAS_FL_EPS);
  const T rol = L[0], vnl = L[1], vtl = L[2], vul = L[3], btl = L[4], bul = L[5], prl = L[6];
  const T ror = R[0], vnr = R[1], vtr = R[2], vur = R[3], btr = R[4], bur = R[5], prr = R[6];
  // ---- forward
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
  const T cbl2a = cl2 + cal2 + T(2.0) * pml * roli;
  const T cbr2a = cr2 + car2 + T(2.0) * pmr * rori;
  const T dl = cbl2a * cbl2a - T(4.0) * cl2 * cal2, dr = cbr2a * cbr2a - T(4.0) * cr2 * car2;
  const T sql = dsqrt(dabs(dl)), sqr = dsqrt(dabs(dr));
  const T cfl2 = T(0.5) * (cbl2a + sql);
  const T cfr2 = T(0.5) * (cbr2a + sqr);
  const T cbl2b = vl2 + cal2 + T(2.0) * pml * roli;
  const T cbr2b = vr2 + car2 + T(2.0) * pmr * rori;
  const T el = cbl2b * cbl2b - T(4.0) * vl2 * cal2, er = cbr2b * cbr2b - T(4.0) * vr2 * car2;
  const T sel = dsqrt(dabs(el)), ser = dsqrt(dabs(er));
  const T ccl2 = T(0.5) * (cbl2b + sel);
  const T ccr2 = T(0.5) * (cbr2b + ser);
  const T cmax = dsqrt(tmax(cfl2, cfr2));
  const T cfn = cmax;
  const T cfni = T(1.0) / cfn;
  const T xsl = tmin(vnl, vnr) - cmax, xsr = tmax(vnl, vnr) + cmax;
  const T sl = tmin(T(0.0), xsl);
  const T sr = tmax(T(0.0), xsr);
  const T mal = vnl * cfni;
  const T mar = vnr * cfni;
  const T snl = (mal > T(0)) ? T(1.0) : T(-1.0);
  const T snr = (mar > T(0)) ? T(1.0) : T(-1.0);
  const T tA = cfn - tmin(dv0, T(0.0)), tB = cfn - tmin(dv1, T(0.0));
  const T tr = tA / tB;
  const T theta0 = tmin(T(1.0), tr);
  T theta = theta0;
  theta *= theta;
  theta *= theta;
  const T alpha = T(0.1875);
  const T beta = T(0.125);
  const T malp1 = mal + T(1.0);
  const T marm1 = mar - T(1.0);
  const T mal2m1 = mal * mal - T(1.0);
  const T mar2m1 = mar * mar - T(1.0);
  const bool supl = snl * mal > T(1.0), supr = snr * mar > T(1.0);
  T bfl, ffl, bfr, ffr;
  if (supl) {
    bfl = T(0.5) * (T(1.0) + snl);
    ffl = bfl * mal;
  } else {
    bfl = +T(0.25) * malp1 * malp1 * (T(2.0) - mal) + alpha * mal * mal2m1 * mal2m1;
    ffl = +T(0.25) * malp1 * malp1 + beta * mal2m1 * mal2m1;
  }
  if (supr) {
    bfr = T(0.5) * (T(1.0) - snr);
    ffr = bfr * mar;
  } else {
    bfr = +T(0.25) * marm1 * marm1 * (T(2.0) + mar) - alpha * mar * mar2m1 * mar2m1;
    ffr = T(-0.25) * marm1 * marm1 - beta * mar2m1 * mar2m1;
  }
  const T mnc = ffl + ffr;
  const T mflux0 = mnc * cfn;
  const T omx = T(1.0) - dabs(mnc);
  const T mx = tmax(omx, T(0.0));
  const T rs = rol + ror;
  const T dp = ptr - ptl;
  const T mflux1 = mflux0 - mx * theta * cfni * dp / rs;
  const T dir = (mflux1 > T(0)) ? T(1.0) : T(-1.0);
  const T dirl = T(0.5) * (T(1.0) + dir);
  const T dirr = T(1.0) - dirl;
  const T rup = dirl * rol + dirr * ror;
  const T mflux = mflux1 * rup;
  const T mfl = mflux * dirl;
  const T mfr = mflux - mfl;
  const T ptot0 = bfl * ptl + bfr * ptr;
  const T cfmod = dsqrt(tmax(ccl2, ccr2));
  const T dvn = vnr - vnl;
  const T ptot1 = ptot0 - T(0.25) * bfl * bfr * rs * cfmod * dvn;
  const T k1 = T(1.0) - cfmod * cfni, k2 = bfl + bfr - T(1.0), k3 = ptl + ptr;
  const T ptot = ptot1 - T(0.5) * k1 * k2 * k3;
  const T slvl = sl - vnl;
  const T srvr = sr - vnr;
  const T snc = dirl * sl + dirr * sr;
  const T vndnm = mflux + dirl * rol * slvl + dirr * ror * srvr;
  const bool vdiv = (dabs(vndnm) > FL_EPS) && (snc != T(0));
  const T vnc = vdiv ? (mflux * snc / vndnm) : (dirl * vnl + dirr * vnr);
  const T slvc = sl - vnc;
  const T srvc = sr - vnc;
  const T slvci = T(1.0) / slvc;
  const T srvci = T(1.0) / srvc;
  const T rhdl = rol * slvl * slvc - bnc2;
  const T rhdr = ror * srvr * srvc - bnc2;
  const bool okl = dabs(rhdl) > FL_EPS, okr = dabs(rhdr) > FL_EPS;
  T ql = T(0.0), qr = T(0.0);
  T dvtl = T(0.0), dvul = T(0.0), dbtl = T(0.0), dbul = T(0.0);
  T dvtr = T(0.0), dvur = T(0.0), dbtr = T(0.0), dbur = T(0.0);
  if (okl) {
    ql = (vnc - vnl) / rhdl;
    dvtl = -bnc * btl * ql;
    dvul = -bnc * bul * ql;
    dbtl = bnc2 * btl * ql * slvci;
    dbul = bnc2 * bul * ql * slvci;
  }
  if (okr) {
    qr = (vnc - vnr) / rhdr;
    dvtr = -bnc * btr * qr;
    dvur = -bnc * bur * qr;
    dbtr = bnc2 * btr * qr * srvci;
    dbur = bnc2 * bur * qr * srvci;
  }
  const T cmpl = slvl * slvci;
  const T cmpr = srvr * srvci;
  const T ro2l = rol * cmpl;
  const T srol = dsqrt(ro2l);
  const T vt2l = dvtl + vtl, vu2l = dvul + vul;
  const T bt2l = dbtl + btl * cmpl, bu2l = dbul + bul * cmpl;
  const T ro2r = ror * cmpr;
  const T sror = dsqrt(ro2r);
  const T vt2r = dvtr + vtr, vu2r = dvur + vur;
  const T bt2r = dbtr + btr * cmpr, bu2r = dbur + bur * cmpr;
  const T sroi = T(1.0) / (srol + sror);
  const T srollri = srol * sroi;
  const T srorlri = sror * sroi;
  // clamp(w) = min(max(0, w), bnca) as tmin(tmax(0, w), bnca); which: 0 -> 0, 1 -> w, 2 -> bnca
  auto clampsel = [&](T w) -> int {
    const T m = (T(0.0) >= w) ? T(0.0) : w;
    return (m >= bnca) ? 2 : ((T(0.0) >= w) ? 0 : 1);
  };
  const T wflb = srollri * (bnca + sror * vnc);
  const T wfrb = srorlri * (bnca - srol * vnc);
  const T wflv = srorlri * (bnca + mflux / sror);
  const T wfrv = srollri * (bnca - mflux / srol);
  const int cflb = clampsel(wflb), cfrb = clampsel(wfrb), cflv = clampsel(wflv), cfrv = clampsel(wfrv);
  auto clampval = [&](T w, int c) -> T { return c == 2 ? bnca : (c == 1 ? w : T(0.0)); };
  const T bnflb = sgn * clampval(wflb, cflb);
  const T bnfrb = sgn * clampval(wfrb, cfrb);
  const T bnflv = sgn * clampval(wflv, cflv);
  const T bnfrv = sgn * clampval(wfrv, cfrv);
  const T zsrv = bnca - dir * vnc * (dirl * srol + dirr * sror);
  const T bnsrv = tmax(zsrv, T(0.0));
  const T ccb = sroi * bnsrv;
  const T ccv = srol * sror * ccb;
  const T bnvt = bnflb * vt2l + bnfrb * vt2r + ccb * (bt2r - bt2l) - vnc * (dirl * dbtl + dirr * dbtr);
  const T bnvu = bnflb * vu2l + bnfrb * vu2r + ccb * (bu2r - bu2l) - vnc * (dirl * dbul + dirr * dbur);
  const T bnbt = bnflv * bt2l + bnfrv * bt2r + ccv * (vt2r - vt2l) - mflux * (dirl * dvtl + dirr * dvtr);
  const T bnbu = bnflv * bu2l + bnfrv * bu2r + ccv * (vu2r - vu2l) - mflux * (dirl * dvul + dirr * dvur);
  const T fmt = mfl * vtl + mfr * vtr - bnbt;
  const T fmu = mfl * vul + mfr * vur - bnbu;
  const T fbt = mfl * btl * roli + mfr * btr * rori - bnvt;
  const T fbu = mfl * bul * roli + mfr * bur * rori - bnvu;
  const T vb1 = dirl * (vtl * btl + vul * bul) + dirr * (vtr * btr + vur * bur);
  const T vb2 = dirl * (vt2l * bt2l + vu2l * bu2l) + dirr * (vt2r * bt2r + vu2r * bu2r);
  const T rvb = mflux * vnc - bnc2;
  const T rv2b2 = rvb * rvb;
  const bool v3 = rv2b2 > FL_EPS * FL_EPS;
  const T a1 = fmt * vnc + fbt * bnc, b1 = fmt * bnc + fbt * mflux;
  const T a2 = fmu * vnc + fbu * bnc, b2 = fmu * bnc + fbu * mflux;
  const T vb3 = v3 ? (a1 * b1 + a2 * b2) / rv2b2 : T(0.0);
  const T bx = snc * vb2 - vnc * vb1;
  const T by = dirl * slvci + dirr * srvci;
  // ---- reverse
  T grol = 0, gvnl = 0, gvtl = 0, gvul = 0, gbtl = 0, gbul = 0, gprl = 0;
  T gror = 0, gvnr = 0, gvtr = 0, gvur = 0, gbtr = 0, gbur = 0, gprr = 0;
  T gbnc = 0, gbnc2 = 0, gbnca = 0;
  T gmflux = gF[0], gfmn = gF[1], gfmt = gF[2], gfmu = gF[3], gfbt = gF[4], gfbu = gF[5];
  // F6 = mfl hnl + mfr hnr - bnvb
  T gmfl = gF[6] * hnl, gmfr = gF[6] * hnr;
  T ghnl = gF[6] * mfl, ghnr = gF[6] * mfr;
  const T gbnvb = -gF[6];
  // bnvb = sgn (bnca bx by + bnsrv (vb3 - vb2))
  T gvb2 = 0, gvb1 = 0, gvb3 = 0, gsnc = 0, gvnc = 0, gslvci = 0, gsrvci = 0, gbnsrv = 0;
  {
    const T t = sgn * gbnvb;
    gbnca += t * bx * by;
    const T gbx = t * bnca * by, gby = t * bnca * bx;
    gbnsrv += t * (vb3 - vb2);
    gvb3 += t * bnsrv;
    gvb2 += -t * bnsrv;
    gsnc += gbx * vb2;
    gvb2 += gbx * snc;
    gvnc += -gbx * vb1;
    gvb1 += -gbx * vnc;
    gslvci += gby * dirl;
    gsrvci += gby * dirr;
  }
  // vb3
  if (v3) {
    const T gN = gvb3 / rv2b2;
    const T gD = -gvb3 * vb3 / rv2b2;
    const T ga1 = gN * b1, gb1 = gN * a1, ga2 = gN * b2, gb2 = gN * a2;
    gfmt += ga1 * vnc + gb1 * bnc;
    gfbt += ga1 * bnc + gb1 * mflux;
    gfmu += ga2 * vnc + gb2 * bnc;
    gfbu += ga2 * bnc + gb2 * mflux;
    gvnc += ga1 * fmt + ga2 * fmu;
    gbnc += ga1 * fbt + gb1 * fmt + ga2 * fbu + gb2 * fmu;
    gmflux += gb1 * fbt + gb2 * fbu;
    const T grvb = T(2.0) * rvb * gD;
    gmflux += grvb * vnc;
    gvnc += grvb * mflux;
    gbnc2 += -grvb;
  }
  // vb2, vb1
  T gvt2l = gvb2 * dirl * bt2l, gbt2l = gvb2 * dirl * vt2l, gvu2l = gvb2 * dirl * bu2l, gbu2l = gvb2 * dirl * vu2l;
  T gvt2r = gvb2 * dirr * bt2r, gbt2r = gvb2 * dirr * vt2r, gvu2r = gvb2 * dirr * bu2r, gbu2r = gvb2 * dirr * vu2r;
  gvtl += gvb1 * dirl * btl;
  gbtl += gvb1 * dirl * vtl;
  gvul += gvb1 * dirl * bul;
  gbul += gvb1 * dirl * vul;
  gvtr += gvb1 * dirr * btr;
  gbtr += gvb1 * dirr * vtr;
  gvur += gvb1 * dirr * bur;
  gbur += gvb1 * dirr * vur;
  // fbu, fbt, fmu, fmt, fmn
  T groli = 0, grori = 0;
  gmfl += gfbu * bul * roli + gfbt * btl * roli + gfmu * vul + gfmt * vtl + gfmn * vnl;
  gmfr += gfbu * bur * rori + gfbt * btr * rori + gfmu * vur + gfmt * vtr + gfmn * vnr;
  gbul += gfbu * mfl * roli;
  gbur += gfbu * mfr * rori;
  gbtl += gfbt * mfl * roli;
  gbtr += gfbt * mfr * rori;
  groli += gfbu * mfl * bul + gfbt * mfl * btl;
  grori += gfbu * mfr * bur + gfbt * mfr * btr;
  gvul += gfmu * mfl;
  gvur += gfmu * mfr;
  gvtl += gfmt * mfl;
  gvtr += gfmt * mfr;
  gvnl += gfmn * mfl;
  gvnr += gfmn * mfr;
  const T gbnvu = -gfbu, gbnvt = -gfbt, gbnbu = -gfmu, gbnbt = -gfmt;
  T gptot = gfmn;
  gbnc2 += T(-0.5) * gfmn;
  // bnbu, bnbt, bnvu, bnvt
  T gdvul = 0, gdvur = 0, gdvtl = 0, gdvtr = 0, gdbul = 0, gdbur = 0, gdbtl = 0, gdbtr = 0;
  gmflux += -gbnbu * (dirl * dvul + dirr * dvur) - gbnbt * (dirl * dvtl + dirr * dvtr);
  gdvul += -gbnbu * mflux * dirl;
  gdvur += -gbnbu * mflux * dirr;
  gdvtl += -gbnbt * mflux * dirl;
  gdvtr += -gbnbt * mflux * dirr;
  gvnc += -gbnvu * (dirl * dbul + dirr * dbur) - gbnvt * (dirl * dbtl + dirr * dbtr);
  gdbul += -gbnvu * vnc * dirl;
  gdbur += -gbnvu * vnc * dirr;
  gdbtl += -gbnvt * vnc * dirl;
  gdbtr += -gbnvt * vnc * dirr;
  T gbnflv = gbnbu * bu2l + gbnbt * bt2l, gbnfrv = gbnbu * bu2r + gbnbt * bt2r;
  gbu2l += gbnbu * bnflv;
  gbu2r += gbnbu * bnfrv;
  gbt2l += gbnbt * bnflv;
  gbt2r += gbnbt * bnfrv;
  T gccv = gbnbu * (vu2r - vu2l) + gbnbt * (vt2r - vt2l);
  gvu2r += gbnbu * ccv;
  gvu2l -= gbnbu * ccv;
  gvt2r += gbnbt * ccv;
  gvt2l -= gbnbt * ccv;
  T gbnflb = gbnvu * vu2l + gbnvt * vt2l, gbnfrb = gbnvu * vu2r + gbnvt * vt2r;
  gvu2l += gbnvu * bnflb;
  gvu2r += gbnvu * bnfrb;
  gvt2l += gbnvt * bnflb;
  gvt2r += gbnvt * bnfrb;
  T gccb = gbnvu * (bu2r - bu2l) + gbnvt * (bt2r - bt2l);
  gbu2r += gbnvu * ccb;
  gbu2l -= gbnvu * ccb;
  gbt2r += gbnvt * ccb;
  gbt2l -= gbnvt * ccb;
  // ccv = srol sror ccb ; ccb = sroi bnsrv
  T gsrol = gccv * sror * ccb, gsror = gccv * srol * ccb;
  gccb += gccv * srol * sror;
  T gsroi = gccb * bnsrv;
  gbnsrv += gccb * sroi;
  // bnsrv = max(zsrv, 0)
  if (zsrv >= T(0.0)) {
    gbnca += gbnsrv;
    gvnc += -gbnsrv * dir * (dirl * srol + dirr * sror);
    gsrol += -gbnsrv * dir * vnc * dirl;
    gsror += -gbnsrv * dir * vnc * dirr;
  }
  // the four clamps
  T gsrollri = 0, gsrorlri = 0;
  {
    T g = sgn * gbnfrv;  // wfrv = srollri (bnca - mflux / srol)
    if (cfrv == 2) gbnca += g;
    else if (cfrv == 1) {
      gsrollri += g * (bnca - mflux / srol);
      gbnca += g * srollri;
      gmflux += -g * srollri / srol;
      gsrol += g * srollri * mflux / (srol * srol);
    }
    g = sgn * gbnflv;  // wflv = srorlri (bnca + mflux / sror)
    if (cflv == 2) gbnca += g;
    else if (cflv == 1) {
      gsrorlri += g * (bnca + mflux / sror);
      gbnca += g * srorlri;
      gmflux += g * srorlri / sror;
      gsror += -g * srorlri * mflux / (sror * sror);
    }
    g = sgn * gbnfrb;  // wfrb = srorlri (bnca - srol vnc)
    if (cfrb == 2) gbnca += g;
    else if (cfrb == 1) {
      gsrorlri += g * (bnca - srol * vnc);
      gbnca += g * srorlri;
      gsrol += -g * srorlri * vnc;
      gvnc += -g * srorlri * srol;
    }
    g = sgn * gbnflb;  // wflb = srollri (bnca + sror vnc)
    if (cflb == 2) gbnca += g;
    else if (cflb == 1) {
      gsrollri += g * (bnca + sror * vnc);
      gbnca += g * srollri;
      gsror += g * srollri * vnc;
      gvnc += g * srollri * sror;
    }
  }
  gsrol += gsrollri * sroi;
  gsroi += gsrollri * srol;
  gsror += gsrorlri * sroi;
  gsroi += gsrorlri * sror;
  {
    const T gsum = -gsroi * sroi * sroi;
    gsrol += gsum;
    gsror += gsum;
  }
  // right states
  T gcmpr = 0, gcmpl = 0;
  gdbur += gbu2r;
  gbur += gbu2r * cmpr;
  gcmpr += gbu2r * bur;
  gdbtr += gbt2r;
  gbtr += gbt2r * cmpr;
  gcmpr += gbt2r * btr;
  gdvur += gvu2r;
  gvur += gvu2r;
  gdvtr += gvt2r;
  gvtr += gvt2r;
  {
    const T gro2r = gsror * (sror > T(0) ? T(0.5) / sror : T(0));
    gror += gro2r * cmpr;
    gcmpr += gro2r * ror;
  }
  gdbul += gbu2l;
  gbul += gbu2l * cmpl;
  gcmpl += gbu2l * bul;
  gdbtl += gbt2l;
  gbtl += gbt2l * cmpl;
  gcmpl += gbt2l * btl;
  gdvul += gvu2l;
  gvul += gvu2l;
  gdvtl += gvt2l;
  gvtl += gvt2l;
  {
    const T gro2l = gsrol * (srol > T(0) ? T(0.5) / srol : T(0));
    grol += gro2l * cmpl;
    gcmpl += gro2l * rol;
  }
  T gsrvr = gcmpr * srvci, gslvl = gcmpl * slvci;
  gsrvci += gcmpr * srvr;
  gslvci += gcmpl * slvl;
  // transverse corrections
  T gslvc = 0, gsrvc = 0;
  if (okr) {
    const T gq = -bnc * btr * gdvtr - bnc * bur * gdvur + bnc2 * btr * srvci * gdbtr + bnc2 * bur * srvci * gdbur;
    gbnc += -btr * qr * gdvtr - bur * qr * gdvur;
    gbtr += -bnc * qr * gdvtr + bnc2 * qr * srvci * gdbtr;
    gbur += -bnc * qr * gdvur + bnc2 * qr * srvci * gdbur;
    gbnc2 += btr * qr * srvci * gdbtr + bur * qr * srvci * gdbur;
    gsrvci += bnc2 * btr * qr * gdbtr + bnc2 * bur * qr * gdbur;
    gvnc += gq / rhdr;
    gvnr += -gq / rhdr;
    const T grhd = -gq * qr / rhdr;
    gror += grhd * srvr * srvc;
    gsrvr += grhd * ror * srvc;
    gsrvc += grhd * ror * srvr;
    gbnc2 += -grhd;
  }
  if (okl) {
    const T gq = -bnc * btl * gdvtl - bnc * bul * gdvul + bnc2 * btl * slvci * gdbtl + bnc2 * bul * slvci * gdbul;
    gbnc += -btl * ql * gdvtl - bul * ql * gdvul;
    gbtl += -bnc * ql * gdvtl + bnc2 * ql * slvci * gdbtl;
    gbul += -bnc * ql * gdvul + bnc2 * ql * slvci * gdbul;
    gbnc2 += btl * ql * slvci * gdbtl + bul * ql * slvci * gdbul;
    gslvci += bnc2 * btl * ql * gdbtl + bnc2 * bul * ql * gdbul;
    gvnc += gq / rhdl;
    gvnl += -gq / rhdl;
    const T grhd = -gq * ql / rhdl;
    grol += grhd * slvl * slvc;
    gslvl += grhd * rol * slvc;
    gslvc += grhd * rol * slvl;
    gbnc2 += -grhd;
  }
  gsrvc += -gsrvci * srvci * srvci;
  gslvc += -gslvci * slvci * slvci;
  T gsl = gslvc, gsr = gsrvc;
  gvnc += -gslvc - gsrvc;
  // vnc
  if (vdiv) {
    gmflux += gvnc * snc / vndnm;
    gsnc += gvnc * mflux / vndnm;
    const T gvd = -gvnc * vnc / vndnm;
    gmflux += gvd;
    grol += gvd * dirl * slvl;
    gslvl += gvd * dirl * rol;
    gror += gvd * dirr * srvr;
    gsrvr += gvd * dirr * ror;
  } else {
    gvnl += gvnc * dirl;
    gvnr += gvnc * dirr;
  }
  gsl += gsnc * dirl + gslvl;
  gsr += gsnc * dirr + gsrvr;
  gvnl -= gslvl;
  gvnr -= gsrvr;
  // ptot
  T gbfl = 0, gbfr = 0, gptl = 0, gptr = 0, gcfmod = 0, gcfni = 0, grs = 0;
  {
    const T gk = T(-0.5) * gptot;
    const T gk1 = gk * k2 * k3, gk2 = gk * k1 * k3, gk3 = gk * k1 * k2;
    gcfmod += -gk1 * cfni;
    gcfni += -gk1 * cfmod;
    gbfl += gk2;
    gbfr += gk2;
    gptl += gk3;
    gptr += gk3;
    const T gP = -gptot;  // ptot1 = ptot0 - P, P = 0.25 bfl bfr rs cfmod dvn
    gbfl += gP * T(0.25) * bfr * rs * cfmod * dvn;
    gbfr += gP * T(0.25) * bfl * rs * cfmod * dvn;
    grs += gP * T(0.25) * bfl * bfr * cfmod * dvn;
    gcfmod += gP * T(0.25) * bfl * bfr * rs * dvn;
    const T gdvn = gP * T(0.25) * bfl * bfr * rs * cfmod;
    gvnr += gdvn;
    gvnl -= gdvn;
    gbfl += gptot * ptl;
    gptl += gptot * bfl;
    gbfr += gptot * ptr;
    gptr += gptot * bfr;
  }
  T gccl2 = 0, gccr2 = 0;
  {
    const T gm = gcfmod * (cfmod > T(0) ? T(0.5) / cfmod : T(0));
    if (ccl2 >= ccr2) gccl2 += gm;
    else gccr2 += gm;
  }
  // mfl = mflux dirl, mfr = mflux - mfl
  gmflux += gmfr;
  gmfl -= gmfr;
  gmflux += gmfl * dirl;
  // mflux = mflux1 rup
  const T gmflux1 = gmflux * rup;
  {
    const T grup = gmflux * mflux1;
    grol += grup * dirl;
    gror += grup * dirr;
  }
  // mflux1 = mflux0 - mx theta cfni dp / rs
  T gmnc = gmflux1 * cfn, gcfn = gmflux1 * mnc, gtheta = 0;
  {
    const T gpd = -gmflux1;
    const T pd = mx * theta * cfni * dp / rs;
    const T gmx = gpd * theta * cfni * dp / rs;
    gtheta += gpd * mx * cfni * dp / rs;
    gcfni += gpd * mx * theta * dp / rs;
    const T gdp = gpd * mx * theta * cfni / rs;
    grs += -gpd * pd / rs;
    gptr += gdp;
    gptl -= gdp;
    if (omx >= T(0.0)) gmnc += -gmx * sgnd(mnc);
  }
  grol += grs;
  gror += grs;
  // mnc = ffl + ffr ; the polynomial branches
  T gmal = 0, gmar = 0;
  {
    const T gffl = gmnc, gffr = gmnc;
    if (supl) gmal += gffl * bfl;
    else {
      const T dbfl = T(0.25) * (T(2.0) * malp1 * (T(2.0) - mal) - malp1 * malp1) +
                     alpha * (mal2m1 * mal2m1 + T(4.0) * mal * mal * mal2m1);
      const T dffl = T(0.5) * malp1 + T(4.0) * beta * mal * mal2m1;
      gmal += gbfl * dbfl + gffl * dffl;
    }
    if (supr) gmar += gffr * bfr;
    else {
      const T dbfr = T(0.25) * (T(2.0) * marm1 * (T(2.0) + mar) + marm1 * marm1) -
                     alpha * (mar2m1 * mar2m1 + T(4.0) * mar * mar * mar2m1);
      const T dffr = T(-0.5) * marm1 - T(4.0) * beta * mar * mar2m1;
      gmar += gbfr * dbfr + gffr * dffr;
    }
  }
  // theta = theta0^4, theta0 = min(1, tA / tB)
  if (T(1.0) >= tr) {
    const T gtr = gtheta * T(4.0) * theta0 * theta0 * theta0;
    const T gtA = gtr / tB, gtB = -gtr * tr / tB;
    gcfn += gtA + gtB;
    if (!(dv0 >= T(0.0))) gdv0 += -gtA;
    if (!(dv1 >= T(0.0))) gdv1 += -gtB;
  }
  // mal = vnl cfni, mar = vnr cfni
  gvnl += gmal * cfni;
  gvnr += gmar * cfni;
  gcfni += gmal * vnl + gmar * vnr;
  // sl, sr
  T gcmax = 0;
  if (T(0.0) >= xsl) {  // sl = xsl
    if (vnl >= vnr) gvnr += gsl;
    else gvnl += gsl;
    gcmax -= gsl;
  }
  if (!(T(0.0) >= xsr)) {  // sr = xsr
    if (vnl >= vnr) gvnl += gsr;
    else gvnr += gsr;
    gcmax += gsr;
  }
  gcfn += -gcfni * cfni * cfni;
  gcmax += gcfn;
  T gcfl2 = 0, gcfr2 = 0;
  {
    const T gm = gcmax * (cmax > T(0) ? T(0.5) / cmax : T(0));
    if (cfl2 >= cfr2) gcfl2 += gm;
    else gcfr2 += gm;
  }
  T gvl2 = 0, gvr2 = 0, gpml = 0, gpmr = 0, gcal2 = 0, gcar2 = 0, gcl2 = 0, gcr2 = 0;
  {  // cc*2 = 0.5 (cb*2b + sqrt|e*|), cb*2b = v*2 + ca*2 + 2 pm* ro*i
    T gcb = T(0.5) * gccl2;
    const T ge = T(0.5) * gccl2 * (sel > T(0) ? T(0.5) / sel : T(0)) * sgnd(el);
    gcb += ge * T(2.0) * cbl2b;
    gvl2 += -T(4.0) * cal2 * ge + gcb;
    gcal2 += -T(4.0) * vl2 * ge + gcb;
    gpml += T(2.0) * roli * gcb;
    groli += T(2.0) * pml * gcb;
    gcb = T(0.5) * gccr2;
    const T ge2 = T(0.5) * gccr2 * (ser > T(0) ? T(0.5) / ser : T(0)) * sgnd(er);
    gcb += ge2 * T(2.0) * cbr2b;
    gvr2 += -T(4.0) * car2 * ge2 + gcb;
    gcar2 += -T(4.0) * vr2 * ge2 + gcb;
    gpmr += T(2.0) * rori * gcb;
    grori += T(2.0) * pmr * gcb;
  }
  {  // cf*2 = 0.5 (cb*2a + sqrt|d*|), cb*2a = c*2 + ca*2 + 2 pm* ro*i
    T gcb = T(0.5) * gcfl2;
    const T gd = T(0.5) * gcfl2 * (sql > T(0) ? T(0.5) / sql : T(0)) * sgnd(dl);
    gcb += gd * T(2.0) * cbl2a;
    gcl2 += -T(4.0) * cal2 * gd + gcb;
    gcal2 += -T(4.0) * cl2 * gd + gcb;
    gpml += T(2.0) * roli * gcb;
    groli += T(2.0) * pml * gcb;
    gcb = T(0.5) * gcfr2;
    const T gd2 = T(0.5) * gcfr2 * (sqr > T(0) ? T(0.5) / sqr : T(0)) * sgnd(dr);
    gcb += gd2 * T(2.0) * cbr2a;
    gcr2 += -T(4.0) * car2 * gd2 + gcb;
    gcar2 += -T(4.0) * cr2 * gd2 + gcb;
    gpmr += T(2.0) * rori * gcb;
    grori += T(2.0) * pmr * gcb;
  }
  gbnc2 += gcal2 * roli + gcar2 * rori;
  groli += gcal2 * bnc2 + gcl2 * gamma * prl;
  grori += gcar2 * bnc2 + gcr2 * gamma * prr;
  gprl += gcl2 * gamma * roli;
  gprr += gcr2 * gamma * rori;
  // h*n = (en* + pt*) ro*i ; en* = gm1i pr* + pm* + ro* v*2 / 2 ; pt* = pr* + pm*
  {
    const T genl = ghnl * roli;
    gptl += ghnl * roli;
    groli += ghnl * (enl + ptl);
    gprl += genl * gammam1i;
    gpml += genl;
    grol += genl * T(0.5) * vl2;
    gvl2 += genl * T(0.5) * rol;
    const T genr = ghnr * rori;
    gptr += ghnr * rori;
    grori += ghnr * (enr + ptr);
    gprr += genr * gammam1i;
    gpmr += genr;
    gror += genr * T(0.5) * vr2;
    gvr2 += genr * T(0.5) * ror;
  }
  gprl += gptl;
  gpml += gptl;
  gprr += gptr;
  gpmr += gptr;
  gbtl += gpml * btl;
  gbul += gpml * bul;
  gbtr += gpmr * btr;
  gbur += gpmr * bur;
  gvnl += T(2.0) * vnl * gvl2;
  gvtl += T(2.0) * vtl * gvl2;
  gvul += T(2.0) * vul * gvl2;
  gvnr += T(2.0) * vnr * gvr2;
  gvtr += T(2.0) * vtr * gvr2;
  gvur += T(2.0) * vur * gvr2;
  grol += -groli * roli * roli;
  gror += -grori * rori * rori;
  gbnc += sgn * gbnca + T(2.0) * bnc * gbnc2;
  gL[0] += grol; gL[1] += gvnl; gL[2] += gvtl; gL[3] += gvul; gL[4] += gbtl; gL[5] += gbul; gL[6] += gprl;
  gR[0] += gror; gR[1] += gvnr; gR[2] += gvtr; gR[3] += gvur; gR[4] += gbtr; gR[5] += gbur; gR[6] += gprr;
  gbn += gbnc;
}

}  // namespace qas
