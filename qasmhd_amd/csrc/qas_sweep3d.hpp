// qas_sweep3d.hpp -- the 3D face sweep kernel, templated on the precision T
// of the reconstruction and Riemann solver (double: FP64 path; float: mixed
// path).  Included by qas_kernels3d.hip and qas_mixed.hip.
#pragma once
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_sweep2d.hpp"

namespace qas {

// dv of axis A (0 x, 1 y, 2 z) at (i, j, k) after its bound()
// (mhd3d_solve.cpp:138-164): flags (-dnxs[1], dnys[1], dnzs[1]) for dvx,
// (dnxs[2], -dnys[2], dnzs[2]) for dvy, (dnxs[3], dnys[3], -dnzs[3]) for dvz.
template <int A, typename VT>
__device__ __forceinline__ double dv3(const VT* v, int i, int j, int k, const Geo& g, const BcFlags& bc)
{
  const int dnx = A == 0 ? -bc.dn[0][1] : bc.dn[0][1 + A];
  const int dny = A == 1 ? -bc.dn[1][2] : bc.dn[1][1 + A];
  const int dnz = A == 2 ? -bc.dn[2][3] : bc.dn[2][1 + A];
  const Src sx = resolve(i, g.nx, 0, dnx);
  const Src sy = resolve(j, g.ny, 0, dny);
  const Src sz = resolve(k, g.nz, 0, dnz);
  const long long q = g.sz * sz.i + g.sy * sy.i + sx.i;
  const long long o = A == 0 ? 1 : (A == 1 ? g.sy : g.sz);
  const VT r = tmin((VT)(v[q + o] - v[q]), (VT)(v[q] - v[q - o]));
  return apply(sz, apply(sy, apply(sx, (double)r)));
}

template <int A, typename VT>
__device__ __forceinline__ double dv3_face(const PrimView<VT>& pv, int i, int j, int k, const Geo& g,
                                           const BcFlags& bc)
{
  const VT* v = A == 0 ? pv.vx : (A == 1 ? pv.vy : pv.vz);
  return dv3<A>(v, i, j, k, g, bc);
}

// the same from the state (velocity m (1/ro) rounded to VT, the stored
// primitive's value), for the sweep that stores the primitives
template <int A, typename VT>
__device__ __forceinline__ double dv3_state(const StagePtrs& p, int i, int j, int k, const Geo& g, const BcFlags& bc)
{
  const VelOf<VT> v{p.in[1 + A], p.in[0]};
  const int dnx = A == 0 ? -bc.dn[0][1] : bc.dn[0][1 + A];
  const int dny = A == 1 ? -bc.dn[1][2] : bc.dn[1][1 + A];
  const int dnz = A == 2 ? -bc.dn[2][3] : bc.dn[2][1 + A];
  const Src sx = resolve(i, g.nx, 0, dnx);
  const Src sy = resolve(j, g.ny, 0, dny);
  const Src sz = resolve(k, g.nz, 0, dnz);
  const long long q = g.sz * sz.i + g.sy * sy.i + sx.i;
  const long long o = A == 0 ? 1 : (A == 1 ? g.sy : g.sz);
  const VT r = tmin((VT)(v[q + o] - v[q]), (VT)(v[q] - v[q - o]));
  return apply(sz, apply(sy, apply(sx, (double)r)));
}

// Cell-centred B (fcen along x, y, z on the staggered components, then the
// bound() index map at ghost cells) and primitives of cell (i, j, k) = flat
// index t from the stage input state (mhd_class.hpp:80-91, mhd_func.c:6-16;
// mhd3d_solve.cpp:49-98), FP64 with the reference's operation order in every
// precision (no FMA contraction here).  INNER: the caller knows the cell is
// no ghost cell, where the index map is the identity.
struct Prim3d {
  double vx, vy, vz, pr, c[3];
};

template <int ODR, bool INNER>
__device__ __forceinline__ Prim3d prim3d_cell(const StagePtrs& p, const Geo& g, const BcFlags& bc, double gam, int i,
                                              int j, int k, long long t)
{
#pragma clang fp contract(off)
  Prim3d r;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double* b = p.in[4 + a];
    const long long o = a == 0 ? 1 : (a == 1 ? g.sy : g.sz);
    if constexpr (INNER) {
      r.c[a] = fcen<ODR>(b[t - o], b[t], b[t + o], b[t + 2 * o]);
    } else {
      const Src sx = resolve(i, g.nx, 0, bc.dn[0][4 + a]);
      const Src sy = resolve(j, g.ny, 0, bc.dn[1][4 + a]);
      const Src sz = resolve(k, g.nz, 0, bc.dn[2][4 + a]);
      const long long q = g.sz * sz.i + g.sy * sy.i + sx.i;
      const double v = fcen<ODR>(b[q - o], b[q], b[q + o], b[q + 2 * o]);
      r.c[a] = apply(sz, apply(sy, apply(sx, v)));
    }
  }
  const double ro = p.in[0][t];
  const double iro = 1.0 / ro;
  r.vx = p.in[1][t] * iro;
  r.vy = p.in[2][t] * iro;
  r.vz = p.in[3][t] * iro;
  const double v2 = r.vx * r.vx + r.vy * r.vy + r.vz * r.vz;
  const double b2 = r.c[0] * r.c[0] + r.c[1] * r.c[1] + r.c[2] * r.c[2];
  r.pr = (gam - 1) * (p.in[7][t] - 0.5 * (ro * v2 + b2));
  return r;
}

// ------------------------------------------------------------------ y, z sweeps
// LDS-staged as the 2D y sweep (qas_sweep2d.hpp): one workgroup = 4
// wavefronts = 4 adjacent x-columns of one line segment along DIR at one
// position of the other transverse axis; stencil inputs loaded
// cooperatively (4 consecutive doubles of a row per 4 threads) into LDS, the
// per-lane computation of k_sweep3d run from LDS, outputs staged in LDS and
// written back (and accumulated / combined into the RK stage sum) by
// coalesced threads.  Same arithmetic: bit-identical to k_sweep3d.
template <int DIR, int RMN, int ODR, typename T>
__device__ __forceinline__ void sweep3d_t(const StagePtrs& p, const Geo& g, const BcFlags& bc, double gam,
                                          int nseg)
{
  static_assert(DIR == 1 || DIR == 2, "transverse sweeps only");
  constexpr int TA = (DIR + 1) % 3;
  constexpr int TB = (DIR + 2) % 3;
  constexpr int LB = DIR == 2 ? 1 : 2;  // the other transverse axis (x is the column axis)
  __shared__ double sh[9][YC][YRP];      // ro vn vt vu bt bu pr cn bf
  __shared__ double so[9][YC][WAVE];     // du[5], 4 face parts
  const int dims[3] = {g.nx, g.ny, g.nz};
  const long long strd[3] = {1, g.sy, g.sz};
  const int ncg = (g.nx + YC - 1) / YC;
  const int nlb = dims[LB];
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  if ((long long)L >= (long long)ncg * nseg * nlb) return;  // whole workgroup
  const int cg = L % ncg;
  const int rest = L / ncg;
  const int seg = rest % nseg, lb = rest / nseg;
  const int i0 = cg * YC;
  const int f0 = 3 + seg * SEG;
  const int n = dims[DIR];
  const long long o = strd[DIR];
  const long long blb = strd[LB] * lb;
  const PrimView<T> pv = prim_view<T>(p);
  const T* vax[3] = {pv.vx, pv.vy, pv.vz};
  const T* cax[3] = {pv.cx, pv.cy, pv.cz};
  if constexpr (DIR == 2) {
    // The z sweep runs first: its staging computes the primitives of the
    // tile's cells from the stage input state and rounds them once to T;
    // the tile's own cells (its x-columns at this y, the z-rows of its
    // segment, the first segment from row 0 and the last to row n-1) are
    // stored for the y and x sweeps, so every cell is written once.
    const int lo = seg == 0 ? 0 : f0;
    const int hi = seg == nseg - 1 ? n : f0 + SEG;
    const bool inner = (i0 >= NG) && (i0 + YC - 1 <= g.nx - NG - 1) && (f0 - 4 >= NG) &&
                       (f0 + YR - 5 <= n - NG - 1) && (lb >= NG) && (lb <= nlb - NG - 1);
    for (int e = threadIdx.x; e < YC * YR; e += blockDim.x) {
      const int c = e % YC, r = e / YC;
      const int rowu = f0 - 4 + r;
      const int row = rowu < 0 ? 0 : (rowu > n - 1 ? n - 1 : rowu);
      const int colu = i0 + c;
      const int col = colu > g.nx - 1 ? g.nx - 1 : colu;
      const long long a = blb + o * row + col;
      const Prim3d q = inner ? prim3d_cell<ODR, true>(p, g, bc, gam, col, lb, row, a)
                             : prim3d_cell<ODR, false>(p, g, bc, gam, col, lb, row, a);
      const T v3[3] = {(T)q.vx, (T)q.vy, (T)q.vz};
      const T c3[3] = {(T)q.c[0], (T)q.c[1], (T)q.c[2]};
      const T pr = (T)q.pr;
      sh[0][c][r] = p.in[0][a];
      sh[1][c][r] = (double)v3[DIR];
      sh[2][c][r] = (double)v3[TA];
      sh[3][c][r] = (double)v3[TB];
      sh[4][c][r] = (double)c3[TA];
      sh[5][c][r] = (double)c3[TB];
      sh[6][c][r] = (double)pr;
      sh[7][c][r] = (double)c3[DIR];
      sh[8][c][r] = p.in[4 + DIR][a];
      if (colu < g.nx && rowu >= lo && rowu < hi) {
        reinterpret_cast<T*>(p.vx)[a] = v3[0];
        reinterpret_cast<T*>(p.vy)[a] = v3[1];
        reinterpret_cast<T*>(p.vz)[a] = v3[2];
        reinterpret_cast<T*>(p.pr)[a] = pr;
        reinterpret_cast<T*>(p.cx)[a] = c3[0];
        reinterpret_cast<T*>(p.cy)[a] = c3[1];
        reinterpret_cast<T*>(p.cz)[a] = c3[2];
      }
    }
  } else {
    const T* src[6] = {vax[DIR], vax[TA], vax[TB], cax[TA], cax[TB], pv.pr};
    for (int e = threadIdx.x; e < YC * YR; e += blockDim.x) {
      const int c = e % YC, r = e / YC;
      int row = f0 - 4 + r;
      row = row < 0 ? 0 : (row > n - 1 ? n - 1 : row);
      int col = i0 + c;
      col = col > g.nx - 1 ? g.nx - 1 : col;
      const long long a = blb + o * row + col;
      sh[0][c][r] = p.in[0][a];
#pragma unroll
      for (int k = 0; k < 6; k++) sh[1 + k][c][r] = (double)src[k][a];
      sh[7][c][r] = (double)cax[DIR][a];
      sh[8][c][r] = p.in[4 + DIR][a];
    }
  }
  __syncthreads();
  const int w = threadIdx.x / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const int la = i0 + w;  // this wavefront's x-column (>= nx: computes, never stores)
  const int q = f0 - 2 + lane;
  const int qc = q < 2 ? 2 : (q > n - 3 ? n - 3 : q);
  const int rb = qc - (f0 - 4);
  T s[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int r = rb + m - 2;
#pragma unroll
    for (int k = 0; k < 7; k++) s[k][m] = (T)sh[k][w][r];
  }
  // B-flux face states for both tangential components (mhd_lr_fb)
  T qlt, qrt, qlu, qru;
  {
    const T bnm = (T)sh[8][w][rb], bnp = (T)sh[8][w][rb + 1];
    T pp[5], qq[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      pp[m] = s[4][m] * s[1][m];
      qq[m] = s[2][m];
    }
    lr_fb<ODR>(pp, qq, bnm, bnp, qlt, qrt);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      pp[m] = s[5][m] * s[1][m];
      qq[m] = s[3][m];
    }
    lr_fb<ODR>(pp, qq, bnm, bnp, qlu, qru);
  }
  T vl[7], vr[7];
  lrstate<ODR>(s, (T)sh[7][w][rb], (T)gam, vl, vr);
  T Lf[7], F[7];
#pragma unroll
  for (int k = 0; k < 7; k++) Lf[k] = __shfl_up(vl[k], 1, WAVE);
  const T qlt_f = __shfl_up(qlt, 1, WAVE);
  const T qlu_f = __shfl_up(qlu, 1, WAVE);
  const int qf = q < 3 ? 3 : (q > n - 3 ? n - 3 : q);
  const int rf = qf - (f0 - 4);
  const T bnf = (T)(0.5 * (sh[8][w][rf] + sh[8][w][rf]));
  T dv0 = T(0.0), dv1 = T(0.0);
  if constexpr (RMN >= 2) {
    dv0 = (T)(sh[1][w][rf] - sh[1][w][rf - 1]);
    const int lac = la > g.nx - 1 ? g.nx - 1 : la;
    int c0[3], c1[3];
    c0[0] = lac; c0[LB] = lb; c0[DIR] = qf - 1;
    c1[0] = lac; c1[LB] = lb; c1[DIR] = qf;
    double a0, a1, b0, b1;
    if constexpr (DIR == 2) {
      // the velocities of the neighbouring columns are being stored by other
      // workgroups of this launch: from the state instead
      a0 = dv3_state<TA, T>(p, c0[0], c0[1], c0[2], g, bc);
      a1 = dv3_state<TA, T>(p, c1[0], c1[1], c1[2], g, bc);
      b0 = dv3_state<TB, T>(p, c0[0], c0[1], c0[2], g, bc);
      b1 = dv3_state<TB, T>(p, c1[0], c1[1], c1[2], g, bc);
    } else {
      a0 = dv3_face<TA>(pv, c0[0], c0[1], c0[2], g, bc);
      a1 = dv3_face<TA>(pv, c1[0], c1[1], c1[2], g, bc);
      b0 = dv3_face<TB>(pv, c0[0], c0[1], c0[2], g, bc);
      b1 = dv3_face<TB>(pv, c1[0], c1[1], c1[2], g, bc);
    }
    dv1 = (T)tmin(tmin(a0, a1), tmin(b0, b1));
  }
  riemann<RMN>(Lf, vr, bnf, (T)gam, dv0, dv1, F);
  const T fct = T(0.5) * (qlt_f + qrt);
  const T fcu = T(0.5) * (qlu_f + qru);
  so[5][w][lane] = (double)(F[4] - fct);
  so[6][w][lane] = (double)fct;
  so[7][w][lane] = (double)(F[5] - fcu);
  so[8][w][lane] = (double)fcu;
  double fv[5];
  fv[0] = F[0];
  fv[1 + DIR] = F[1];
  fv[1 + TA] = F[2];
  fv[1 + TB] = F[3];
  fv[4] = F[6];
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const double fm1 = __shfl_up(fv[v], 1, WAVE);
    const double f1 = __shfl_down(fv[v], 1, WAVE);
    const double f2 = __shfl_down(fv[v], 2, WAVE);
    so[v][w][lane] = dfc<ODR>(fm1, fv[v], f1, f2);
  }
  __syncthreads();
  {
    const int c = threadIdx.x % YC, r = threadIdx.x / YC;
    const int qq = f0 - 2 + r;
    const int i = i0 + c;
    const bool ownr = (r >= 2) && (r < 2 + SEG) && (qq <= n - 3) && (i < g.nx);
    if (ownr) {
      const long long a = blb + o * qq + i;
#pragma unroll
      for (int k = 0; k < 4; k++) store_as<T>(p.emf[DIR * 4 + k], a, so[5 + k][c][r]);
      const bool upd = (qq >= NG) && (qq <= n - NG - 1) && (i >= NG) && (i <= g.nx - NG - 1) && (lb >= NG) &&
                       (lb <= nlb - NG - 1);
      if (upd) {
        // dt/dy Dy (slots 0-4) or dt/dz Dz (slots 5-9) of (ro, mx, my, mz,
        // en); the x sweep, which runs last, subtracts both
        const double dtd = *p.dt / (DIR == 1 ? g.dy : g.dz);
        double* const* acc = p.acc + (DIR == 1 ? 0 : 5);
#pragma unroll
        for (int v = 0; v < 5; v++) store_as<T>(acc[v], a, dtd * so[v][c][r]);
      }
    }
  }
}

// FMA: build tag of the instantiation (see k_sweep2d_y)
template <int DIR, int RMN, int ODR, typename T, int FMA = 0>
__global__ __launch_bounds__(256, 2) void k_sweep3d_t(StagePtrs p, Geo g, BcFlags bc, double gam, int nseg)
{
  sweep3d_t<DIR, RMN, ODR, T>(p, g, bc, gam, nseg);
}

// ------------------------------------------------------------------ x sweep
// One wavefront per 60-face segment of an x line; runs after the z and y
// sweeps and applies the RK stage sum.
template <int DIR, int RMN, int ODR, typename T, int FMA = 0>
__global__ __launch_bounds__(256) void k_sweep3d(StagePtrs p, Geo g, BcFlags bc, double gam, int nseg)
{
  constexpr int TA = (DIR + 1) % 3;  // first transverse axis (t)
  constexpr int TB = (DIR + 2) % 3;  // second transverse axis (u)
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
  const int dims[3] = {g.nx, g.ny, g.nz};
  const long long strd[3] = {1, g.sy, g.sz};
  // lines: x -> (j, k); y -> (i, k); z -> (i, j); the faster line index is
  // the lower axis for coalescing
  constexpr int LA = DIR == 0 ? 1 : 0;  // faster transverse axis
  constexpr int LB = DIR == 2 ? 1 : 2;  // slower transverse axis
  const int nla = dims[LA], nlb = dims[LB];
  const long long nline = (long long)nla * nlb;
  if ((long long)wave >= nline * nseg) return;
  const long long line = wave / nseg;
  const int seg = (int)(wave - line * nseg);
  const int la = (int)(line % nla), lb = (int)(line / nla);
  const int n = dims[DIR];
  const long long o = strd[DIR];
  const long long base = strd[LA] * la + strd[LB] * lb;
  const int f0 = 3 + seg * SEG;
  const int q = f0 - 2 + lane;
  const int qc = q < 2 ? 2 : (q > n - 3 ? n - 3 : q);
  const long long sc = base + o * qc;

  const PrimView<T> pv = prim_view<T>(p);
  const T* vax[3] = {pv.vx, pv.vy, pv.vz};
  const T* cax[3] = {pv.cx, pv.cy, pv.cz};
  const T* vn = vax[DIR];
  const T* vt = vax[TA];
  const T* vu = vax[TB];
  const T* bt = cax[TA];
  const T* bu = cax[TB];
  const double* bf = p.in[4 + DIR];
  T s[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const long long a = sc + (m - 2) * o;
    s[0][m] = (T)p.in[0][a];
    s[1][m] = vn[a];
    s[2][m] = vt[a];
    s[3][m] = vu[a];
    s[4][m] = bt[a];
    s[5][m] = bu[a];
    s[6][m] = pv.pr[a];
  }
  T vl[7], vr[7];
  lrstate<ODR>(s, (T)cax[DIR][sc], (T)gam, vl, vr);

  // B-flux face states for both tangential components (mhd_lr_fb)
  T qlt, qrt, qlu, qru;
  {
    T pp[5], qq[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const long long a = sc + (m - 2) * o;
      pp[m] = bt[a] * vn[a];
      qq[m] = vt[a];
    }
    lr_fb<ODR>(pp, qq, (T)bf[sc], (T)bf[sc + o], qlt, qrt);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const long long a = sc + (m - 2) * o;
      pp[m] = bu[a] * vn[a];
      qq[m] = vu[a];
    }
    lr_fb<ODR>(pp, qq, (T)bf[sc], (T)bf[sc + o], qlu, qru);
  }

  T L[7], F[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = __shfl_up(vl[k], 1, WAVE);
  const T qlt_f = __shfl_up(qlt, 1, WAVE);
  const T qlu_f = __shfl_up(qlu, 1, WAVE);
  const int qf = q < 3 ? 3 : (q > n - 3 ? n - 3 : q);
  const long long sf = base + o * qf;
  const T bnf = (T)(0.5 * (bf[sf] + bf[sf]));
  T dv0 = T(0.0), dv1 = T(0.0);
  if constexpr (RMN >= 2) {
    dv0 = (T)(vn[sf] - vn[sf - o]);
    int c0[3], c1[3];
    c0[LA] = la; c0[LB] = lb; c0[DIR] = qf - 1;
    c1[LA] = la; c1[LB] = lb; c1[DIR] = qf;
    double a0, a1, b0, b1;
    if constexpr (DIR == 2) {
      // the velocities of the neighbouring columns are being stored by other
      // workgroups of this launch: from the state instead
      a0 = dv3_state<TA, T>(p, c0[0], c0[1], c0[2], g, bc);
      a1 = dv3_state<TA, T>(p, c1[0], c1[1], c1[2], g, bc);
      b0 = dv3_state<TB, T>(p, c0[0], c0[1], c0[2], g, bc);
      b1 = dv3_state<TB, T>(p, c1[0], c1[1], c1[2], g, bc);
    } else {
      a0 = dv3_face<TA>(pv, c0[0], c0[1], c0[2], g, bc);
      a1 = dv3_face<TA>(pv, c1[0], c1[1], c1[2], g, bc);
      b0 = dv3_face<TB>(pv, c0[0], c0[1], c0[2], g, bc);
      b1 = dv3_face<TB>(pv, c1[0], c1[1], c1[2], g, bc);
    }
    dv1 = (T)tmin(tmin(a0, a1), tmin(b0, b1));
  }
  riemann<RMN>(L, vr, bnf, (T)gam, dv0, dv1, F);
  const T fct = T(0.5) * (qlt_f + qrt);
  const T fcu = T(0.5) * (qlu_f + qru);
  const bool own = (lane >= 2) && (lane < 2 + SEG) && (q <= n - 3);
  if (own) {
    const long long fi = base + o * q;
    store_as<T>(p.emf[DIR * 4 + 0], fi, (double)(F[4] - fct));  // B_t flux, upwind part
    store_as<T>(p.emf[DIR * 4 + 1], fi, (double)fct);           // B_t flux, central part
    store_as<T>(p.emf[DIR * 4 + 2], fi, (double)(F[5] - fcu));  // B_u flux, upwind part
    store_as<T>(p.emf[DIR * 4 + 3], fi, (double)fcu);           // B_u flux, central part
  }
  // divergence of ro, mx, my, mz, en: (ro, mn, mt, mu) permuted back
  double fv[5];
  fv[0] = F[0];
  fv[1 + DIR] = F[1];
  fv[1 + TA] = F[2];
  fv[1 + TB] = F[3];
  fv[4] = F[6];
  double du[5];
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const double fm1 = __shfl_up(fv[v], 1, WAVE);
    const double f1 = __shfl_down(fv[v], 1, WAVE);
    const double f2 = __shfl_down(fv[v], 2, WAVE);
    du[v] = dfc<ODR>(fm1, fv[v], f1, f2);
  }
  const bool upd = own && (q >= NG) && (q <= n - NG - 1) && (la >= NG) && (la <= nla - NG - 1) && (lb >= NG) &&
                   (lb <= nlb - NG - 1);
  if (!upd) return;
  // the z and y sweeps ran first (dt/dz Dz in slots 5-9, dt/dy Dy in 0-4);
  // the stage sum u = a u0 + b (u + (((-dt/dx Dx) - dt/dy Dy) - dt/dz Dz))
  // is the reference's expression in its order (mhd3d_solve.cpp:658-676)
  static_assert(DIR == 0, "the per-lane sweep applies the stage sum along x");
  const long long c = base + o * q;
  const double dtdx = *p.dt / g.dx;
  const int vid[5] = {0, 1, 2, 3, 7};
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const int m = vid[v];
    p.out[m][c] = rk_upd(p.in[m][c], p.u0[m][c], ((-dtdx * du[v]) - load_as<T>(p.acc[v], c)) - load_as<T>(p.acc[5 + v], c),
                         p.a, p.b);
  }
}

}  // namespace qas
