// qas_sweep2d.hpp -- the 2D face sweep kernel, templated on the precision T
// of the reconstruction and Riemann solver (double: the FP64 path, bit-
// identical to the reference; float: the mixed path).  Included by
// qas_kernels2d.hip (T = double, -ffp-contract=off) and qas_mixed.hip
// (T = float).  Lane layout: one 64-lane wavefront per 64 consecutive cells
// of a line; lane l reconstructs cell f0-2+l, neighbour states and fluxes
// move between lanes by cross-lane shuffles, and the wavefront owns the 60
// faces/cells [f0, f0+60).
#pragma once
#include "qas_batch.hpp"
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_stage2d.hpp"

namespace qas {

// velocity component m (1/ro) of a cell from the state, rounded to T as
// the stored primitive is
template <typename T>
struct VelOf {
  const double* m;
  const double* ro;
  __device__ __forceinline__ T operator[](long long q) const { return (T)(m[q] * (1.0 / ro[q])); }
};

// shock sensor dv of mhd2d_solve.cpp:105-125 at (i, j) after its bound():
// DIR 0 -> dvx (flags -dnxs[1], dnys[1]); DIR 1 -> dvy (dnxs[2], -dnys[2]).
// v: the velocity array, or a VelOf<VT> reading it from the state.
template <int DIR, typename V>
__device__ __forceinline__ double dv_at(const V& v, int i, int j, const Geo& g, const BcFlags& bc)
{
  using VT = typename std::remove_cv<typename std::remove_reference<decltype(v[0])>::type>::type;
  const int dnx = DIR == 0 ? -bc.dn[0][1] : bc.dn[0][2];
  const int dny = DIR == 0 ? bc.dn[1][1] : -bc.dn[1][2];
  const Src sx = resolve(i, g.nx, 0, dnx);
  const Src sy = resolve(j, g.ny, 0, dny);
  const long long q = (long long)g.nx * sy.i + sx.i;
  const long long o = DIR == 0 ? 1 : g.nx;
  const VT r = tmin((VT)(v[q + o] - v[q]), (VT)(v[q] - v[q - o]));
  return apply(sy, apply(sx, (double)r));
}

// y sweep, LDS-staged; it runs before the x sweep of the stage.  One
// workgroup = 4 wavefronts = 4 adjacent columns of one 60-face segment.  The stencil inputs of the 68 rows f0-4..f0+63 of
// the 4 columns are loaded cooperatively (4 consecutive threads read 4
// consecutive doubles of a row) into LDS, column-major; each wavefront then
// runs the same per-lane computation as the x sweep on its column from LDS;
// the outputs (dt/dy times the y divergence of 6 variables, 2 face parts)
// go back through LDS and are written by coalesced threads.  Same
// arithmetic as a direct per-lane sweep: bit-identical.
constexpr int YC = 4;          // columns per workgroup
constexpr int YR = WAVE + 4;   // stencil rows per segment
constexpr int YRP = YR + 1;    // padded LDS column length

//
// One kernel serves a single domain (BATCH false) and a batch of domains
// (BATCH true), as the x sweep below; with the body as a device function
// called from thin kernels the FP64 HLLD / ODR 4 instance compiled to 250
// VGPRs (2 waves per SIMD) instead of 163 (3 waves).
// FMA: a tag (0 or 1) naming the build of the instantiation: the FP64
// sweeps built with FMA contraction (qas_fma64.hip) are the same template
// under FMA = 1, so their kernel symbols differ from the bit-exact ones.
// FP32 (mixed) LHLLD and MLAU instances may take up to 170 VGPRs (3 waves
// per SIMD): held to 128 with the FP64 primitives of the staging they
// spilled 36-72 bytes per lane (the MLAU y sweep at 2048^2: 314 us per
// launch against 272 us at 3 waves).  With the wavefront-uniform skip of
// the fast-only reconstruction terms they compile to 128 VGPRs without a
// spill (4 waves per SIMD): 276.6 -> about 265 us.
template <int RMN, int ODR, typename T, bool BATCH, bool GRAV = false, int FMA = 0>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? (RMN >= 2 ? 3 : 4) : (RMN != 1 ? 2 : 1)) void k_sweep2d_y(
    StagePtrs p_arg, Geo g_arg, BcFlags bc_arg, double gam, int nseg_arg, const BatchInst* I, const int* off,
    int ninst, int stage)
{
  StagePtrs p;
  Geo g;
  BcFlags bc;
  int nseg, L;
  if constexpr (BATCH) {
    const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int i = batch_find(off, ninst, b);
    p = I[i].st[stage];
    g = I[i].g;
    bc = I[i].bc;
    nseg = I[i].nsy;
    L = b - off[i];
  } else {
    p = p_arg;
    g = g_arg;
    bc = bc_arg;
    nseg = nseg_arg;
    L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  }
  __shared__ double sh[9][YC][YRP];  // ro vy vz vx bz cx pr cy by
  __shared__ double so[8][YC][WAVE];  // du[6], upwind and central B-flux parts
  const int ncg = (g.nx + YC - 1) / YC;
  if (L >= ncg * nseg) return;  // whole workgroup
  const int cg = L % ncg, seg = L / ncg;
  const int i0 = cg * YC;
  const int f0 = 3 + seg * SEG;
  const int n = g.ny;
  const long long sy = g.sy;
  {
    // Staging computes the primitives of the tile's cells from the stage
    // input state (prim2d_cell, FP64) and rounds them once to the sweep
    // precision T.  The tile's own cells -- its columns, and the rows of
    // its segment (the first segment from row 0, the last to row n-1) --
    // are also stored for the x sweep, so every cell of the domain is
    // written by exactly one workgroup.
    const int lo = seg == 0 ? 0 : f0;
    const int hi = seg == nseg - 1 ? n : f0 + SEG;
    // a tile of interior cells only skips the ghost-index map (uniform)
    const bool inner = (i0 >= NG) && (i0 + YC - 1 <= g.nx - NG - 1) && (f0 - 4 >= NG) && (f0 + YR - 5 <= n - NG - 1);
    for (int e = threadIdx.x; e < YC * YR; e += blockDim.x) {
      const int c = e % YC, r = e / YC;
      const int rowu = f0 - 4 + r;
      const int row = rowu < 0 ? 0 : (rowu > n - 1 ? n - 1 : rowu);
      const int colu = i0 + c;
      const int col = colu > g.nx - 1 ? g.nx - 1 : colu;
      const long long a = sy * row + col;
      const Prim2d q = inner ? prim2d_cell<ODR, GRAV, true>(p, g, bc, gam, col, row, a)
                             : prim2d_cell<ODR, GRAV, false>(p, g, bc, gam, col, row, a);
      const T vx = (T)q.vx, vy = (T)q.vy, vz = (T)q.vz, pr = (T)q.pr, cx = (T)q.cx, cy = (T)q.cy;
      sh[0][c][r] = G(p.in[0])[a];
      sh[1][c][r] = (double)vy;
      sh[2][c][r] = (double)vz;
      sh[3][c][r] = (double)vx;
      sh[4][c][r] = G(p.in[6])[a];
      sh[5][c][r] = (double)cx;
      sh[6][c][r] = (double)pr;
      sh[7][c][r] = (double)cy;
      sh[8][c][r] = G(p.in[5])[a];
      if (colu < g.nx && rowu >= lo && rowu < hi) {
        G(reinterpret_cast<T*>(p.vx))[a] = vx;
        G(reinterpret_cast<T*>(p.vy))[a] = vy;
        G(reinterpret_cast<T*>(p.vz))[a] = vz;
        G(reinterpret_cast<T*>(p.pr))[a] = pr;
        G(reinterpret_cast<T*>(p.cx))[a] = cx;
        G(reinterpret_cast<T*>(p.cy))[a] = cy;
        if constexpr (GRAV) p.enrt[a] = q.enrt;
      }
    }
  }
  __syncthreads();
  const int w = threadIdx.x / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const int line = i0 + w;  // this wavefront's column (>= nx: computes, never stores)
  const int q = f0 - 2 + lane;
  const int qc = q < 2 ? 2 : (q > n - 3 ? n - 3 : q);
  const int rb = qc - (f0 - 4);
  T s[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int r = rb + m - 2;
    s[0][m] = (T)sh[0][w][r];
    s[1][m] = (T)sh[1][w][r];
    s[2][m] = (T)sh[2][w][r];
    s[3][m] = (T)sh[3][w][r];
    s[4][m] = (T)sh[4][w][r];
    s[5][m] = (T)sh[5][w][r];
    s[6][m] = (T)sh[6][w][r];
  }
  // B-flux face states first (their inputs are already in s[]), then the
  // characteristic reconstruction
  T ql, qr;
  {
    T pp[5], qq[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      pp[m] = s[5][m] * s[1][m];  // cx * vy
      qq[m] = s[3][m];            // vx
    }
    lr_fb<ODR>(pp, qq, (T)sh[8][w][rb], (T)sh[8][w][rb + 1], ql, qr);
  }
  T vl[7], vr[7];
  // (FP32 LHLLD / MLAU: the wavefront-uniform skip of the fast-only terms
  // brings the kernel to 128 VGPRs, 4 waves per SIMD)
  lrstate<ODR, T, 0, true, (sizeof(T) == 8 || RMN >= 2)>(s, (T)sh[7][w][rb], (T)gam, vl, vr);
  T Lf[7], Rf[7], F[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    Lf[k] = __shfl_up(vl[k], 1, WAVE);
    Rf[k] = vr[k];
  }
  const T qlf = __shfl_up(ql, 1, WAVE);
  const int sf = q < 3 ? 3 : (q > n - 3 ? n - 3 : q);
  const int rf = sf - (f0 - 4);
  const T bnf = (T)(0.5 * (sh[8][w][rf] + sh[8][w][rf]));
  T dv0 = T(0.0), dv1 = T(0.0);
  if constexpr (RMN >= 2) {
    // the transverse sensor needs vx of the neighbouring columns, which
    // other workgroups of this launch store: taken from the state instead,
    // with the primitive's own expression and rounding
    const int lc = line > g.nx - 1 ? g.nx - 1 : line;
    dv0 = (T)(sh[1][w][rf] - sh[1][w][rf - 1]);
    const VelOf<T> vxs{p.in[1], p.in[0]};
    dv1 = (T)tmin(dv_at<0>(vxs, lc, sf - 1, g, bc), dv_at<0>(vxs, lc, sf, g, bc));
  }
  riemann<RMN>(Lf, Rf, bnf, (T)gam, dv0, dv1, F);
  const T fc = T(0.5) * (qlf + qr);
  const T fbu = F[5] - fc;
  const bool own = (lane >= 2) && (lane < 2 + SEG) && (q <= n - 3);
  // y: (F0, F3, F1, F2, F4, F6) feed (ro, mx, my, mz, bz, en)
  double fv[6];
  fv[0] = F[0]; fv[1] = F[3]; fv[2] = F[1]; fv[3] = F[2]; fv[4] = F[4]; fv[5] = F[6];
  if constexpr (GRAV) {  // work done by gravity (2D/SERIAL/RTI/mhd2d_solve.cpp:305)
    const long long lc = line > g.nx - 1 ? g.nx - 1 : line;
    fv[5] += F[0] * 0.5 * (p.phi[lc + sy * (sf - 1)] + p.phi[lc + sy * sf]);
  }
#pragma unroll
  for (int v = 0; v < 6; v++) {
    const double fm1 = __shfl_up(fv[v], 1, WAVE);
    const double f1 = __shfl_down(fv[v], 1, WAVE);
    const double f2 = __shfl_down(fv[v], 2, WAVE);
    so[v][w][lane] = dfc<ODR>(fm1, fv[v], f1, f2);
  }
  so[6][w][lane] = (double)fbu;
  so[7][w][lane] = (double)fc;
  (void)own;
  __syncthreads();
  // coalesced write-back: thread -> (column c, lane r) of the 4 x 64 tile
  {
    const int c = threadIdx.x % YC, r = threadIdx.x / YC;
    const int qq = f0 - 2 + r;
    const int i = i0 + c;
    const bool ownr = (r >= 2) && (r < 2 + SEG) && (qq <= n - 3) && (i < g.nx);
    if (ownr) {
      const long long a = sy * qq + i;
      store_as<T>(p.emf[4], a, so[6][c][r]);
      store_as<T>(p.emf[5], a, so[7][c][r]);
      const bool upd = (qq >= NG) && (qq <= n - NG - 1) && (i >= NG) && (i <= g.nx - NG - 1);
      if (upd) {
        // dt/dy * Dy of (ro, mx, my, mz, bz, en); the x sweep, which runs
        // next, subtracts it in the stage sum
        const double dtdy = G(p.dt)[0] / g.dy;
#pragma unroll
        for (int v = 0; v < 6; v++) store_as<T>(p.acc[v], a, dtdy * so[v][c][r]);
      }
    }
  }
}

// ------------------------------------------------------------------ x sweep
// x faces for every row j (one wavefront per 60-face segment of a row); it
// runs after the y sweep and applies the RK stage sum.  One kernel serves a
// single domain (BATCH false: the stage descriptor is the kernel argument)
// and a batch of domains (BATCH true: the instance table, as kb_* kernels).
// The body is the kernel itself: as a force-inlined device function called
// from thin kernels it compiled to 246 VGPRs (2 waves per SIMD) instead of
// 158 (3 waves) for FP64 HLLD / ODR 4.  The FP64 Roe and MLAU instances ask
// for 2 waves per SIMD, which they otherwise miss by a few registers.
template <int DIR, int RMN, int ODR, typename T, bool BATCH, bool GRAV = false, int FMA = 0>
__global__ __launch_bounds__(256, (sizeof(T) == 8 && RMN != 1) ? 2 : 1) void k_sweep2d(StagePtrs p_arg, Geo g_arg, BcFlags bc_arg, double gam,
                                                 int nseg_arg, const BatchInst* I, const int* off, int ninst,
                                                 int stage)
{
  StagePtrs p;
  Geo g;
  BcFlags bc;
  int nseg, wave;
  if constexpr (BATCH) {
    const int b = (int)blockIdx.x;
    const int i = batch_find(off, ninst, b);
    p = I[i].st[stage];
    g = I[i].g;
    bc = I[i].bc;
    nseg = I[i].nsx;
    wave = (b - off[i]) * (int)(blockDim.x / WAVE) + (int)(threadIdx.x / WAVE);
  } else {
    p = p_arg;
    g = g_arg;
    bc = bc_arg;
    nseg = nseg_arg;
    wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
  }
  const int lane = threadIdx.x & (WAVE - 1);
  const int nline = DIR == 0 ? g.ny : g.nx;
  if (wave >= nseg * nline) return;  // whole wavefront exits together
  const int line = wave / nseg;
  const int seg = wave - line * nseg;
  const int n = DIR == 0 ? g.nx : g.ny;       // points along the sweep
  const long long o = DIR == 0 ? 1 : g.sy;    // stride along the sweep
  const long long base = DIR == 0 ? g.sy * line : (long long)line;
  const int f0 = 3 + seg * SEG;
  const int q = f0 - 2 + lane;                // this lane's cell
  const int qc = q < 2 ? 2 : (q > n - 3 ? n - 3 : q);
  const long long sc = base + o * qc;

  // primitive stencil, cyclic order (vn, vt, vu), (bt, bu)
  const PrimView<T> pv = prim_view<T>(p);
  const auto vn = G(DIR == 0 ? pv.vx : pv.vy);
  const auto vt = G(DIR == 0 ? pv.vy : pv.vz);
  const auto vu = G(DIR == 0 ? pv.vz : pv.vx);
  const auto cn = G(DIR == 0 ? pv.cx : pv.cy);
  const auto cb = G(DIR == 0 ? pv.cy : pv.cx);  // the cell-centred tangential B (the other one is bz)
  const auto bf = G(DIR == 0 ? p.in[4] : p.in[5]);  // staggered normal B
  const auto ro = G(p.in[0]);
  const auto bz = G(p.in[6]);
  const auto pr = G(pv.pr);
  T s[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const long long a = sc + (m - 2) * o;
    s[0][m] = (T)ro[a];
    s[1][m] = vn[a];
    s[2][m] = vt[a];
    s[3][m] = vu[a];
    s[4][m] = DIR == 0 ? cb[a] : (T)bz[a];
    s[5][m] = DIR == 0 ? (T)bz[a] : cb[a];
    s[6][m] = pr[a];
  }
  T vl[7], vr[7];
  lrstate<ODR>(s, (T)cn[sc], (T)gam, vl, vr);

  // B-flux face states (mhd_lr_fb): b_t v_n - b_n v_t with the tangential
  // component that pairs with this sweep (x: By, y: Bx)
  T ql, qr;
  {
    const auto vv = G(DIR == 0 ? pv.vy : pv.vx);
    T pp[5], qq[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const long long a = sc + (m - 2) * o;
      pp[m] = cb[a] * vn[a];
      qq[m] = vv[a];
    }
    lr_fb<ODR>(pp, qq, (T)bf[sc], (T)bf[sc + o], ql, qr);
  }

  // flux at face q (between q-1 and q): left state from lane-1
  T L[7], R[7], F[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    L[k] = __shfl_up(vl[k], 1, WAVE);
    R[k] = vr[k];
  }
  const T qlf = __shfl_up(ql, 1, WAVE);
  const long long sf = base + o * (q < 3 ? 3 : (q > n - 3 ? n - 3 : q));  // face index (clamped)
  const T bnf = (T)(0.5 * (bf[sf] + bf[sf]));
  T dv0 = T(0.0), dv1 = T(0.0);
  if constexpr (RMN >= 2) {
    const int qq = q < 3 ? 3 : (q > n - 3 ? n - 3 : q);
    dv0 = (T)(vn[sf] - vn[sf - o]);
    if (DIR == 0) {
      dv1 = (T)tmin(dv_at<1>(pv.vy, qq - 1, line, g, bc), dv_at<1>(pv.vy, qq, line, g, bc));
    } else {
      dv1 = (T)tmin(dv_at<0>(pv.vx, line, qq - 1, g, bc), dv_at<0>(pv.vx, line, qq, g, bc));
    }
  }
  riemann<RMN>(L, R, bnf, (T)gam, dv0, dv1, F);
  const T fc = T(0.5) * (qlf + qr);
  // split of the B flux: x sweep -> By flux (F[4]); y sweep -> Bx flux (F[5])
  const T fbu = (DIR == 0 ? F[4] : F[5]) - fc;

  const bool own = (lane >= 2) && (lane < 2 + SEG) && (q <= n - 3);
  if (own) {
    const long long fi = base + o * q;
    store_as<T>(p.emf[DIR * 4 + 0], fi, (double)fbu);
    store_as<T>(p.emf[DIR * 4 + 1], fi, (double)fc);
  }

  // flux divergence of ro, mx, my, mz, bz, en at cell q: faces q-1 .. q+2
  // x: (F0, F1, F2, F3, F5, F6); y: (F0, F3, F1, F2, F4, F6)
  // (fluxes are differenced in FP64 in both precisions)
  double fv[6];
  if (DIR == 0) {
    fv[0] = F[0]; fv[1] = F[1]; fv[2] = F[2]; fv[3] = F[3]; fv[4] = F[5]; fv[5] = F[6];
  } else {
    fv[0] = F[0]; fv[1] = F[3]; fv[2] = F[1]; fv[3] = F[2]; fv[4] = F[4]; fv[5] = F[6];
  }
  if constexpr (GRAV) fv[5] += F[0] * 0.5 * (p.phi[sf - o] + p.phi[sf]);  // work done by gravity (RTI:197)
  double du[6];
#pragma unroll
  for (int v = 0; v < 6; v++) {
    const double fm1 = __shfl_up(fv[v], 1, WAVE);
    const double f1 = __shfl_down(fv[v], 1, WAVE);
    const double f2 = __shfl_down(fv[v], 2, WAVE);
    du[v] = dfc<ODR>(fm1, fv[v], f1, f2);
  }
  const int tl = line;
  const int ntl = DIR == 0 ? g.ny : g.nx;
  const bool upd = own && (q >= NG) && (q <= n - NG - 1) && (tl >= NG) && (tl <= ntl - NG - 1);
  if (!upd) return;
  // the y sweep ran first and left acc = dt/dy * Dy; the stage sum
  // u = a u0 + b (u + ((-dt/dx Dx) - dt/dy Dy)) is the reference's
  // expression in its order (mhd2d_solve.cpp:199, 341-358)
  static_assert(DIR == 0, "the per-lane sweep applies the stage sum along x");
  const long long c = base + o * q;
  const double dtdx = G(p.dt)[0] / g.dx;
  const int vid[6] = {0, 1, 2, 3, 6, 7};
  double un[6];
#pragma unroll
  for (int v = 0; v < 6; v++) {
    const int m = vid[v];
    un[v] = rk_upd(G(p.in[m])[c], G(p.u0[m])[c], (-dtdx * du[v]) - load_as<T>(p.acc[v], c), p.a, p.b);
  }
  if constexpr (GRAV) {
    // gravity on the momentum with the updated density (RTI/mhd2d_solve.cpp:361-362)
    const double dtdy = *p.dt / g.dy;
    un[1] += -p.b * un[0] * 0.5 * (p.phi[c + 1] - p.phi[c - 1]) * dtdx;
    un[2] += -p.b * un[0] * 0.5 * (p.phi[c + g.sy] - p.phi[c - g.sy]) * dtdy;
  }
#pragma unroll
  for (int v = 0; v < 6; v++) G(p.out[vid[v]])[c] = un[v];
}

// kernel tables: y sweep and x sweep (single domain or batch)
template <typename T, int FMA = 0>
xsweep2d_fn pick_ysweep2d_t(int rmn, int odr, bool batch)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return batch ? k_sweep2d_y<R, O, T, true, false, FMA> : k_sweep2d_y<R, O, T, false, false, FMA>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

template <typename T, int FMA = 0>
xsweep2d_fn pick_xsweep2d_t(int rmn, int odr, bool batch)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return batch ? k_sweep2d<0, R, O, T, true, false, FMA> : k_sweep2d<0, R, O, T, false, false, FMA>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

}  // namespace qas
