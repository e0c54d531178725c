// qas_stage2d.hpp -- 2D stage bodies other than the sweeps' own: the
// primitives of one cell (used by the y sweep's staging), and the corner
// EMF and CT bodies (per thread, flat index t of the domain), shared by the
// single-domain kernels of qas_kernels2d.hip and the batched kernels of
// qas_batch.hip.
#pragma once
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_common.hpp"

namespace qas {

// ------------------------------------------------------------------ prim
// Cell-centred B and primitives of cell (i, j) (flat index t) from the
// stage input state, for any cell including ghosts: fcen of the staggered
// bx (along x) and by (along y) on the cells bcell covers, then the
// bound() index map (mhd_class.hpp:86-91 with common_func.c:60-138);
// v = m (1/ro) and the pressure recovery of mhd_ieos (mhd_func.c:6-16).
// FP64 with the reference's operation order in every precision (no FMA
// contraction here, also in a translation unit built with it).  With
// GRAV the potential energy ro phi is taken out of en for the pressure and
// put back (2D/SERIAL/RTI/mhd2d_solve.cpp:86-88): enrt is that round-
// tripped energy, the stage input energy of the sweeps.
struct Prim2d {
  double vx, vy, vz, pr, cx, cy, enrt;
};

// INNER: the caller knows 4 <= i <= nx-5 and 4 <= j <= ny-5, where the
// bound() index map is the identity (no ghost cell), and skips it.
template <int ODR, bool GRAV, bool INNER = false>
__device__ __forceinline__ Prim2d prim2d_cell(const StagePtrs& p, const Geo& g, const BcFlags& bc, double gam, int i,
                                              int j, long long t)
{
#pragma clang fp contract(off)
  const long long nx = g.nx;
  const auto bx = G(p.in[4]);
  const auto by = G(p.in[5]);
  Prim2d r;
  if constexpr (INNER) {
    r.cx = fcen<ODR>(bx[t - 1], bx[t], bx[t + 1], bx[t + 2]);
    r.cy = fcen<ODR>(by[t - nx], by[t], by[t + nx], by[t + 2 * nx]);
  } else {
  {
    const Src sx = resolve(i, g.nx, 0, bc.dn[0][4]);
    const Src sy = resolve(j, g.ny, 0, bc.dn[1][4]);
    const long long q = nx * sy.i + sx.i;
    const double v = fcen<ODR>(bx[q - 1], bx[q], bx[q + 1], bx[q + 2]);
    r.cx = apply(sy, apply(sx, v));
  }
  {
    const Src sx = resolve(i, g.nx, 0, bc.dn[0][5]);
    const Src sy = resolve(j, g.ny, 0, bc.dn[1][5]);
    const long long q = nx * sy.i + sx.i;
    const double v = fcen<ODR>(by[q - nx], by[q], by[q + nx], by[q + 2 * nx]);
    r.cy = apply(sy, apply(sx, v));
  }
  }
  const double ro = G(p.in[0])[t];
  const double iro = 1.0 / ro;
  r.vx = G(p.in[1])[t] * iro;
  r.vy = G(p.in[2])[t] * iro;
  r.vz = G(p.in[3])[t] * iro;
  const double cz = G(p.in[6])[t];
  const double v2 = r.vx * r.vx + r.vy * r.vy + r.vz * r.vz;
  const double b2 = r.cx * r.cx + r.cy * r.cy + cz * cz;
  double en = G(p.in[7])[t];
  r.enrt = en;
  if constexpr (GRAV) {
    const double rp = ro * p.phi[t];
    en = en - rp;
    r.enrt = en + rp;
  }
  r.pr = (gam - 1) * (en - 0.5 * (ro * v2 + b2));
  return r;
}

// velocity component of cell c from the state: the same expression as the
// prim kernel (m * (1/ro)), so the FP64 value is bit-identical to it
__device__ __forceinline__ double vel(const StagePtrs& p, int comp, long long c)
{
  return p.in[1 + comp][c] * (1.0 / p.in[0][c]);
}

// ------------------------------------------------------------------ EMF
// FT: element type of the face parts (the sweep precision)
template <int ODR, typename FT>
__device__ __forceinline__ void emf2d_at(const StagePtrs& p, const Geo& g, int ctw, long long t)
{
  const int ni = g.nx - 5, nj = g.ny - 5;   // corners [3, n-3]
  if (t >= (long long)ni * nj) return;
  const int i = 3 + (int)(t % ni);
  const int j = 3 + (int)(t / ni);
  const long long nx = g.nx;
  const long long ss = nx * j + i;
  double ct = 0.5;
  if (ctw) {
    const double* ro = p.in[0];
    const long long c0 = ss - 1 - nx, c1 = ss - nx, c2 = ss - 1, c3 = ss;
    // the cells' velocities: in FP64 the primitives the y sweep stored
    // (m * (1/ro) of the same state, so the same values); in the mixed path
    // those are FP32, and the weight takes m * (1/ro) of the state instead
    auto v = [&](int comp, long long c) -> double {
      if constexpr (sizeof(FT) == 8) return (comp == 0 ? p.vx : p.vy)[c];
      else return vel(p, comp, c);
    };
    ct = cuct_weight(ro[c0], ro[c1], ro[c2], ro[c3], v(0, c0), v(0, c1), v(0, c2), v(0, c3), v(1, c0), v(1, c1),
                     v(1, c2), v(1, c3), p.in[4][ss - nx], p.in[4][ss], p.in[5][ss - 1], p.in[5][ss]);
  }
  double ez = 0.0;
  {  // from the x sweep: By flux interpolated along y to (i-1/2, j-1/2)
    const double* fu = p.emf[0];
    const double* fcn = p.emf[1];
    double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(fu, ss + (m - 3) * nx); b[m] = load_as<FT>(fcn, ss + (m - 3) * nx); }
    lin5<ODR>(a, ul0, dum);
    lin5<ODR>(b, ul1, dum);
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(fu, ss + (m - 2) * nx); b[m] = load_as<FT>(fcn, ss + (m - 2) * nx); }
    lin5<ODR>(a, dum, ur0);
    lin5<ODR>(b, dum, ur1);
    ez += -0.5 * ((ul0 + ur0) + (1.0 - ct) * (ul1 + ur1));
  }
  {  // from the y sweep: Bx flux interpolated along x
    const double* fu = p.emf[4];
    const double* fcn = p.emf[5];
    double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(fu, ss + (m - 3)); b[m] = load_as<FT>(fcn, ss + (m - 3)); }
    lin5<ODR>(a, ul0, dum);
    lin5<ODR>(b, ul1, dum);
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(fu, ss + (m - 2)); b[m] = load_as<FT>(fcn, ss + (m - 2)); }
    lin5<ODR>(a, dum, ur0);
    lin5<ODR>(b, dum, ur1);
    ez += +0.5 * ((ul0 + ur0) + ct * (ul1 + ur1));
  }
  p.ez[ss] = ez;
}

// CJ corners (i, j0 .. j0 + CJ - 1) of one column per thread, each with
// emf2d_at's arithmetic: the x-sweep face parts along y and the cell and
// face values of the CUCT weight are loaded once into register windows
// that the column's corners share (per corner 40 loads in emf2d_at).
// Window rows past the last corner's stencil are clamped to the array and
// not used.
template <int ODR, typename FT, int CJ>
__device__ __forceinline__ void emf2d_col(const StagePtrs& p, const Geo& g, int ctw, long long t)
{
  const int ni = g.nx - 5, nj = g.ny - 5;   // corners [3, n-3]
  const int ncol = (nj + CJ - 1) / CJ;
  if (t >= (long long)ni * ncol) return;
  const int i = 3 + (int)(t % ni);
  const int j0 = 3 + CJ * (int)(t / ni);
  const int jend = 3 + nj;                  // one past the last corner row
  const long long nx = g.nx;
  auto at = [&](int jj, int di) -> long long { return nx * (jj < g.ny - 1 ? jj : g.ny - 1) + i + di; };
  // x-sweep face parts at rows j0-3 .. j0+CJ+1 of column i
  double fu[CJ + 5], fc[CJ + 5];
#pragma unroll
  for (int r = 0; r < CJ + 5; r++) {
    fu[r] = load_as<FT>(p.emf[0], at(j0 - 3 + r, 0));
    fc[r] = load_as<FT>(p.emf[1], at(j0 - 3 + r, 0));
  }
  // CUCT cells (i-1, i) x rows j0-1 .. j0+CJ-1, bx at (i, row), by at (i-1, i) x rows j0 .. j0+CJ-1
  double ro[CJ + 1][2], va[CJ + 1][2], vb[CJ + 1][2], bx[CJ + 1], by[CJ][2];
  if (ctw) {
#pragma unroll
    for (int r = 0; r < CJ + 1; r++) {
#pragma unroll
      for (int d = 0; d < 2; d++) {
        const long long c = at(j0 - 1 + r, d - 1);
        ro[r][d] = G(p.in[0])[c];
        if constexpr (sizeof(FT) == 8) {
          va[r][d] = G(p.vx)[c];
          vb[r][d] = G(p.vy)[c];
        } else {
          va[r][d] = vel(p, 0, c);
          vb[r][d] = vel(p, 1, c);
        }
      }
      bx[r] = G(p.in[4])[at(j0 - 1 + r, 0)];
    }
#pragma unroll
    for (int r = 0; r < CJ; r++) {
      by[r][0] = G(p.in[5])[at(j0 + r, -1)];
      by[r][1] = G(p.in[5])[at(j0 + r, 0)];
    }
  }
  const double* gu = p.emf[4];
  const double* gc = p.emf[5];
  const auto ezo = G(p.ez);
  double ezv[CJ];
#pragma unroll
  for (int c = 0; c < CJ; c++) {
    const int j = j0 + c;
    // (a corner past the last row computes on clamped rows, never stored)
    const long long ss = nx * (j < jend ? j : jend - 1) + i;
    double ct = 0.5;
    if (ctw)
      ct = cuct_weight(ro[c][0], ro[c][1], ro[c + 1][0], ro[c + 1][1], va[c][0], va[c][1], va[c + 1][0],
                       va[c + 1][1], vb[c][0], vb[c][1], vb[c + 1][0], vb[c + 1][1], bx[c], bx[c + 1], by[c][0],
                       by[c][1]);
    double ez = 0.0;
    {  // from the x sweep: By flux interpolated along y to (i-1/2, j-1/2)
      double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
      for (int m = 0; m < 5; m++) { a[m] = fu[c + m]; b[m] = fc[c + m]; }
      lin5<ODR>(a, ul0, dum);
      lin5<ODR>(b, ul1, dum);
#pragma unroll
      for (int m = 0; m < 5; m++) { a[m] = fu[c + 1 + m]; b[m] = fc[c + 1 + m]; }
      lin5<ODR>(a, dum, ur0);
      lin5<ODR>(b, dum, ur1);
      ez += -0.5 * ((ul0 + ur0) + (1.0 - ct) * (ul1 + ur1));
    }
    {  // from the y sweep: Bx flux interpolated along x
      double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
      for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(gu, ss + (m - 3)); b[m] = load_as<FT>(gc, ss + (m - 3)); }
      lin5<ODR>(a, ul0, dum);
      lin5<ODR>(b, ul1, dum);
#pragma unroll
      for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(gu, ss + (m - 2)); b[m] = load_as<FT>(gc, ss + (m - 2)); }
      lin5<ODR>(a, dum, ur0);
      lin5<ODR>(b, dum, ur1);
      ez += +0.5 * ((ul0 + ur0) + ct * (ul1 + ur1));
    }
    ezv[c] = ez;
  }
  // the stores after every load (a batched launch reads its pointers from
  // the instance table, which the compiler must otherwise reload)
#pragma unroll
  for (int c = 0; c < CJ; c++)
    if (j0 + c < jend) ezo[nx * (j0 + c) + i] = ezv[c];
}

// ------------------------------------------------------------------ CT
// the new bx (comp 0, faces i in [4, nx-4], rows j in [4, ny-5]) or by
// (comp 1, i in [4, nx-5], j in [4, ny-4]) at face ss
template <int ODR>
__host__ __device__ __forceinline__ double ct2d_face(const StagePtrs& p, const Geo& g, long long ss, int comp)
{
  const long long nx = g.nx;
  const double dtv = G(p.dt)[0];
  const auto ez = G(p.ez);
  if (comp == 0) {
    const double dtdy = dtv / g.dy;
    const double du = dfc<ODR>(ez[ss - nx], ez[ss], ez[ss + nx], ez[ss + 2 * nx]);
    return rk_upd(G(p.in[4])[ss], G(p.u0[4])[ss], -dtdy * du, p.a, p.b);
  }
  const double mdtdx = -(dtv / g.dx);  // mhd_updt2d_ctb is called with -dtdx
  const double du = dfc<ODR>(ez[ss - 1], ez[ss], ez[ss + 1], ez[ss + 2]);
  return rk_upd(G(p.in[5])[ss], G(p.u0[5])[ss], -mdtdx * du, p.a, p.b);
}

// CT alone (the batched stages, whose ghost fill is its own launch)
template <int ODR>
__device__ __forceinline__ void ct2d_at(const StagePtrs& p, const Geo& g, long long t)
{
  // faces i, j in [4, n-4]: bx needs j <= ny-5, by needs i <= nx-5
  const int ni = g.nx - 7, nj = g.ny - 7;
  if (t >= (long long)ni * nj) return;
  const int i = 4 + (int)(t % ni);
  const int j = 4 + (int)(t / ni);
  const long long ss = (long long)g.nx * j + i;
  // both updates before the first store: a batched launch reads the stage
  // pointers from its instance table, which the compiler must otherwise
  // reload after a store
  const bool wx = j <= g.ny - 5, wy = i <= g.nx - 5;
  const double vx = wx ? ct2d_face<ODR>(p, g, ss, 0) : 0.0;
  const double vy = wy ? ct2d_face<ODR>(p, g, ss, 1) : 0.0;
  if (wx) G(p.out[4])[ss] = vx;
  if (wy) G(p.out[5])[ss] = vy;
}

// true where bound() writes face (i, j) of variable m (4: bx, 5: by)
__host__ __device__ __forceinline__ bool bound_writes(const Geo& g, const BcFlags& bc, int m, int i, int j)
{
  return resolve(i, g.nx, bc.st[0][m], bc.dn[0][m]).i != i || resolve(j, g.ny, bc.st[1][m], bc.dn[1][m]).i != j;
}

// the CT face range of variable m (4: bx, 5: by) holds cell (i, j)
__host__ __device__ __forceinline__ bool ct_range(const Geo& g, int m, int i, int j)
{
  return i >= 4 && j >= 4 && i <= g.nx - 4 - (m == 5) && j <= g.ny - 4 - (m == 4);
}

// The ghost columns (or rows) of one axis that bound() fills from the
// kept cell i, as a mask over the 2 NG ghost slots: slot a < NG is column
// a, slot a >= NG column n - 2 NG + a (resolve() inverted: periodic images,
// the mirrored cell, the edge cell of an open boundary for all its ghosts)
__host__ __device__ __forceinline__ unsigned ghost_images(int i, int n, int st, int dn)
{
  unsigned msk = 0;
  if (dn == 0) {
    if (i >= NG && i < 2 * NG) msk |= 1u << i;
    const int l = i - (n - 2 * NG);
    if (l >= 0 && l < NG) msk |= 1u << l;
  } else if (dn == 1 || dn == -1) {
    const int l = (2 * NG - 1 + st) - i;
    if (l >= 0 && l < NG) msk |= 1u << l;
    const int r = (n - 2 * NG + st) + (n - 1 - i);
    if (r >= n - NG + st && r < n) msk |= 1u << (r - (n - 2 * NG));
  } else if (dn == 2 || dn == -2) {
    if (i == NG) msk |= (1u << NG) - 1;
    if (i == n - 1 - NG + st) msk |= ((1u << (NG - st)) - 1) << (NG + st);
  }
  return msk;
}

// writes v (the new value of a kept face (i, j)) into every ghost whose
// bound() source it is: mx, my are its ghost_images on the two axes and
// fx, fy the factors bound() applies there (the same on both sides of an
// axis), applied in bound_at_f's order (x, then y)
__host__ __device__ __forceinline__ void push_ghosts(double* f, int nx, int ny, int i, int j, unsigned mx, unsigned my,
                                            const Src& fx, const Src& fy, double v)
{
  for (unsigned rx = mx; rx; rx &= rx - 1) {
    const int a = __builtin_ctz(rx);
    const int gx = a < NG ? a : nx - 2 * NG + a;
    const double vx = apply(fx, v);
    f[(long long)nx * j + gx] = vx;
    for (unsigned ry = my; ry; ry &= ry - 1) {
      const int b = __builtin_ctz(ry);
      f[(long long)nx * (b < NG ? b : ny - 2 * NG + b) + gx] = apply(fy, vx);
    }
  }
  for (unsigned ry = my; ry; ry &= ry - 1) {
    const int b = __builtin_ctz(ry);
    f[(long long)nx * (b < NG ? b : ny - 2 * NG + b) + i] = apply(fy, v);
  }
}

// CT fused with the ghost fill of the stage output (one launch instead of
// ct2d + bound).  Thread t < nc updates the faces (i, j) with i, j in
// [4, n-4] (bx for j <= ny-5, by for i <= nx-5), except those that bound()
// overwrites (the periodic image of the last face), and writes each
// updated face's ghost images itself; thread nc + r fills ghost r as
// bound_at_f does, except the bx / by ghosts whose source a CT thread
// updates.  Cell-centred sources were written by the x sweep (an earlier
// launch); no thread reads a cell another thread of this launch writes.
template <int ODR>
__host__ __device__ __forceinline__ void ctb2d_at(const StagePtrs& p, const Geo& g, const BcFlags& bc, const double* phi,
                                         double gfac, long long nghost, long long t, bool ct_part)
{
  if (ct_part) {
    const int nx = g.nx, ny = g.ny;
    const int ni = nx - 7, nj = ny - 7;
    if (t >= (long long)ni * nj) return;
    const int i = 4 + (int)(t % ni);
    const int j = 4 + (int)(t / ni);
    const long long ss = (long long)nx * j + i;
    // everything read from the arguments before the first store (a batched
    // launch reads them from its instance table, which the compiler must
    // otherwise reload after each store)
    // only faces within 2 NG + 1 of an edge are ghost sources or ghosts
    bool w4 = j <= ny - 5, w5 = i <= nx - 5;
    unsigned mx4 = 0, my4 = 0, mx5 = 0, my5 = 0;
    if (i <= 2 * NG || i >= nx - 2 * NG - 1 || j <= 2 * NG || j >= ny - 2 * NG - 1) {
      w4 = w4 && !bound_writes(g, bc, 4, i, j);
      w5 = w5 && !bound_writes(g, bc, 5, i, j);
      mx4 = ghost_images(i, nx, bc.st[0][4], bc.dn[0][4]);
      my4 = ghost_images(j, ny, bc.st[1][4], bc.dn[1][4]);
      mx5 = ghost_images(i, nx, bc.st[0][5], bc.dn[0][5]);
      my5 = ghost_images(j, ny, bc.st[1][5], bc.dn[1][5]);
    }
    const Src fx4 = resolve(0, nx, bc.st[0][4], bc.dn[0][4]), fy4 = resolve(0, ny, bc.st[1][4], bc.dn[1][4]);
    const Src fx5 = resolve(0, nx, bc.st[0][5], bc.dn[0][5]), fy5 = resolve(0, ny, bc.st[1][5], bc.dn[1][5]);
    double* o4 = p.out[4];
    double* o5 = p.out[5];
    const double v4 = w4 ? ct2d_face<ODR>(p, g, ss, 0) : 0.0;
    const double v5 = w5 ? ct2d_face<ODR>(p, g, ss, 1) : 0.0;
    if (w4) o4[ss] = v4;
    if (w5) o5[ss] = v5;
    if ((mx4 | my4 | mx5 | my5) != 0) {
      if (w4) push_ghosts(o4, nx, ny, i, j, mx4, my4, fx4, fy4, v4);
      if (w5) push_ghosts(o5, nx, ny, i, j, mx5, my5, fx5, fy5, v5);
    }
    return;
  }
  const long long nx = g.nx;
  auto pushed = [&](int m, long long s) {
    return (m == 4 || m == 5) && ct_range(g, m, (int)(s % nx), (int)(s / nx));
  };
  bound_at_f([&](int m) { return p.out[m]; }, pushed, 8, phi, gfac, g, bc, nghost, NG, NG, 0, t);
}

}  // namespace qas
