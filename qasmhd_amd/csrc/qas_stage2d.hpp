// qas_stage2d.hpp -- 2D stage bodies other than the sweeps' own: the
// primitives of one cell (used by the y sweep's staging) and the tiled
// corner-EMF + CT kernel body, shared by the single-domain kernels of
// qas_kernels2d.hip and the batched kernels of qas_batch.hip.
#pragma once
#include "qas_device.hpp"
#include "qas_internal.hpp"

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
  const double* bx = p.in[4];
  const double* by = p.in[5];
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
  const double ro = p.in[0][t];
  const double iro = 1.0 / ro;
  r.vx = p.in[1][t] * iro;
  r.vy = p.in[2][t] * iro;
  r.vz = p.in[3][t] * iro;
  const double cz = p.in[6][t];
  const double v2 = r.vx * r.vx + r.vy * r.vy + r.vz * r.vz;
  const double b2 = r.cx * r.cx + r.cy * r.cy + cz * cz;
  double en = p.in[7][t];
  r.enrt = en;
  if constexpr (GRAV) {
    const double rp = ro * p.phi[t];
    en = en - rp;
    r.enrt = en + rp;
  }
  r.pr = (gam - 1) * (en - 0.5 * (ro * v2 + b2));
  return r;
}

// ------------------------------------------------------------------ EMF + CT
// Corner EMF (mhd2d_solve.cpp:91-103, 201-232, 308-339) at corner (i-1/2,
// j-1/2) stored at ss = nx j + i, from both sweeps' face parts, with the
// CUCT weight ct.  FT: element type of the face parts (the sweep precision).
template <int ODR, typename FT>
__device__ __forceinline__ double emf2d_corner(const StagePtrs& p, long long nx, long long ss, double ct)
{
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
  return ez;
}

// One tile of the corner-EMF + CT stage (mhd_updt2d_ctb, mhd_func.c:286-300,
// mhd2d_solve.cpp:359-380): the staggered bx, by of the ET_X x ET_Y
// interior faces [I0, I0+ET_X) x [J0, J0+ET_Y) (the last tile along x also
// takes bx of face column nx-4, the last along y by of face row ny-4).
//   1. cells [I0-2, I0+ET_X+1] x [J0-2, J0+ET_Y+1]: ro and v = m (1/ro) into
//      LDS (CUCT on), once per cell rather than once per corner using it;
//   2. corners [I0-1, I0+ET_X+1] x [J0-1, J0+ET_Y+1]: CUCT weight and Ez
//      into LDS (the edge EMF never goes to memory);
//   3. faces: the 4th-order curl of Ez and the RK stage sum.
// Same expressions and operands as one thread per corner and per face:
// bit-identical.  256 threads.
constexpr int ET_X = 64, ET_Y = 16;
constexpr int ET_CX = ET_X + 4, ET_CY = ET_Y + 4;  // cell region
constexpr int ET_KX = ET_X + 3, ET_KY = ET_Y + 3;  // corner region

__host__ __device__ inline int emfct2d_tiles(int nx, int ny)
{
  return ((nx - 2 * NG + ET_X - 1) / ET_X) * ((ny - 2 * NG + ET_Y - 1) / ET_Y);
}

template <int ODR, typename FT>
__device__ __forceinline__ void emfct2d_tile(const StagePtrs& p, const Geo& g, int ctw, int tile)
{
  __shared__ double sro[ET_CY][ET_CX], svx[ET_CY][ET_CX], svy[ET_CY][ET_CX];
  __shared__ double sez[ET_KY][ET_KX + 1];
  const int ntx = (g.nx - 2 * NG + ET_X - 1) / ET_X;
  const int nty = (g.ny - 2 * NG + ET_Y - 1) / ET_Y;
  if (tile >= ntx * nty) return;  // whole workgroup
  const int tx = tile % ntx, ty = tile / ntx;
  const int I0 = NG + tx * ET_X, J0 = NG + ty * ET_Y;
  const long long nx = g.nx;
  const int t = (int)threadIdx.x;
  if (ctw) {
    for (int e = t; e < ET_CX * ET_CY; e += blockDim.x) {
      const int cx = e % ET_CX, cy = e / ET_CX;
      int i = I0 - 2 + cx, j = J0 - 2 + cy;  // cells outside the domain are never used
      i = i > g.nx - 1 ? g.nx - 1 : i;
      j = j > g.ny - 1 ? g.ny - 1 : j;
      const long long c = nx * j + i;
      const double ro = p.in[0][c];
      const double iro = 1.0 / ro;  // m * (1/ro), the primitives' expression
      sro[cy][cx] = ro;
      svx[cy][cx] = p.in[1][c] * iro;
      svy[cy][cx] = p.in[2][c] * iro;
    }
    __syncthreads();
  }
  for (int e = t; e < ET_KX * ET_KY; e += blockDim.x) {
    const int kx = e % ET_KX, ky = e / ET_KX;
    const int i = I0 - 1 + kx, j = J0 - 1 + ky;
    double ez = 0.0;
    if (i <= g.nx - 3 && j <= g.ny - 3) {  // corners [3, n-3]
      const long long ss = nx * j + i;
      double ct = 0.5;
      if (ctw) {
        // cells (i-1, j-1), (i, j-1), (i-1, j), (i, j) at region (kx, ky) + (0|1, 0|1)
        ct = cuct_weight(sro[ky][kx], sro[ky][kx + 1], sro[ky + 1][kx], sro[ky + 1][kx + 1], svx[ky][kx],
                         svx[ky][kx + 1], svx[ky + 1][kx], svx[ky + 1][kx + 1], svy[ky][kx], svy[ky][kx + 1],
                         svy[ky + 1][kx], svy[ky + 1][kx + 1], p.in[4][ss - nx], p.in[4][ss], p.in[5][ss - 1],
                         p.in[5][ss]);
      }
      ez = emf2d_corner<ODR, FT>(p, nx, ss, ct);
    }
    sez[ky][kx] = ez;
  }
  __syncthreads();
  const double dtv = *p.dt;
  const double dtdy = dtv / g.dy;
  const double mdtdx = -(dtv / g.dx);  // mhd_updt2d_ctb is called with -dtdx
  // bx at (i, j): Ez at (i, j-1 .. j+2); by at (i, j): Ez at (i-1 .. i+2, j)
  auto upd_bx = [&](int i, int j) {
    const int kx = i - (I0 - 1), ky = j - (J0 - 1);
    const long long ss = nx * j + i;
    const double du = dfc<ODR>(sez[ky - 1][kx], sez[ky][kx], sez[ky + 1][kx], sez[ky + 2][kx]);
    p.out[4][ss] = rk_upd(p.in[4][ss], p.u0[4][ss], -dtdy * du, p.a, p.b);
  };
  auto upd_by = [&](int i, int j) {
    const int kx = i - (I0 - 1), ky = j - (J0 - 1);
    const long long ss = nx * j + i;
    const double du = dfc<ODR>(sez[ky][kx - 1], sez[ky][kx], sez[ky][kx + 1], sez[ky][kx + 2]);
    p.out[5][ss] = rk_upd(p.in[5][ss], p.u0[5][ss], -mdtdx * du, p.a, p.b);
  };
  const int i = I0 + t % ET_X;
#pragma unroll
  for (int k = 0; k < ET_Y / 4; k++) {
    const int j = J0 + t / ET_X + 4 * k;
    if (i <= g.nx - NG - 1 && j <= g.ny - NG - 1) {
      upd_bx(i, j);
      upd_by(i, j);
    }
  }
  if (tx == ntx - 1 && t < ET_Y) {  // bx of the face column nx-4
    const int j = J0 + t;
    if (j <= g.ny - NG - 1) upd_bx(g.nx - NG, j);
  }
  if (ty == nty - 1 && t >= ET_X && t < 2 * ET_X) {  // by of the face row ny-4
    const int ii = I0 + t - ET_X;
    if (ii <= g.nx - NG - 1) upd_by(ii, g.ny - NG);
  }
}

}  // namespace qas
