// qas_stage2d.hpp -- per-thread bodies of the 2D stage kernels other than
// the sweeps (prim, corner EMF, CT), shared by the single-domain kernels of
// qas_kernels2d.hip and the batched kernels of qas_batch.hip.  Each body
// takes the flat thread index t of its domain.
#pragma once
#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

// ------------------------------------------------------------------ prim
// PT: storage type of the primitives (double, or float for the mixed path);
// the arithmetic is FP64 either way.
template <int ODR, typename PT>
__device__ __forceinline__ void prim2d_at(const StagePtrs& p, const Geo& g, const BcFlags& bc, double gam, long long t)
{
  if (t >= g.nd) return;
  const int i = (int)(t % g.nx);
  const int j = (int)(t / g.nx);
  const long long nx = g.nx;
  const double* bx = p.in[4];
  const double* by = p.in[5];
  // cx: fcen along x on i in [1, nx-3], then bound with (st 0, dnxs[4], dnys[4])
  double cx, cy;
  {
    const Src sx = resolve(i, g.nx, 0, bc.dn[0][4]);
    const Src sy = resolve(j, g.ny, 0, bc.dn[1][4]);
    const long long q = nx * sy.i + sx.i;
    double v = fcen<ODR>(bx[q - 1], bx[q], bx[q + 1], bx[q + 2]);
    cx = apply(sy, apply(sx, v));
  }
  {
    const Src sx = resolve(i, g.nx, 0, bc.dn[0][5]);
    const Src sy = resolve(j, g.ny, 0, bc.dn[1][5]);
    const long long q = nx * sy.i + sx.i;
    double v = fcen<ODR>(by[q - nx], by[q], by[q + nx], by[q + 2 * nx]);
    cy = apply(sy, apply(sx, v));
  }
  const double ro = p.in[0][t];
  const double iro = 1.0 / ro;
  const double vx = p.in[1][t] * iro;
  const double vy = p.in[2][t] * iro;
  const double vz = p.in[3][t] * iro;
  const double cz = p.in[6][t];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double b2 = cx * cx + cy * cy + cz * cz;
  double en = p.in[7][t];
  if (p.phi) {
    // gravity: the potential energy is taken out for the pressure and put
    // back in the stage's energy (2D/SERIAL/RTI/mhd2d_solve.cpp:86-88)
    const double rp = ro * p.phi[t];
    en = en - rp;
    p.enrt[t] = en + rp;
  }
  store_prim<PT>(p.pr, t, (gam - 1) * (en - 0.5 * (ro * v2 + b2)));
  store_prim<PT>(p.vx, t, vx);
  store_prim<PT>(p.vy, t, vy);
  store_prim<PT>(p.vz, t, vz);
  store_prim<PT>(p.cx, t, cx);
  store_prim<PT>(p.cy, t, cy);
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
    ct = cuct_weight(ro[c0], ro[c1], ro[c2], ro[c3], vel(p, 0, c0), vel(p, 0, c1), vel(p, 0, c2), vel(p, 0, c3),
                     vel(p, 1, c0), vel(p, 1, c1), vel(p, 1, c2), vel(p, 1, c3), p.in[4][ss - nx], p.in[4][ss],
                     p.in[5][ss - 1], p.in[5][ss]);
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

// ------------------------------------------------------------------ CT
template <int ODR>
__device__ __forceinline__ void ct2d_at(const StagePtrs& p, const Geo& g, long long t)
{
  // faces i, j in [4, n-4]: bx needs j <= ny-5, by needs i <= nx-5
  const int ni = g.nx - 7, nj = g.ny - 7;
  if (t >= (long long)ni * nj) return;
  const int i = 4 + (int)(t % ni);
  const int j = 4 + (int)(t / ni);
  const long long nx = g.nx;
  const long long ss = nx * j + i;
  const double dtv = *p.dt;
  const double* ez = p.ez;
  if (j <= g.ny - 5) {
    const double dtdy = dtv / g.dy;
    const double du = dfc<ODR>(ez[ss - nx], ez[ss], ez[ss + nx], ez[ss + 2 * nx]);
    p.out[4][ss] = rk_upd(p.in[4][ss], p.u0[4][ss], -dtdy * du, p.a, p.b);
  }
  if (i <= g.nx - 5) {
    const double mdtdx = -(dtv / g.dx);  // mhd_updt2d_ctb is called with -dtdx
    const double du = dfc<ODR>(ez[ss - 1], ez[ss], ez[ss + 1], ez[ss + 2]);
    p.out[5][ss] = rk_upd(p.in[5][ss], p.u0[5][ss], -mdtdx * du, p.a, p.b);
  }
}

}  // namespace qas
