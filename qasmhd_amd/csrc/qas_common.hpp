// qas_common.hpp -- bodies of the boundary and CFL kernels, shared by the
// single-domain kernels of qas_common.hip and the batched kernels of
// qas_batch.hip.
#pragma once
#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

struct BoundArgs {
  double* f[8];
  int nvar;
  const double* phi = nullptr;  // 2D gravity: hydrostatic energy ghosts along y
  double gfac = 0.0;            // (2 - gam) / (gam - 1)
};

// fv(m): the array of variable m (a kernel argument's entry, or a pointer
// read from an instance table in device memory: a local copy of the
// pointer array indexed by m would sit in scratch memory); skip(m, s):
// the ghosts of variable m whose source is cell s are written elsewhere
template <typename FV, typename SKIP>
__host__ __device__ __forceinline__ void bound_at_f(const FV& fv, const SKIP& skip, int nvar, const double* phi,
                                           double gfac, const Geo& g, const BcFlags& bc, long long nghost, int ox,
                                           int oy, int oz, long long t)
{
  if (t >= nghost * nvar) return;
  const int m = (int)(t / nghost);
  long long r = t - (long long)m * nghost;
  // decode ghost cell r: z slabs, then y slabs of the inner z range, then
  // x slabs of the inner (y, z) range
  const long long nx = g.nx, ny = g.ny, nz = g.nz;
  const long long zs = 2LL * oz * ny * nx;
  const long long ys = 2LL * oy * nx * (nz - 2 * oz);
  int i, j, k;
  if (r < zs) {
    const long long plane = r / (ny * nx);
    const long long rem = r - plane * ny * nx;
    k = (int)(plane < oz ? plane : nz - 2 * oz + plane);
    j = (int)(rem / nx);
    i = (int)(rem % nx);
  } else if (r < zs + ys) {
    r -= zs;
    const long long row = r / nx;  // 0 .. 2*oy*(nz-2oz)-1
    i = (int)(r % nx);
    k = oz + (int)(row / (2 * oy));
    const int jj = (int)(row % (2 * oy));
    j = jj < oy ? jj : (int)ny - 2 * oy + jj;
  } else {
    r -= zs + ys;
    const long long col = r % (2 * ox);
    const long long line = r / (2 * ox);
    i = (int)(col < ox ? col : nx - 2 * ox + col);
    j = oy + (int)(line % (ny - 2 * oy));
    k = oz + (int)(line / (ny - 2 * oy));
  }
  if (phi && m == 7 && nvar == 8 && oz == 0 && (j < oy || j >= ny - oy)) {
    // with gravity along y the energy's y ghosts (x ghosts included) are
    // the hydrostatic extrapolation from the first / last interior row, as
    // 2D/SERIAL/RTI/mhd2d_class.cpp:10-32 sets them after bc2d; that row's
    // ro and en are taken through their x boundary maps
    const int jb = j < oy ? oy : (int)ny - oy - 1;
    const Src ex = resolve(i, g.nx, bc.st[0][7], bc.dn[0][7]);
    const Src rx = resolve(i, g.nx, bc.st[0][0], bc.dn[0][0]);
    double* en = fv(7);
    const double enb = apply(ex, en[nx * jb + ex.i]);
    const double rob = apply(rx, fv(0)[nx * jb + rx.i]);
    const long long ss = nx * j + i, sb = nx * jb + i;
    en[ss] = enb - gfac * rob * (phi[ss] - phi[sb]);
    return;
  }
  const Src sx = ox ? resolve(i, g.nx, bc.st[0][m], bc.dn[0][m]) : Src{i, 1.0, false};
  const Src sy = oy ? resolve(j, g.ny, bc.st[1][m], bc.dn[1][m]) : Src{j, 1.0, false};
  const Src sz = oz ? resolve(k, g.nz, bc.st[2][m], bc.dn[2][m]) : Src{k, 1.0, false};
  if (sx.i == i && sy.i == j && sz.i == k) return;  // a kept cell (staggered boundary face)
  const long long s = (long long)nx * (ny * sz.i + sy.i) + sx.i;
  if (skip(m, s)) return;
  double v = fv(m)[s];
  v = apply(sx, v);
  v = apply(sy, v);
  v = apply(sz, v);
  fv(m)[(long long)nx * (ny * k + j) + i] = v;
}

template <typename FV>
__host__ __device__ __forceinline__ void bound_at_f(const FV& fv, int nvar, const double* phi, double gfac, const Geo& g,
                                           const BcFlags& bc, long long nghost, int ox, int oy, int oz, long long t)
{
  bound_at_f(fv, [](int, long long) { return false; }, nvar, phi, gfac, g, bc, nghost, ox, oy, oz, t);
}

__device__ __forceinline__ void bound_at(const BoundArgs& a, const Geo& g, const BcFlags& bc, long long nghost,
                                         int ox, int oy, int oz, long long t)
{
  bound_at_f([&](int m) { return a.f[m]; }, a.nvar, a.phi, a.gfac, g, bc, nghost, ox, oy, oz, t);
}

struct CflArgs {
  const double* u[8];
  double* part;  // one maximum per block
  double gam;
  int stx4, sty5, stz6;
  const double* phi = nullptr;  // 2D gravity: potential round trip of the energy
  double* enw = nullptr;        // the energy array it is written back to (u[7])
};

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = (w > v) ? w : v;
  }
  return v;
}

// max over the block's 4 wavefronts, valid in thread 0
__device__ __forceinline__ double block_max(double v)
{
  __shared__ double wm[4];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) v = (wm[w] > v) ? wm[w] : v;
  }
  return v;
}

// grid-stride over the interior; each block writes its maximum of
// |v| + vfast (floor 1.0, the reference's initial vmax) to part[block]
__device__ __forceinline__ void cfl_block(const CflArgs& a, const Geo& g, int blk, int nblk)
{
  const int ox = qas::NG;
  const int oy = g.ndim >= 2 ? qas::NG : 0;
  const int oz = g.ndim >= 3 ? qas::NG : 0;
  const int ni = g.nx - 2 * ox, nj = g.ny - 2 * oy, nk = g.nz - 2 * oz;
  const long long ntot = (long long)ni * nj * nk;
  double vm = 1.0;
  for (long long t = (long long)blk * blockDim.x + threadIdx.x; t < ntot; t += (long long)nblk * blockDim.x) {
    const int i = ox + (int)(t % ni);
    const int j = oy + (int)((t / ni) % nj);
    const int k = oz + (int)(t / ((long long)ni * nj));
    const long long ss = (long long)g.nx * ((long long)g.ny * k + j) + i;
    const double gam = a.gam;
    const double ro = a.u[0][ss];
    const double* bx = a.u[4];
    const double* by = a.u[5];
    const double* bz = a.u[6];
    double cx, cy, cz;
    if (g.ndim == 1) {
      cx = bx[ss];
      cy = by[ss];
      cz = bz[ss];
    } else {
      cx = 0.5 * (bx[ss] + bx[ss + a.stx4]);
      cy = 0.5 * (by[ss] + by[ss + g.sy * a.sty5]);
      cz = (g.ndim == 3) ? 0.5 * (bz[ss] + bz[ss + g.sz * a.stz6]) : bz[ss];
    }
    const double iro = 1.0 / ro;
    const double vx = a.u[1][ss] * iro;
    const double vy = a.u[2][ss] * iro;
    const double vz = a.u[3][ss] * iro;
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double b2 = cx * cx + cy * cy + cz * cz;
    double en = a.u[7][ss];
    if (a.phi) {  // setdt's potential round trip, in place (RTI/mhd2d_class.cpp:46-48)
      const double rp = ro * a.phi[ss];
      en = en - rp;
      a.enw[ss] = en + rp;
    }
    const double pr = (gam - 1) * (en - 0.5 * (ro * v2 + b2));
    const double vf = sqrt((gam * pr + (cx * cx + cy * cy + cz * cz)) / ro);
    double v;
    if (g.ndim == 1) v = fabs(vx) + vf;
    else if (g.ndim == 2) v = sqrt(vx * vx + vy * vy) + vf;
    else v = sqrt(vx * vx + vy * vy + vz * vz) + vf;
    if (v > vm) vm = v;
  }
  vm = block_max(vm);
  if (threadIdx.x == 0) a.part[blk] = vm;
}

// dt = cfl * dr / max(part[0..nb)); one block of 256 threads.  max is
// exact and order-independent, so dt equals the reference's serial loop.
__device__ __forceinline__ void cfl_final_block(const double* part, int nb, double* dt, double* hist_slot,
                                                double cfl_dr)
{
  double vm = 1.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) vm = (part[b] > vm) ? part[b] : vm;
  vm = block_max(vm);
  if (threadIdx.x == 0) {
    const double d = cfl_dr / vm;
    *dt = d;
    if (hist_slot) *hist_slot = d;
  }
}

}  // namespace qas
