// qas_common.hip -- boundary fill and CFL reduction kernels (all dimensions).
//
//   k_bound     bound(val) = bc1d / bc2d / bc3d (common/common_func.c:34-278)
//               on one device-resident domain: each ghost cell is written
//               once from the kept cell that the x -> y -> z passes of the
//               reference would copy into it, times the pass factors.
//   k_cfl       setdt(1)'s max of |v| + vfast over the interior with the
//               2nd-order cell-centred B (mhd2d_class.cpp:11-33,
//               mhd1d_class.cpp:11-30, mhd3d_class.cpp:15-46): grid-stride
//               loop, wavefront max by cross-lane shuffles, block max in
//               LDS, one partial maximum per block (no atomics).
//   k_cfl_final dt = cfl*dr/vmax into device memory from the partials,
//               with the reference's floor vmax >= 1.0.
#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

struct BoundArgs {
  double* f[8];
  int nvar;
};

__global__ __launch_bounds__(256) void k_bound(BoundArgs a, Geo g, BcFlags bc, long long nghost, int ox, int oy,
                                               int oz)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nghost * a.nvar) return;
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
  const Src sx = ox ? resolve(i, g.nx, bc.st[0][m], bc.dn[0][m]) : Src{i, 1.0, false};
  const Src sy = oy ? resolve(j, g.ny, bc.st[1][m], bc.dn[1][m]) : Src{j, 1.0, false};
  const Src sz = oz ? resolve(k, g.nz, bc.st[2][m], bc.dn[2][m]) : Src{k, 1.0, false};
  if (sx.i == i && sy.i == j && sz.i == k) return;  // a kept cell (staggered boundary face)
  double* f = a.f[m];
  double v = f[(long long)nx * (ny * sz.i + sy.i) + sx.i];
  v = apply(sx, v);
  v = apply(sy, v);
  v = apply(sz, v);
  f[(long long)nx * (ny * k + j) + i] = v;
}

hipError_t launch_bound(qas_ctx* c, double* const f[8], int nvar)
{
  const Geo& g = c->geo;
  const int ox = qas::NG;
  const int oy = g.ndim >= 2 ? qas::NG : 0;
  const int oz = g.ndim >= 3 ? qas::NG : 0;
  const long long nghost = (long long)g.nx * g.ny * g.nz -
                           (long long)(g.nx - 2 * ox) * (g.ny - 2 * oy) * (g.nz - 2 * oz);
  BoundArgs a;
  for (int m = 0; m < 8; m++) a.f[m] = f[m];
  a.nvar = nvar;
  const long long tot = nghost * nvar;
  QAS_LAUNCH(c, QAS_K_BOUND, k_bound, (unsigned)((tot + 255) / 256), 256, 0, c->stream, a, g, c->bc, nghost, ox, oy, oz);
  return hipGetLastError();
}

// ------------------------------------------------------------------ CFL
struct CflArgs {
  const double* u[8];
  double* part;  // one maximum per block
  double gam;
  int stx4, sty5, stz6;
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
__global__ __launch_bounds__(256) void k_cfl(CflArgs a, Geo g)
{
  const int ox = qas::NG;
  const int oy = g.ndim >= 2 ? qas::NG : 0;
  const int oz = g.ndim >= 3 ? qas::NG : 0;
  const int ni = g.nx - 2 * ox, nj = g.ny - 2 * oy, nk = g.nz - 2 * oz;
  const long long ntot = (long long)ni * nj * nk;
  double vm = 1.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ntot;
       t += (long long)gridDim.x * blockDim.x) {
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
    const double pr = (gam - 1) * (a.u[7][ss] - 0.5 * (ro * v2 + b2));
    const double vf = sqrt((gam * pr + (cx * cx + cy * cy + cz * cz)) / ro);
    double v;
    if (g.ndim == 1) v = fabs(vx) + vf;
    else if (g.ndim == 2) v = sqrt(vx * vx + vy * vy) + vf;
    else v = sqrt(vx * vx + vy * vy + vz * vz) + vf;
    if (v > vm) vm = v;
  }
  vm = block_max(vm);
  if (threadIdx.x == 0) a.part[blockIdx.x] = vm;
}

// dt = cfl * dr / max(part[0..nb)); one block of 256 threads.  max is
// exact and order-independent, so dt equals the reference's serial loop.
__global__ __launch_bounds__(256) void k_cfl_final(const double* part, int nb, double* dt, double* hist_slot,
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

__global__ void k_set_dt(double* dt, double v) { *dt = v; }
__global__ void k_copy1(const double* src, double* dst) { *dst = *src; }
// dt history for graph replays: hist[*cnt] = *dt; ++*cnt (one thread)
__global__ void k_record(const double* dt, double* hist, long long* cnt, long long cap)
{
  const long long n = *cnt;
  if (n < cap) hist[n] = *dt;
  *cnt = n + 1;
}

hipError_t launch_cfl(qas_ctx* c, const double* const U[8])
{
  const Geo& g = c->geo;
  CflArgs a;
  for (int m = 0; m < 8; m++) a.u[m] = U[m];
  a.part = c->d_cflpart;
  a.gam = c->s.gam;
  a.stx4 = c->s.st[0][4];
  a.sty5 = c->s.st[1][5];
  a.stz6 = c->s.st[2][6];
  const int ox = qas::NG, oy = g.ndim >= 2 ? qas::NG : 0, oz = g.ndim >= 3 ? qas::NG : 0;
  const long long n = (long long)(g.nx - 2 * ox) * (g.ny - 2 * oy) * (g.nz - 2 * oz);
  long long nb = (n + 255) / 256;
  if (nb > CFL_MAX_BLOCKS) nb = CFL_MAX_BLOCKS;
  c->cfl_blocks = (int)nb;
  QAS_LAUNCH(c, QAS_K_CFL, k_cfl, (unsigned)nb, 256, 0, c->stream, a, g);
  return hipGetLastError();
}

hipError_t launch_cfl_finalize(qas_ctx* c, double* hist_slot)
{
  const Geo& g = c->geo;
  double dr = g.dx;
  if (g.ndim == 2) dr = (g.dx >= g.dy) ? g.dy : g.dx;
  if (g.ndim == 3) {
    const double t = (g.dx >= g.dy) ? g.dy : g.dx;
    dr = (t >= g.dz) ? g.dz : t;
  }
  const double cfl_dr = c->s.cfl * dr;
  hipLaunchKernelGGL(k_cfl_final, 1, 256, 0, c->stream, c->d_cflpart, c->cfl_blocks, c->d_dt, hist_slot, cfl_dr);
  return hipGetLastError();
}

hipError_t launch_set_dt(qas_ctx* c, double dt)
{
  hipLaunchKernelGGL(k_set_dt, 1, 1, 0, c->stream, c->d_dt, dt);
  return hipGetLastError();
}

hipError_t launch_record_dt(qas_ctx* c, double* slot)
{
  hipLaunchKernelGGL(k_copy1, 1, 1, 0, c->stream, c->d_dt, slot);
  return hipGetLastError();
}

hipError_t launch_record_counted(qas_ctx* c)
{
  hipLaunchKernelGGL(k_record, 1, 1, 0, c->stream, c->d_dt, c->d_hist, c->d_hcount, (long long)c->hist_cap);
  return hipGetLastError();
}

__global__ void k_zero_ll(long long* p) { *p = 0; }

hipError_t launch_zero_counter(qas_ctx* c)
{
  hipLaunchKernelGGL(k_zero_ll, 1, 1, 0, c->stream, c->d_hcount);
  return hipGetLastError();
}

}  // namespace qas
