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
#include "qas_common.hpp"
#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

__global__ __launch_bounds__(256) void k_bound(BoundArgs a, Geo g, BcFlags bc, long long nghost, int ox, int oy,
                                               int oz)
{
  bound_at(a, g, bc, nghost, ox, oy, oz, (long long)blockIdx.x * blockDim.x + threadIdx.x);
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
  if (c->d_phi && nvar == 8) {
    a.phi = c->d_phi;
    a.gfac = (2.0 - c->s.gam) / (c->s.gam - 1.0);
  }
  const long long tot = nghost * nvar;
  QAS_LAUNCH(c, QAS_K_BOUND, k_bound, (unsigned)((tot + 255) / 256), 256, 0, c->stream, a, g, c->bc, nghost, ox, oy, oz);
  return hipGetLastError();
}

// ------------------------------------------------------------------ CFL
__global__ __launch_bounds__(256) void k_cfl(CflArgs a, Geo g) { cfl_block(a, g, (int)blockIdx.x, (int)gridDim.x); }

__global__ __launch_bounds__(256) void k_cfl_final(const double* part, int nb, double* dt, double* hist_slot,
                                                   double cfl_dr)
{
  cfl_final_block(part, nb, dt, hist_slot, cfl_dr);
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
  if (c->d_phi) {
    a.phi = c->d_phi;
    a.enw = const_cast<double*>(U[7]);
    c->bounded = false;  // the interior energy is rewritten: its ghosts are stale
  }
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
