// qas_kernels1d.hip -- one SSP-RK stage of the 1D ideal-MHD step on gfx950.
//
// Replaces the stage body of MHD1D::ideal (1D/SERIAL/shock/mhd1d_solve.cpp:
// 99-184): primitives at every cell (:104-112), then one sweep that
// reconstructs each cell (:113-141), solves each face (:142-162) and applies
// the 4th-order difference and RK stage sum to the seven evolved variables
// (:163-179).  In 1D the cell-centred B is the stored B (cx=bx, cy=by,
// cz=bz, :89-91) and bx is not evolved (the prim kernel copies it into
// the stage output buffer).  Lane layout as in the 2D sweeps.
#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

__global__ __launch_bounds__(256) void k_prim1d(StagePtrs p, Geo g, double gam)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nd) return;
  const double ro = p.in[0][t];
  const double iro = 1.0 / ro;
  const double vx = p.in[1][t] * iro;
  const double vy = p.in[2][t] * iro;
  const double vz = p.in[3][t] * iro;
  const double bx = p.in[4][t], by = p.in[5][t], bz = p.in[6][t];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double b2 = bx * bx + by * by + bz * bz;
  p.pr[t] = (gam - 1) * (p.in[7][t] - 0.5 * (ro * v2 + b2));
  p.vx[t] = vx;
  p.vy[t] = vy;
  p.vz[t] = vz;
  // bx is not evolved in 1D (the reference leaves val[4] in place); the
  // stage output is another buffer, so carry it over at every cell
  p.out[4][t] = bx;
}

template <int RMN, int ODR>
__global__ __launch_bounds__(256) void k_sweep1d(StagePtrs p, Geo g, double gam, int nseg)
{
  const int lane = threadIdx.x & (WAVE - 1);
  const int seg = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
  if (seg >= nseg) return;
  const int n = g.nx;
  const int f0 = 3 + seg * SEG;
  const int q = f0 - 2 + lane;
  const int qc = q < 2 ? 2 : (q > n - 3 ? n - 3 : q);
  const double* bx = p.in[4];
  double s[7][5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int a = qc + m - 2;
    s[0][m] = p.in[0][a];
    s[1][m] = p.vx[a];
    s[2][m] = p.vy[a];
    s[3][m] = p.vz[a];
    s[4][m] = p.in[5][a];
    s[5][m] = p.in[6][a];
    s[6][m] = p.pr[a];
  }
  double vl[7], vr[7];
  lrstate<ODR>(s, bx[qc], gam, vl, vr);
  double L[7], F[7];
#pragma unroll
  for (int k = 0; k < 7; k++) L[k] = __shfl_up(vl[k], 1, WAVE);
  const int qf = q < 3 ? 3 : (q > n - 3 ? n - 3 : q);
  const double bnf = 0.5 * (bx[qf] + bx[qf]);
  const double dv0 = p.vx[qf] - p.vx[qf - 1];
  riemann<RMN>(L, vr, bnf, gam, dv0, 0.0, F);
  // evolved: ro, mx, my, mz, by, bz, en = F0..F6
  double du[7];
#pragma unroll
  for (int v = 0; v < 7; v++) {
    const double fm1 = __shfl_up(F[v], 1, WAVE);
    const double f1 = __shfl_down(F[v], 1, WAVE);
    const double f2 = __shfl_down(F[v], 2, WAVE);
    du[v] = dfc<ODR>(fm1, F[v], f1, f2);
  }
  const bool upd = (lane >= 2) && (lane < 2 + SEG) && (q >= NG) && (q <= n - NG - 1);
  if (!upd) return;
  const double dtdx = *p.dt / g.dx;
  const int vid[7] = {0, 1, 2, 3, 5, 6, 7};
#pragma unroll
  for (int v = 0; v < 7; v++) {
    const int m = vid[v];
    p.out[m][q] = rk_upd(p.in[m][q], p.u0[m][q], -dtdx * du[v], p.a, p.b);
  }
}

typedef void (*sweep1_fn)(StagePtrs, Geo, double, int);

static sweep1_fn pick1(int rmn, int odr)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return k_sweep1d<R, O>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

hipError_t launch_stage1d(qas_ctx* c, const StagePtrs& p)
{
  const Geo& g = c->geo;
  QAS_LAUNCH(c, QAS_K_PRIM, k_prim1d, (unsigned)((g.nd + 255) / 256), 256, 0, c->stream, p, g, c->s.gam);
  const int nseg = (g.nx - 5 + SEG - 1) / SEG;
  sweep1_fn f = pick1(c->s.rmn, c->s.odr);
  if (!f) return hipErrorInvalidValue;
  QAS_LAUNCH(c, QAS_K_SWEEP_X, f, (unsigned)((nseg * WAVE + 255) / 256), 256, 0, c->stream, p, g, c->s.gam, nseg);
  return hipGetLastError();
}

}  // namespace qas
