// qas_dissipative.hip -- the visco-resistive step of the 2D dissipative
// driver: DMHD2D::dsptv (2D/SERIAL/MRX/dmhd2d_solve.cpp:3-151) and its
// sub-cycling after ideal() in DMHD2D::exec_ (dmhd2d_class.cpp:15-33).
//
// One SSP-RK stage is three launches plus the ghost fill of its output:
//   k_dsp_edge  every cell: v = m (1/ro) (vlcty, common/mhd_class.hpp:59-66);
//               CT corners [2, n-2]^2: the resistive field eta j
//               (mhd_ct_eres, common/mhd_func.c:78-106)
//   k_dsp_flux  viscous + resistive fluxes (calc_flux_viscous /
//               calc_flux_resistive, common/mhd_flux.c:700-782) at the x faces
//               i in [3, nx-3] of interior rows and the y faces j in
//               [3, ny-3] of interior columns -- the faces the update reads
//   k_dsp_upd   interior cells: divergence + RK sum of ro mx my mz bz en
//               (mhd_updt2d, mhd_func.c:222-248); kept faces: CT update of
//               bx, by from Ez (mhd_updt2d_ctb, mhd_func.c:286-300)
// The stage buffers rotate as in ideal() (qas_capi.hip), so the reference's
// copy of the stage-0 state (`ut`) is the step base read in place.  Every
// expression keeps the reference's operand order, including the additions
// of the zero flux components, so results are bit-identical.
#include <cstring>
#include <string>

#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

struct DspPtrs {
  const double* in[8];
  double* out[8];
  const double* u0[8];
  double *vx, *vy, *vz;
  double *ex, *ey, *ez;
  double* fx[5];  // x-face fluxes of mx my mz bz en
  double* fy[5];  // y-face fluxes of mx my mz bz en
  const double *nu, *eta;
  double idx, idy, dtdx, dtdy, a, b;
};

// calc_flux_viscous (mhd_flux.c:700-753) at face i-1/2 of the cell at s
template <int ODR>
__device__ __forceinline__ void visc_flux(const double* vx, const double* vy, const double* vz, long long s,
                                          double rnu, double idx, double idy, double idz, long long xo,
                                          long long yo, long long zo, double& fmx, double& fmy, double& fmz,
                                          double& fen)
{
  const long long ss[4] = {s - 2 * xo, s - xo, s, s + xo};
  const double c2 = 2. / 3., c4 = 4. / 3.;
  const double vx0 = fcen<ODR>(vx[ss[0]], vx[ss[1]], vx[ss[2]], vx[ss[3]]);
  const double vy0 = fcen<ODR>(vy[ss[0]], vy[ss[1]], vy[ss[2]], vy[ss[3]]);
  const double vz0 = fcen<ODR>(vz[ss[0]], vz[ss[1]], vz[ss[2]], vz[ss[3]]);
  const double dvxdx = dfc<ODR>(vx[ss[0]], vx[ss[1]], vx[ss[2]], vx[ss[3]]) * idx;
  const double dvydx = dfc<ODR>(vy[ss[0]], vy[ss[1]], vy[ss[2]], vy[ss[3]]) * idx;
  const double dvzdx = dfc<ODR>(vz[ss[0]], vz[ss[1]], vz[ss[2]], vz[ss[3]]) * idx;
  double dv[4];
#pragma unroll
  for (int k = 0; k < 4; k++) dv[k] = 0.5 * (vx[ss[k] + yo] - vx[ss[k] - yo]);
  const double dvxdy = fcen<ODR>(dv[0], dv[1], dv[2], dv[3]) * idy;
#pragma unroll
  for (int k = 0; k < 4; k++) dv[k] = 0.5 * (vy[ss[k] + yo] - vy[ss[k] - yo]);
  const double dvydy = fcen<ODR>(dv[0], dv[1], dv[2], dv[3]) * idy;
#pragma unroll
  for (int k = 0; k < 4; k++) dv[k] = 0.5 * (vx[ss[k] + zo] - vx[ss[k] - zo]);
  const double dvxdz = fcen<ODR>(dv[0], dv[1], dv[2], dv[3]) * idz;
#pragma unroll
  for (int k = 0; k < 4; k++) dv[k] = 0.5 * (vz[ss[k] + zo] - vz[ss[k] - zo]);
  const double dvzdz = fcen<ODR>(dv[0], dv[1], dv[2], dv[3]) * idz;
  fmx = -rnu * (c4 * dvxdx - c2 * (dvydy + dvzdz));
  fmy = -rnu * (dvydx + dvxdy);
  fmz = -rnu * (dvzdx + dvxdz);
  fen = fmx * vx0 + fmy * vy0 + fmz * vz0;
}

// calc_flux_resistive (mhd_flux.c:755-782) at face i-1/2 of the cell at s
template <int ODR>
__device__ __forceinline__ void res_flux(const double* by, const double* bz, const double* ey, const double* ez,
                                         long long s, long long xo, long long yo, long long zo, double& fby,
                                         double& fbz, double& fen)
{
  const long long ss[4] = {s - 2 * xo, s - xo, s, s + xo};
  const double ey0 = 0.5 * (ey[s] + ey[s + zo]);
  const double ez0 = 0.5 * (ez[s] + ez[s + yo]);
  double byi[4], bzi[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    byi[k] = 0.5 * (by[ss[k]] + by[ss[k] + yo]);
    bzi[k] = 0.5 * (bz[ss[k]] + bz[ss[k] + zo]);
  }
  const double by0 = fcen<ODR>(byi[0], byi[1], byi[2], byi[3]);
  const double bz0 = fcen<ODR>(bzi[0], bzi[1], bzi[2], bzi[3]);
  fby = -ez0;
  fbz = +ey0;
  fen = ey0 * bz0 - ez0 * by0;
}

template <int ODR>
__global__ __launch_bounds__(256) void k_dsp_edge(DspPtrs p, Geo g)
{
  const long long t = (long long)xcd_remap((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= g.nd) return;
  const double iro = 1.0 / p.in[0][t];
  p.vx[t] = p.in[1][t] * iro;
  p.vy[t] = p.in[2][t] * iro;
  p.vz[t] = p.in[3][t] * iro;
  const int i = (int)(t % g.nx), j = (int)(t / g.nx);
  if (i < 2 || i > g.nx - 2 || j < 2 || j > g.ny - 2) return;
  // mhd_ct_eres with (idx, idy, idz) = (idx, idy, 0), offsets (1, nx, 0)
  const double* bx = p.in[4];
  const double* by = p.in[5];
  const double* bz = p.in[6];
  const double* eta = p.eta;
  const long long nx = g.nx;
  const double idx = p.idx, idy = p.idy, idz = 0.0;
  const double dbzj = dfc<ODR>(bz[t - 2 * nx], bz[t - nx], bz[t], bz[t + nx]);
  const double dbyk = dfc<ODR>(by[t], by[t], by[t], by[t]);
  const double dbxk = dfc<ODR>(bx[t], bx[t], bx[t], bx[t]);
  const double dbzi = dfc<ODR>(bz[t - 2], bz[t - 1], bz[t], bz[t + 1]);
  const double dbyi = dfc<ODR>(by[t - 2], by[t - 1], by[t], by[t + 1]);
  const double dbxj = dfc<ODR>(bx[t - 2 * nx], bx[t - nx], bx[t], bx[t + nx]);
  const double jx = dbzj * idy - dbyk * idz;
  const double jy = dbxk * idz - dbzi * idx;
  const double jz = dbyi * idx - dbxj * idy;
  p.ex[t] = 0.25 * (eta[t - nx] + eta[t] + eta[t - nx] + eta[t]) * jx;
  p.ey[t] = 0.25 * (eta[t - 1] + eta[t - 1] + eta[t] + eta[t]) * jy;
  p.ez[t] = 0.25 * (eta[t - 1 - nx] + eta[t - nx] + eta[t - 1] + eta[t]) * jz;
}

template <int ODR>
__global__ __launch_bounds__(256) void k_dsp_flux(DspPtrs p, Geo g)
{
  // cells (i, j) in [3, nx-3] x [3, ny-3]
  const int ni = g.nx - 5, nj = g.ny - 5;
  const long long t = (long long)xcd_remap((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj) return;
  const int i = 3 + (int)(t % ni), j = 3 + (int)(t / ni);
  const long long nx = g.nx;
  const long long s = nx * j + i;
  const double* ro = p.in[0];
  const double* nu = p.nu;
  if (j >= NG && j <= g.ny - NG - 1) {  // x face i-1/2
    const double rnu = 0.5 * (ro[s - 1] * nu[s - 1] + ro[s] * nu[s]);
    double fmx, fmy, fmz, fenv, fby, fbz, fenr;
    visc_flux<ODR>(p.vx, p.vy, p.vz, s, rnu, p.idx, p.idy, 0.0, 1, nx, 0, fmx, fmy, fmz, fenv);
    res_flux<ODR>(p.in[5], p.in[6], p.ey, p.ez, s, 1, nx, 0, fby, fbz, fenr);
    (void)fby;
    p.fx[0][s] = fmx + 0.0;
    p.fx[1][s] = fmy + 0.0;
    p.fx[2][s] = fmz + 0.0;
    p.fx[3][s] = 0.0 + fbz;
    p.fx[4][s] = fenv + fenr;
  }
  if (i >= NG && i <= g.nx - NG - 1) {  // y face j-1/2: the reference's call with the axes rotated
    const double rnu = 0.5 * (ro[s - nx] * nu[s - nx] + ro[s] * nu[s]);
    double fmy, fmz, fmx, fenv, fbz, fbx, fenr;
    visc_flux<ODR>(p.vy, p.vz, p.vx, s, rnu, p.idy, 0.0, p.idx, nx, 0, 1, fmy, fmz, fmx, fenv);
    res_flux<ODR>(p.in[6], p.in[4], p.ez, p.ex, s, nx, 0, 1, fbz, fbx, fenr);
    (void)fbx;
    p.fy[0][s] = fmx + 0.0;
    p.fy[1][s] = fmy + 0.0;
    p.fy[2][s] = fmz + 0.0;
    p.fy[3][s] = 0.0 + fbz;
    p.fy[4][s] = fenv + fenr;
  }
}

template <int ODR>
__global__ __launch_bounds__(256) void k_dsp_upd(DspPtrs p, Geo g)
{
  // (i, j) in [4, nx-4] x [4, ny-4]
  const int ni = g.nx - 7, nj = g.ny - 7;
  const long long t = (long long)xcd_remap((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj) return;
  const int i = NG + (int)(t % ni), j = NG + (int)(t / ni);
  const long long nx = g.nx;
  const long long s = nx * j + i;
  const bool ci = i <= g.nx - NG - 1, cj = j <= g.ny - NG - 1;
  if (ci && cj) {
    // ro: both fluxes are 0.0 + 0.0
    {
      const double z = 0.0 + 0.0;
      const double du0 = dfc<ODR>(z, z, z, z), du1 = dfc<ODR>(z, z, z, z);
      p.out[0][s] = rk_upd(p.in[0][s], p.u0[0][s], (-p.dtdx * du0 - p.dtdy * du1), p.a, p.b);
    }
    const int vm[5] = {1, 2, 3, 6, 7};
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const double* fx = p.fx[v];
      const double* fy = p.fy[v];
      const double du0 = dfc<ODR>(fx[s - 1], fx[s], fx[s + 1], fx[s + 2]);
      const double du1 = dfc<ODR>(fy[s - nx], fy[s], fy[s + nx], fy[s + 2 * nx]);
      const int m = vm[v];
      p.out[m][s] = rk_upd(p.in[m][s], p.u0[m][s], (-p.dtdx * du0 - p.dtdy * du1), p.a, p.b);
    }
  }
  const double* ez = p.ez;
  if (cj) {  // bx at face i-1/2, i in [4, nx-4]
    const double du = dfc<ODR>(ez[s - nx], ez[s], ez[s + nx], ez[s + 2 * nx]);
    const double d = +p.dtdy;
    p.out[4][s] = rk_upd(p.in[4][s], p.u0[4][s], -d * du, p.a, p.b);
  }
  if (ci) {  // by at face j-1/2, j in [4, ny-4]
    const double du = dfc<ODR>(ez[s - 1], ez[s], ez[s + 1], ez[s + 2]);
    const double d = -p.dtdx;
    p.out[5][s] = rk_upd(p.in[5][s], p.u0[5][s], -d * du, p.a, p.b);
  }
}

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

template <int ODR>
static hipError_t launch_dsp_stage(qas_ctx* c, const DspPtrs& p)
{
  const Geo& g = c->geo;
  hipStream_t st = c->stream;
  hipLaunchKernelGGL((k_dsp_edge<ODR>), nblk(g.nd, 256), 256, 0, st, p, g);
  hipLaunchKernelGGL((k_dsp_flux<ODR>), nblk((long long)(g.nx - 5) * (g.ny - 5), 256), 256, 0, st, p, g);
  hipLaunchKernelGGL((k_dsp_upd<ODR>), nblk((long long)(g.nx - 7) * (g.ny - 7), 256), 256, 0, st, p, g);
  return hipGetLastError();
}

}  // namespace qas

using namespace qas;

static int dfail(qas_ctx* c, int code, const std::string& m)
{
  if (c) c->err = m;
  return code;
}

#define QAS_DCHECK(c, expr)                                                                               \
  do {                                                                                                    \
    hipError_t _e = (expr);                                                                               \
    if (_e != hipSuccess) return dfail((c), QAS_E_HIP, std::string(#expr) + ": " + hipGetErrorName(_e)); \
  } while (0)

namespace qas {
void dsp_release(qas_ctx* c)
{
  if (c && c->d_nu) {
    (void)hipFree(c->d_nu);
    c->d_nu = nullptr;
    c->d_eta = nullptr;
  }
}

// one DMHD2D::dsptv(dt) call on the current state (bounded on entry); the
// stage buffers rotate like ideal()'s
int dsp_step(qas_ctx* c, double dt)
{
  const int R = c->s.r_k;
  const int A = c->cur, B = (c->cur + 1) % 3, C = (c->cur + 2) % 3;
  int seq_in[3], seq_out[3];
  if (R == 1) { seq_in[0] = A; seq_out[0] = B; }
  else if (R == 2) { seq_in[0] = A; seq_out[0] = B; seq_in[1] = B; seq_out[1] = A; }
  else { seq_in[0] = A; seq_out[0] = B; seq_in[1] = B; seq_out[1] = C; seq_in[2] = C; seq_out[2] = A; }
  DspPtrs p;
  std::memset(&p, 0, sizeof(p));
  p.vx = c->vx;
  p.vy = c->vy;
  p.vz = c->vz;
  p.ex = c->emf[6];
  p.ey = c->emf[7];
  p.ez = c->ez;
  for (int k = 0; k < 5; k++) {
    p.fx[k] = c->acc[k];
    p.fy[k] = c->emf[k];
  }
  p.nu = c->d_nu;
  p.eta = c->d_eta;
  // dmhd2d_solve.cpp:9-10
  p.idx = 1.0 / c->geo.dx;
  p.idy = 1.0 / c->geo.dy;
  p.dtdx = dt * p.idx;
  p.dtdy = dt * p.idy;
  for (int rk = 0; rk < R; rk++) {
    for (int m = 0; m < 8; m++) {
      p.in[m] = c->U[seq_in[rk]][m];
      p.out[m] = c->U[seq_out[rk]][m];
      p.u0[m] = c->U[A][m];
    }
    if (rk == 0) { p.a = 0.0; p.b = 1.0; }
    else if (rk == 1) { p.a = 0.5 + (R - 2) * 0.25; p.b = 0.5 - (R - 2) * 0.25; }
    else { p.a = 1. / 3.; p.b = 2. / 3.; }
    hipError_t e = c->s.odr <= 2 ? launch_dsp_stage<2>(c, p) : launch_dsp_stage<4>(c, p);
    if (e != hipSuccess) return dfail(c, QAS_E_HIP, std::string("dsptv stage launch: ") + hipGetErrorName(e));
    e = launch_bound(c, c->U[seq_out[rk]], 8);
    if (e != hipSuccess) return dfail(c, QAS_E_HIP, std::string("dsptv bound launch: ") + hipGetErrorName(e));
  }
  c->cur = seq_out[R - 1];
  c->bounded = true;
  return QAS_OK;
}

// the dissipative part of one DMHD2D::exec_ step after ideal(dt), with dt in
// device memory: ncmax = 1 + (int)(2 (dcmax dt) / dr^2) sub-steps of
// {bound; dsptv(dt / ncmax)} (dmhd2d_class.cpp:23-30)
int dsp_after_ideal(qas_ctx* c)
{
  if (!c->d_nu || !(c->dcmax > 1e-15)) return QAS_OK;
  QAS_DCHECK(c, hipMemcpyAsync(c->h_pinned, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  QAS_DCHECK(c, hipStreamSynchronize(c->stream));
  const double dt = c->h_pinned[0];
  const double dr = c->geo.dx < c->geo.dy ? c->geo.dx : c->geo.dy;  // min(dx, dy): common_func.h max/min
  const int ncmax = 1 + (int)(2 * (c->dcmax * dt) / (dr * dr));
  for (int nc = 0; nc < ncmax; nc++) {
    // bound() first, as exec_ does; the state is bounded after ideal()/dsptv()
    if (!c->bounded) QAS_DCHECK(c, launch_bound(c, c->U[c->cur], 8));
    int r = dsp_step(c, dt / ncmax);
    if (r) return r;
  }
  return QAS_OK;
}
}  // namespace qas

extern "C" int qas_set_dissipation(qas_ctx* c, const double* nu, const double* eta, double dcmax)
{
  if (!c || !c->slab) return dfail(c, QAS_E_ARG, "qas_set_dissipation: null context");
  if (!nu) {
    dsp_release(c);
    c->dcmax = 0.0;
    c->graph_flg = -1;  // the step graph (if any) must be rebuilt
    return QAS_OK;
  }
  if (!eta) return dfail(c, QAS_E_ARG, "qas_set_dissipation: eta is null");
  if (c->g.ndim != 2) return dfail(c, QAS_E_ARG, "qas_set_dissipation: the dissipative step covers 2D domains");
  if (c->s.precision != 0) return dfail(c, QAS_E_ARG, "qas_set_dissipation: the dissipative step is FP64 only");
  if (!(dcmax >= 0.0)) return dfail(c, QAS_E_ARG, "qas_set_dissipation: dcmax must be >= 0");
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  if (!c->d_nu) {
    double* d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * bytes);
    if (e != hipSuccess)
      return dfail(c, QAS_E_NOMEM, std::string("qas_set_dissipation: ") + hipGetErrorName(e));
    c->d_nu = d;
    c->d_eta = d + c->geo.nd;
  }
  QAS_DCHECK(c, hipMemcpyAsync(c->d_nu, nu, bytes, hipMemcpyHostToDevice, c->stream));
  QAS_DCHECK(c, hipMemcpyAsync(c->d_eta, eta, bytes, hipMemcpyHostToDevice, c->stream));
  QAS_DCHECK(c, hipStreamSynchronize(c->stream));
  c->dcmax = dcmax;
  c->graph_flg = -1;
  return QAS_OK;
}

extern "C" int qas_dsptv(qas_ctx* c, double dt)
{
  if (!c || !c->slab) return dfail(c, QAS_E_ARG, "qas_dsptv: null context");
  if (!c->uploaded) return dfail(c, QAS_E_STATE, "qas_dsptv: no state uploaded");
  if (!c->d_nu) return dfail(c, QAS_E_STATE, "qas_dsptv: no viscosity / resistivity set (qas_set_dissipation)");
  if (!(dt > 0.0)) return dfail(c, QAS_E_ARG, "qas_dsptv: dt must be positive");
  if (!c->bounded) {
    QAS_DCHECK(c, launch_bound(c, c->U[c->cur], 8));
    c->bounded = true;
  }
  return dsp_step(c, dt);
}
