// qas_testpoints.hip -- batched device evaluation of the per-point
// functions, for checking them against vectors written by the reference's
// own functions (common/common_interp.c:7-150, common/common_interp.h:10-24,
// common/mhd_func.c:56-76 and 108-161, common/mhd_flux.c:7-698).
//
// Each entry point copies a host batch to the device, runs one thread per
// item through exactly the device functions the stage kernels use
// (qas_device.hpp, FP64, -ffp-contract=off), and copies the results back.
// Test-only: every call allocates and frees its own device buffers.
#include <vector>

#include "../../include/qasmhd.h"
#include "qas_device.hpp"
#include "qas_stage2d.hpp"

namespace qas {

template <int ODR, bool LIN>
__global__ void k_t_interp(int n, const double* f5, double* fl, double* fr)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double f[5], a, b;
  for (int m = 0; m < 5; m++) f[m] = f5[5 * q + m];
  if (LIN) lin5<ODR>(f, a, b);
  else interp5<ODR>(f, a, b);
  fl[q] = a;
  fr[q] = b;
}

template <int ODR, bool DF>
__global__ void k_t_f4(int n, const double* f4, double* out)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double* f = f4 + 4 * q;
  out[q] = DF ? dfc<ODR>(f[0], f[1], f[2], f[3]) : fcen<ODR>(f[0], f[1], f[2], f[3]);
}

template <int ODR>
__global__ void k_t_lrstate(int n, const double* st, const double* bn, double gam, double* vl, double* vr)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double s[7][5], l[7], r[7];
  for (int k = 0; k < 7; k++)
    for (int m = 0; m < 5; m++) s[k][m] = st[35 * (long long)q + 5 * k + m];
  lrstate<ODR>(s, bn[q], gam, l, r);
  for (int k = 0; k < 7; k++) {
    vl[7 * (long long)q + k] = l[k];
    vr[7 * (long long)q + k] = r[k];
  }
}

template <int RMN>
__global__ void k_t_flux(int n, const double* ql, const double* qr, const double* bn, double gam, const double* dv,
                         double* f)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double L[7], R[7], F[7];
  for (int k = 0; k < 7; k++) {
    L[k] = ql[7 * (long long)q + k];
    R[k] = qr[7 * (long long)q + k];
  }
  riemann<RMN>(L, R, bn[q], gam, dv[2 * q], dv[2 * q + 1], F);
  for (int k = 0; k < 7; k++) f[7 * (long long)q + k] = F[k];
}

__global__ void k_t_cuct(int n, const double* c, const double* b, double* out)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double* x = c + 12 * q;
  const double* y = b + 4 * q;
  out[q] = cuct_weight(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], y[0], y[1], y[2],
                       y[3]);
}

// rcp_inrange(d) and the IEEE division 1.0 / d side by side
__global__ void k_t_rcp(int n, const double* d, double* fast, double* div)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double x = d[q];
  fast[q] = rcp_inrange(x);
  div[q] = 1.0 / x;
}

// device copies of host arrays for one call
struct Batch {
  std::vector<void*> bufs;
  hipError_t err = hipSuccess;
  double* in(const double* h, size_t n)
  {
    void* d = nullptr;
    if (err == hipSuccess) err = hipMalloc(&d, n * sizeof(double) + 8);
    if (err == hipSuccess) err = hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice);
    bufs.push_back(d);
    return (double*)d;
  }
  double* out(size_t n)
  {
    void* d = nullptr;
    if (err == hipSuccess) err = hipMalloc(&d, n * sizeof(double) + 8);
    bufs.push_back(d);
    return (double*)d;
  }
  void get(double* h, const double* d, size_t n)
  {
    if (err == hipSuccess) err = hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost);
  }
  int finish()
  {
    if (err == hipSuccess) err = hipDeviceSynchronize();
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    return err == hipSuccess ? QAS_OK : QAS_E_HIP;
  }
};

static inline unsigned nb(int n) { return (unsigned)((n + 127) / 128); }

}  // namespace qas

using namespace qas;

extern "C" int qas_test_interp(int kind, int n, const double* f5, double* fl, double* fr)
{
  if (n <= 0 || !f5 || !fl || !fr) return QAS_E_ARG;
  const bool lin = kind >= 10;
  const int odr = (lin ? kind - 10 : kind) + 1;
  if (odr < 1 || odr > 4) return QAS_E_ARG;
  Batch b;
  const double* d5 = b.in(f5, 5 * (size_t)n);
  double *dl = b.out(n), *dr = b.out(n);
  if (b.err == hipSuccess) {
#define QAS_TI(O)                                                                        \
  case O:                                                                                \
    if (lin) hipLaunchKernelGGL((k_t_interp<O, true>), nb(n), 128, 0, 0, n, d5, dl, dr); \
    else hipLaunchKernelGGL((k_t_interp<O, false>), nb(n), 128, 0, 0, n, d5, dl, dr);    \
    break;
    switch (odr) { QAS_TI(1) QAS_TI(2) QAS_TI(3) QAS_TI(4) }
#undef QAS_TI
    b.err = hipGetLastError();
  }
  b.get(fl, dl, n);
  b.get(fr, dr, n);
  return b.finish();
}

extern "C" int qas_test_f4(int op, int odr, int n, const double* f4, double* out)
{
  if (n <= 0 || !f4 || !out || odr < 1 || odr > 4 || op < 0 || op > 1) return QAS_E_ARG;
  Batch b;
  const double* d4 = b.in(f4, 4 * (size_t)n);
  double* d = b.out(n);
  if (b.err == hipSuccess) {
#define QAS_TF(O)                                                                   \
  case O:                                                                           \
    if (op) hipLaunchKernelGGL((k_t_f4<O, true>), nb(n), 128, 0, 0, n, d4, d);      \
    else hipLaunchKernelGGL((k_t_f4<O, false>), nb(n), 128, 0, 0, n, d4, d);        \
    break;
    switch (odr) { QAS_TF(1) QAS_TF(2) QAS_TF(3) QAS_TF(4) }
#undef QAS_TF
    b.err = hipGetLastError();
  }
  b.get(out, d, n);
  return b.finish();
}

extern "C" int qas_test_lrstate(int odr, int n, const double* st, const double* bn, double gam, double* vl,
                                double* vr)
{
  if (n <= 0 || !st || !bn || !vl || !vr || odr < 1 || odr > 4) return QAS_E_ARG;
  Batch b;
  const double* ds = b.in(st, 35 * (size_t)n);
  const double* db = b.in(bn, n);
  double *dl = b.out(7 * (size_t)n), *dr = b.out(7 * (size_t)n);
  if (b.err == hipSuccess) {
    switch (odr) {
      case 1: hipLaunchKernelGGL((k_t_lrstate<1>), nb(n), 128, 0, 0, n, ds, db, gam, dl, dr); break;
      case 2: hipLaunchKernelGGL((k_t_lrstate<2>), nb(n), 128, 0, 0, n, ds, db, gam, dl, dr); break;
      case 3: hipLaunchKernelGGL((k_t_lrstate<3>), nb(n), 128, 0, 0, n, ds, db, gam, dl, dr); break;
      default: hipLaunchKernelGGL((k_t_lrstate<4>), nb(n), 128, 0, 0, n, ds, db, gam, dl, dr); break;
    }
    b.err = hipGetLastError();
  }
  b.get(vl, dl, 7 * (size_t)n);
  b.get(vr, dr, 7 * (size_t)n);
  return b.finish();
}

extern "C" int qas_test_flux(int rmn, int n, const double* ql, const double* qr, const double* bn, double gam,
                             const double* dv, double* f)
{
  if (n <= 0 || !ql || !qr || !bn || !dv || !f || rmn < 0 || rmn > 3) return QAS_E_ARG;
  Batch b;
  const double* dl = b.in(ql, 7 * (size_t)n);
  const double* dr = b.in(qr, 7 * (size_t)n);
  const double* db = b.in(bn, n);
  const double* dd = b.in(dv, 2 * (size_t)n);
  double* df = b.out(7 * (size_t)n);
  if (b.err == hipSuccess) {
    switch (rmn) {
      case 0: hipLaunchKernelGGL((k_t_flux<0>), nb(n), 128, 0, 0, n, dl, dr, db, gam, dd, df); break;
      case 1: hipLaunchKernelGGL((k_t_flux<1>), nb(n), 128, 0, 0, n, dl, dr, db, gam, dd, df); break;
      case 2: hipLaunchKernelGGL((k_t_flux<2>), nb(n), 128, 0, 0, n, dl, dr, db, gam, dd, df); break;
      default: hipLaunchKernelGGL((k_t_flux<3>), nb(n), 128, 0, 0, n, dl, dr, db, gam, dd, df); break;
    }
    b.err = hipGetLastError();
  }
  b.get(f, df, 7 * (size_t)n);
  return b.finish();
}

extern "C" int qas_test_cuct(int n, const double* cells, const double* faces, double* out)
{
  if (n <= 0 || !cells || !faces || !out) return QAS_E_ARG;
  Batch b;
  const double* dc = b.in(cells, 12 * (size_t)n);
  const double* df = b.in(faces, 4 * (size_t)n);
  double* d = b.out(n);
  if (b.err == hipSuccess) {
    hipLaunchKernelGGL(k_t_cuct, nb(n), 128, 0, 0, n, dc, df, d);
    b.err = hipGetLastError();
  }
  b.get(out, d, n);
  return b.finish();
}

extern "C" int qas_test_rcp(int n, const double* d, double* fast, double* div)
{
  if (n <= 0 || !d || !fast || !div) return QAS_E_ARG;
  Batch b;
  const double* dd = b.in(d, n);
  double *df = b.out(n), *dv = b.out(n);
  if (b.err == hipSuccess) {
    hipLaunchKernelGGL(k_t_rcp, nb(n), 128, 0, 0, n, dd, df, dv);
    b.err = hipGetLastError();
  }
  b.get(fast, df, n);
  b.get(div, dv, n);
  return b.finish();
}

// ---------------------------------------------------------------- host runs
// The 2D boundary fill and the CT launch with the fused ghost fill, run on
// the host through the same __host__ __device__ functions the kernels use,
// one "thread" after another.  dn, st: 2 x 8 boundary flags (x, y); the
// arrays are 8 x (nx * ny) doubles, x fastest.
namespace {
qas::Geo host_geo(int nx, int ny)
{
  qas::Geo g{};
  g.ndim = 2;
  g.nx = nx;
  g.ny = ny;
  g.nz = 1;
  g.sy = nx;
  g.sz = (long long)nx * ny;
  g.nd = (long long)nx * ny;
  g.dx = g.dy = g.dz = 1.0;
  return g;
}
qas::BcFlags host_bc(const int* dn, const int* st)
{
  qas::BcFlags bc{};
  for (int a = 0; a < 2; a++)
    for (int m = 0; m < 8; m++) {
      bc.dn[a][m] = (signed char)dn[8 * a + m];
      bc.st[a][m] = (signed char)st[8 * a + m];
    }
  return bc;
}
bool host_args_ok(int nx, int ny, const int* dn, const int* st)
{
  if (nx < 2 * qas::NG + 8 || ny < 2 * qas::NG + 8 || !dn || !st) return false;
  for (int k = 0; k < 16; k++)
    if (dn[k] < -2 || dn[k] > 2 || st[k] < 0 || st[k] > 1) return false;
  return true;
}
}  // namespace

extern "C" int qas_test_bound_host(int nx, int ny, const int* dn, const int* st, double* f)
{
  if (!host_args_ok(nx, ny, dn, st) || !f) return QAS_E_ARG;
  const qas::Geo g = host_geo(nx, ny);
  const qas::BcFlags bc = host_bc(dn, st);
  const long long nd = g.nd;
  const long long nghost = nd - (long long)(nx - 2 * qas::NG) * (ny - 2 * qas::NG);
  for (long long t = 0; t < 8 * nghost; t++)
    qas::bound_at_f([&](int m) { return f + m * nd; }, 8, nullptr, 0.0, g, bc, nghost, qas::NG, qas::NG, 0, t);
  return QAS_OK;
}

extern "C" int qas_test_ctb2d_host(int nx, int ny, const int* dn, const int* st, const double* in, const double* u0,
                                   double* out, double a, double b, int alias, int ghosts_first)
{
  if (!host_args_ok(nx, ny, dn, st) || !in || !out || (!u0 && !alias)) return QAS_E_ARG;
  const qas::Geo g = host_geo(nx, ny);
  const qas::BcFlags bc = host_bc(dn, st);
  const long long nd = g.nd;
  std::vector<double> ez((size_t)nd, 0.0);
  const double dt = 0.5;
  qas::StagePtrs p{};
  for (int m = 0; m < 8; m++) {
    p.in[m] = in + m * nd;
    p.out[m] = out + m * nd;
    p.u0[m] = alias ? out + m * nd : u0 + m * nd;
  }
  p.ez = ez.data();
  p.dt = &dt;
  p.a = a;
  p.b = b;
  const long long nc = (long long)(nx - 7) * (ny - 7);
  const long long nghost = nd - (long long)(nx - 2 * qas::NG) * (ny - 2 * qas::NG);
  for (int part = 0; part < 2; part++) {
    const bool ct = (part == 0) != (ghosts_first != 0);
    const long long n = ct ? nc : 8 * nghost;
    for (long long t = 0; t < n; t++) qas::ctb2d_at<4>(p, g, bc, nullptr, 0.0, nghost, t, ct);
  }
  return QAS_OK;
}
