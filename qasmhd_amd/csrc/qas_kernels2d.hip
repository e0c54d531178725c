// qas_kernels2d.hip -- one SSP-RK stage of the 2D ideal-MHD step on gfx950.
//
// Replaces the stage body of MHD2D::ideal (2D/SERIAL/OTvortex/
// mhd2d_solve.cpp:44-386).  A stage is five launches:
//   prim2d   cell-centred B (fcen, ghosts by the bound() index map) and
//            primitives at every cell                       (:51-89)
//   sweep<X> characteristic WCNS + Riemann flux per x face, for every row;
//            writes the x flux divergence (times -dt/dx) of the interior
//            and the upwind/central parts of the By flux    (:127-199)
//   sweep<Y> same along y; adds the y divergence and applies the RK stage
//            sum to ro, mx, my, mz, bz, en                  (:234-306, 341-358)
//   emf2d    CUCT corner EMF from both sweeps' face parts   (:91-103, 201-232, 308-339)
//   ct2d     staggered bx, by update                        (:359-380)
// followed by the bound() kernel (:384-385).
// A sweep assigns one 64-lane wavefront to 64 consecutive cells of a line;
// lane l reconstructs cell f0-2+l, neighbour states and fluxes move between
// lanes by cross-lane shuffles, and the wavefront owns the 60 faces/cells
// [f0, f0+60).  Everything runs in FP64 with the reference's operation
// order (the library is built with -ffp-contract=off).
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_sweep2d.hpp"

namespace qas {

// ------------------------------------------------------------------ prim
// PT: storage type of the primitives (double, or float for the mixed path);
// the arithmetic is FP64 either way.
template <int ODR, typename PT>
__global__ __launch_bounds__(256) void k_prim2d(StagePtrs p, Geo g, BcFlags bc, double gam)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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
  store_prim<PT>(p.pr, t, (gam - 1) * (p.in[7][t] - 0.5 * (ro * v2 + b2)));
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

// the face sweeps live in qas_sweep2d.hpp (shared with the mixed path)
// ------------------------------------------------------------------ EMF
template <int ODR>
__global__ __launch_bounds__(256) void k_emf2d(StagePtrs p, Geo g, int ctw)
{
  const int ni = g.nx - 5, nj = g.ny - 5;   // corners [3, n-3]
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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
    for (int m = 0; m < 5; m++) { a[m] = fu[ss + (m - 3) * nx]; b[m] = fcn[ss + (m - 3) * nx]; }
    lin5<ODR>(a, ul0, dum);
    lin5<ODR>(b, ul1, dum);
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = fu[ss + (m - 2) * nx]; b[m] = fcn[ss + (m - 2) * nx]; }
    lin5<ODR>(a, dum, ur0);
    lin5<ODR>(b, dum, ur1);
    ez += -0.5 * ((ul0 + ur0) + (1.0 - ct) * (ul1 + ur1));
  }
  {  // from the y sweep: Bx flux interpolated along x
    const double* fu = p.emf[4];
    const double* fcn = p.emf[5];
    double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = fu[ss + (m - 3)]; b[m] = fcn[ss + (m - 3)]; }
    lin5<ODR>(a, ul0, dum);
    lin5<ODR>(b, ul1, dum);
#pragma unroll
    for (int m = 0; m < 5; m++) { a[m] = fu[ss + (m - 2)]; b[m] = fcn[ss + (m - 2)]; }
    lin5<ODR>(a, dum, ur0);
    lin5<ODR>(b, dum, ur1);
    ez += +0.5 * ((ul0 + ur0) + ct * (ul1 + ur1));
  }
  p.ez[ss] = ez;
}

// ------------------------------------------------------------------ CT
template <int ODR>
__global__ __launch_bounds__(256) void k_ct2d(StagePtrs p, Geo g)
{
  // faces i, j in [4, n-4]: bx needs j <= ny-5, by needs i <= nx-5
  const int ni = g.nx - 7, nj = g.ny - 7;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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

// ------------------------------------------------------------------ launch
typedef void (*sweep_fn)(StagePtrs, Geo, BcFlags, double, int);

template <int DIR>
static sweep_fn pick_sweep(int rmn, int odr)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return DIR == 0 ? k_sweep2d<0, R, O, double> : k_sweep2d_y<R, O, double>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_stage2d(qas_ctx* c, const StagePtrs& p)
{
  const Geo& g = c->geo;
  const int odr = c->s.odr;
  const double gam = c->s.gam;
  hipStream_t st = c->stream;
  // 1) cell-centred B and primitives
  const bool f32 = c->s.precision == 1;
  if (odr <= 2) {
    if (f32) QAS_LAUNCH(c, QAS_K_PRIM, (k_prim2d<2, float>), nblk(g.nd, 256), 256, 0, st, p, g, c->bc, gam);
    else QAS_LAUNCH(c, QAS_K_PRIM, (k_prim2d<2, double>), nblk(g.nd, 256), 256, 0, st, p, g, c->bc, gam);
  } else {
    if (f32) QAS_LAUNCH(c, QAS_K_PRIM, (k_prim2d<4, float>), nblk(g.nd, 256), 256, 0, st, p, g, c->bc, gam);
    else QAS_LAUNCH(c, QAS_K_PRIM, (k_prim2d<4, double>), nblk(g.nd, 256), 256, 0, st, p, g, c->bc, gam);
  }
  // 2) x sweep, 3) y sweep
  const int nsx = (g.nx - 5 + SEG - 1) / SEG;
  const int nsy = (g.ny - 5 + SEG - 1) / SEG;
  const bool mixed = c->s.precision == 1;
  sweep_fn fx = mixed ? pick_sweep_f32(2, 0, c->s.rmn, odr) : pick_sweep<0>(c->s.rmn, odr);
  sweep_fn fy = mixed ? pick_sweep_f32(2, 1, c->s.rmn, odr) : pick_sweep<1>(c->s.rmn, odr);
  if (!fx || !fy) return hipErrorInvalidValue;
  QAS_LAUNCH(c, QAS_K_SWEEP_X, fx, nblk((long long)nsx * g.ny * WAVE, 256), 256, 0, st, p, g, c->bc, gam, nsx);
  // y sweep: one 256-thread workgroup per (4-column group, segment)
  QAS_LAUNCH(c, QAS_K_SWEEP_Y, fy, (unsigned)(((g.nx + YC - 1) / YC) * nsy), 256, 0, st, p, g, c->bc, gam, nsy);
  // 4) corner EMF, 5) CT
  const long long ne = (long long)(g.nx - 5) * (g.ny - 5);
  const long long nc = (long long)(g.nx - 7) * (g.ny - 7);
  switch (odr) {
    case 1: QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<1>), nblk(ne, 256), 256, 0, st, p, g, c->s.ctw);
            QAS_LAUNCH(c, QAS_K_CT, (k_ct2d<1>), nblk(nc, 256), 256, 0, st, p, g); break;
    case 2: QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<2>), nblk(ne, 256), 256, 0, st, p, g, c->s.ctw);
            QAS_LAUNCH(c, QAS_K_CT, (k_ct2d<2>), nblk(nc, 256), 256, 0, st, p, g); break;
    case 3: QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<3>), nblk(ne, 256), 256, 0, st, p, g, c->s.ctw);
            QAS_LAUNCH(c, QAS_K_CT, (k_ct2d<3>), nblk(nc, 256), 256, 0, st, p, g); break;
    default: QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<4>), nblk(ne, 256), 256, 0, st, p, g, c->s.ctw);
             QAS_LAUNCH(c, QAS_K_CT, (k_ct2d<4>), nblk(nc, 256), 256, 0, st, p, g); break;
  }
  return hipGetLastError();
}

}  // namespace qas
