// qas_kernels3d.hip -- one SSP-RK stage of the 3D ideal-MHD step on gfx950.
//
// Replaces the stage body of MHD3D::ideal (3D/MPI/OTvortex/mhd3d_solve.cpp:
// 42-715) on one device-resident domain:
//   sweep<d>  for d = z, y, x: characteristic WCNS + Riemann flux on every
//             line; z and y store dt/dz Dz and dt/dy Dy of ro, mx, my, mz,
//             en, the x sweep applies the RK stage sum with the reference's
//             ((-dtdx Dx) - dtdy Dy) - dtdz Dz; upwind/central parts of both
//             tangential B fluxes stored per face          (:166-258, 330-422, 494-586, 658-676);
//             the z sweep first computes the cell-centred B (fcen along x,
//             y, z; ghosts by the bound() index map) and primitives of its
//             LDS tile and stores them for the y and x sweeps (:49-98)
//   emf3d     Ex, Ey, Ez at cell edges from the face parts (CUCT)
//                                                            (:100-136, 259-328, 423-492, 587-656)
//   ct3d      staggered bx, by, bz update                    (:677-709)
// followed by the bound() kernel (:713-714).
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_sweep3d.hpp"

namespace qas {

// the face sweeps live in qas_sweep3d.hpp (shared with the mixed path)

// ------------------------------------------------------------------ EMF
// One contribution: interpolate the face parts (upwind fu, central fc) of a
// sweep along axis o to the edge and combine with weight w:
//   sgn * ((uL + uR) + w (cL + cR))
// (FT: element type of the face parts, the sweep precision)
template <int ODR, typename FT>
__device__ __forceinline__ double emf_part(const double* fu, const double* fc, long long ss, long long o, double w,
                                           double sgn)
{
  double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
  for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(fu, ss + (m - 3) * o); b[m] = load_as<FT>(fc, ss + (m - 3) * o); }
  lin5<ODR>(a, ul0, dum);
  lin5<ODR>(b, ul1, dum);
#pragma unroll
  for (int m = 0; m < 5; m++) { a[m] = load_as<FT>(fu, ss + (m - 2) * o); b[m] = load_as<FT>(fc, ss + (m - 2) * o); }
  lin5<ODR>(a, dum, ur0);
  lin5<ODR>(b, dum, ur1);
  return sgn * ((ul0 + ur0) + w * (ul1 + ur1));
}

// Edge EMFs Ex, Ey, Ez (CUCT, mhd3d_solve.cpp:100-136, 259-328, 423-492,
// 587-656) marching along z: one thread per (i, j) edge column of a chunk
// of EMF_KC planes.  The four z-interpolated face-part arrays (Ey's and
// Ex's contributions from the x and y sweeps' partner sweeps along z) are
// kept as 6-point register windows shifted one plane per step, and the cell
// values of plane k-1 the CUCT weights of Ex and Ey need are carried from
// the previous step, so each is read from memory once per point instead of
// 6 (resp. 2) times from planes far apart.  Per edge the operations and
// their order are the reference's.
// FT: the sweep precision.  FP64: the CUCT velocities are m (1/ro) of the
// state (the primitives' expression) and the edge EMFs are stored in FP64.
// Mixed (FT = float): the velocities are the FP32 primitives the z sweep
// stored, and the edge EMFs -- built from FP32 face parts -- are stored in
// FP32 for the CT kernel, which halves both kernels' traffic in them.
constexpr int EMF_KC = 32;     // planes per thread
constexpr int EMF_TX = 64;     // tile: 64 x 4 columns per 256-thread workgroup
constexpr int EMF_TY = 4;

template <int ODR>
__device__ __forceinline__ double emf_part_w(const double (&u)[6], const double (&c)[6], double w, double sgn)
{
  double a[5], b[5], ul0, ur0, ul1, ur1, dum;
#pragma unroll
  for (int m = 0; m < 5; m++) { a[m] = u[m]; b[m] = c[m]; }
  lin5<ODR>(a, ul0, dum);
  lin5<ODR>(b, ul1, dum);
#pragma unroll
  for (int m = 0; m < 5; m++) { a[m] = u[m + 1]; b[m] = c[m + 1]; }
  lin5<ODR>(a, dum, ur0);
  lin5<ODR>(b, dum, ur1);
  return sgn * ((ul0 + ur0) + w * (ul1 + ur1));
}

template <int ODR, typename FT>
__global__ __launch_bounds__(256) void k_emf3d_z(StagePtrs p, Geo g, int ctw)
{
  const int ni = g.nx - 5, nj = g.ny - 5, nk = g.nz - 5;  // edges on [3, n-3]^3
  const int ntx = (ni + EMF_TX - 1) / EMF_TX, nty = (nj + EMF_TY - 1) / EMF_TY;
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int tile = L % (ntx * nty), chunk = L / (ntx * nty);
  const int i = 3 + (tile % ntx) * EMF_TX + (int)(threadIdx.x % EMF_TX);
  const int j = 3 + (tile / ntx) * EMF_TY + (int)(threadIdx.x / EMF_TX);
  const int k0 = 3 + chunk * EMF_KC;
  if (i > g.nx - 3 || j > g.ny - 3 || k0 > g.nz - 3) return;  // no barriers below
  const int k1 = (k0 + EMF_KC < 3 + nk) ? k0 + EMF_KC : 3 + nk;
  const long long sy = g.sy, sz = g.sz;
  const double* ro = p.in[0];
  const double *mx = p.in[1], *my = p.in[2], *mz = p.in[3];
  const double *bx = p.in[4], *by = p.in[5], *bz = p.in[6];
  constexpr bool F32 = sizeof(FT) == 4;
  const float *fvx = reinterpret_cast<const float*>(p.vx), *fvy = reinterpret_cast<const float*>(p.vy),
              *fvz = reinterpret_cast<const float*>(p.vz);
  auto vel = [&](const double* m, const float* fv, long long c) -> double {
    if constexpr (F32) return (double)fv[c];
    else return m[c] * (1.0 / ro[c]);
  };
  // z windows at planes k-3 .. k+2
  double u2[6], c2[6], u4[6], c4[6];
  {
    const long long s0 = sz * k0 + sy * j + i;
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const long long a = s0 + (m - 3) * sz;
      u2[m] = load_as<FT>(p.emf[2], a);
      c2[m] = load_as<FT>(p.emf[3], a);
      u4[m] = load_as<FT>(p.emf[4], a);
      c4[m] = load_as<FT>(p.emf[5], a);
    }
  }
  // plane k0-1 cell values used by the Ex / Ey weights
  double ro_p = 0, ro_pi = 0, ro_pj = 0, vx_p = 0, vx_pi = 0, vy_p = 0, vy_pj = 0, vz_p = 0, vz_pi = 0, vz_pj = 0,
         bx_p = 0, by_p = 0;
  if (ctw) {
    const long long s = sz * (k0 - 1) + sy * j + i;
    ro_p = ro[s]; ro_pi = ro[s - 1]; ro_pj = ro[s - sy];
    vx_p = vel(mx, fvx, s); vx_pi = vel(mx, fvx, s - 1);
    vy_p = vel(my, fvy, s); vy_pj = vel(my, fvy, s - sy);
    vz_p = vel(mz, fvz, s); vz_pi = vel(mz, fvz, s - 1); vz_pj = vel(mz, fvz, s - sy);
    bx_p = bx[s]; by_p = by[s];
  }
  for (int k = k0; k < k1; k++) {
    const long long ss = sz * k + sy * j + i;
    {
      const long long a = ss + 2 * sz;
      u2[5] = load_as<FT>(p.emf[2], a);
      c2[5] = load_as<FT>(p.emf[3], a);
      u4[5] = load_as<FT>(p.emf[4], a);
      c4[5] = load_as<FT>(p.emf[5], a);
    }
    double ctz = 0.5, ctx = 0.5, cty = 0.5;
    double ro0 = 0, roi = 0, roj = 0, vx0 = 0, vxi = 0, vy0 = 0, vyj = 0, vz0 = 0, vzi = 0, vzj = 0, bx0 = 0,
           by0 = 0;
    if (ctw) {
      ro0 = ro[ss]; roi = ro[ss - 1]; roj = ro[ss - sy];
      const double roij = ro[ss - 1 - sy];
      vx0 = vel(mx, fvx, ss); vxi = vel(mx, fvx, ss - 1);
      const double vxj = vel(mx, fvx, ss - sy), vxij = vel(mx, fvx, ss - 1 - sy);
      vy0 = vel(my, fvy, ss); vyj = vel(my, fvy, ss - sy);
      const double vyi = vel(my, fvy, ss - 1), vyij = vel(my, fvy, ss - 1 - sy);
      vz0 = vel(mz, fvz, ss); vzi = vel(mz, fvz, ss - 1); vzj = vel(mz, fvz, ss - sy);
      bx0 = bx[ss]; by0 = by[ss];
      // x-y plane for Ez: strides (1, ny)
      ctz = cuct_weight(roij, roj, roi, ro0, vxij, vxj, vxi, vx0, vyij, vyj, vyi, vy0, bx[ss - sy], bx0, by[ss - 1],
                        by0);
      // y-z plane for Ex: strides (sy, sz)
      ctx = cuct_weight(ro_pj, ro_p, roj, ro0, vy_pj, vy_p, vyj, vy0, vz_pj, vz_p, vzj, vz0, by_p, by0, bz[ss - sy],
                        bz[ss]);
      // z-x plane for Ey: strides (sz, 1)
      cty = cuct_weight(ro_pi, roi, ro_p, ro0, vz_pi, vzi, vz_p, vz0, vx_pi, vxi, vx_p, vx0, bz[ss - 1], bz[ss], bx_p,
                        bx0);
    }
    // Ez: x sweep B_y (t) along y, then y sweep B_x (u) along x
    double ez = 0.0;
    ez += emf_part<ODR, FT>(p.emf[0], p.emf[1], ss, sy, 1.0 - ctz, -0.5);
    ez += emf_part<ODR, FT>(p.emf[6], p.emf[7], ss, 1, ctz, +0.5);
    // Ey: x sweep B_z (u) along z, then z sweep B_x (t) along x
    double ey = 0.0;
    ey += emf_part_w<ODR>(u2, c2, cty, +0.5);
    ey += emf_part<ODR, FT>(p.emf[8], p.emf[9], ss, 1, 1.0 - cty, -0.5);
    // Ex: y sweep B_z (t) along z, then z sweep B_y (u) along y
    double ex = 0.0;
    ex += emf_part_w<ODR>(u4, c4, 1.0 - ctx, -0.5);
    ex += emf_part<ODR, FT>(p.emf[10], p.emf[11], ss, sy, ctx, +0.5);
    store_as<FT>(p.ex, ss, ex);
    store_as<FT>(p.ey, ss, ey);
    store_as<FT>(p.ez, ss, ez);
    // next plane: shift the windows, carry this plane's values
#pragma unroll
    for (int m = 0; m < 5; m++) {
      u2[m] = u2[m + 1];
      c2[m] = c2[m + 1];
      u4[m] = u4[m + 1];
      c4[m] = c4[m + 1];
    }
    ro_p = ro0; ro_pi = roi; ro_pj = roj;
    vx_p = vx0; vx_pi = vxi;
    vy_p = vy0; vy_pj = vyj;
    vz_p = vz0; vz_pi = vzi; vz_pj = vzj;
    bx_p = bx0; by_p = by0;
  }
}

// ------------------------------------------------------------------ CT
// Staggered-B update (mhd_updt3d_ctb, common/mhd_func.c:302-325), marching
// along z: one thread per (i, j) column of a 64 x 4 tile over CT_KC planes.
// The z-direction E values (Ey for bx, Ex for by) sit in 4-plane register
// windows, so each of those planes is read once per chunk instead of four
// times; the in-plane neighbours come through the caches.  Per face the
// operations and their order are those of the reference.
constexpr int CT_KC = 32;

// FT: element type of the stored edge EMFs (the sweep precision)
template <int ODR, typename FT>
__global__ __launch_bounds__(256) void k_ct3d(StagePtrs p, Geo g)
{
  const int ni = g.nx - 7, nj = g.ny - 7, nk = g.nz - 7;  // cells [4, n-4]
  const int ntx = (ni + EMF_TX - 1) / EMF_TX, nty = (nj + EMF_TY - 1) / EMF_TY;
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int tile = L % (ntx * nty), chunk = L / (ntx * nty);
  const int i = 4 + (tile % ntx) * EMF_TX + (int)(threadIdx.x % EMF_TX);
  const int j = 4 + (tile / ntx) * EMF_TY + (int)(threadIdx.x / EMF_TX);
  const int k0 = 4 + chunk * CT_KC;
  if (i > g.nx - 4 || j > g.ny - 4 || k0 > g.nz - 4) return;  // no barriers below
  const int k1 = (k0 + CT_KC < 4 + nk) ? k0 + CT_KC : 4 + nk;
  const long long sy = g.sy, sz = g.sz;
  const long long c0 = sy * j + i;
  const double dtv = *p.dt;
  const double dtdx = dtv / g.dx, dtdy = dtv / g.dy, dtdz = dtv / g.dz;
  const FT *ex = reinterpret_cast<const FT*>(p.ex), *ey = reinterpret_cast<const FT*>(p.ey),
           *ez = reinterpret_cast<const FT*>(p.ez);
  const bool ji = j <= g.ny - 5, ii = i <= g.nx - 5;
  // windows: planes k-1 .. k+2 of ey and ex at (i, j); planes [3, nz-2]
  double eyw[4], exw[4];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    eyw[m] = ey[c0 + sz * (k0 - 1 + m)];
    exw[m] = ex[c0 + sz * (k0 - 1 + m)];
  }
  for (int k = k0; k < k1; k++) {
    const long long ss = sz * k + c0;
    eyw[3] = ey[ss + 2 * sz];
    exw[3] = ex[ss + 2 * sz];
    const bool ki = k <= g.nz - 5;
    if (ji && ki) {  // bx: one extra face in x
      const double du0 = +dfc<ODR>((double)ez[ss - sy], (double)ez[ss], (double)ez[ss + sy], (double)ez[ss + 2 * sy]);
      const double du1 = -dfc<ODR>(eyw[0], eyw[1], eyw[2], eyw[3]);
      p.out[4][ss] = rk_upd(p.in[4][ss], p.u0[4][ss], -dtdy * du0 - dtdz * du1, p.a, p.b);
    }
    if (ii && ki) {  // by: one extra face in y
      const double du0 = +dfc<ODR>(exw[0], exw[1], exw[2], exw[3]);
      const double du1 = -dfc<ODR>((double)ez[ss - 1], (double)ez[ss], (double)ez[ss + 1], (double)ez[ss + 2]);
      p.out[5][ss] = rk_upd(p.in[5][ss], p.u0[5][ss], -dtdz * du0 - dtdx * du1, p.a, p.b);
    }
    if (ii && ji) {  // bz: one extra face in z
      const double du0 = +dfc<ODR>((double)ey[ss - 1], eyw[1], (double)ey[ss + 1], (double)ey[ss + 2]);
      const double du1 = -dfc<ODR>((double)ex[ss - sy], exw[1], (double)ex[ss + sy], (double)ex[ss + 2 * sy]);
      p.out[6][ss] = rk_upd(p.in[6][ss], p.u0[6][ss], -dtdx * du0 - dtdy * du1, p.a, p.b);
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
      eyw[m] = eyw[m + 1];
      exw[m] = exw[m + 1];
    }
  }
}

// ------------------------------------------------------------------ launch
typedef void (*sweep3_fn)(StagePtrs, Geo, BcFlags, double, int);

template <int DIR>
static sweep3_fn pick3(int rmn, int odr)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return DIR == 0 ? k_sweep3d<0, R, O, double> : k_sweep3d_t<DIR == 0 ? 1 : DIR, R, O, double>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

static inline unsigned nblk3(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_stage3d(qas_ctx* c, const StagePtrs& p)
{
  const Geo& g = c->geo;
  const int odr = c->s.odr;
  const double gam = c->s.gam;
  hipStream_t st = c->stream;
  const bool f32 = c->s.precision == 1;
  const int dims[3] = {g.nx, g.ny, g.nz};
  sweep3_fn f[3] = {pick3<0>(c->s.rmn, odr), pick3<1>(c->s.rmn, odr), pick3<2>(c->s.rmn, odr)};
  if (c->s.precision == 1)
    for (int d = 0; d < 3; d++) f[d] = pick_sweep_f32(3, d, c->s.rmn, odr);
  if (c->s.precision == 2)
    for (int d = 0; d < 3; d++) f[d] = pick_sweep_fma(3, d, c->s.rmn, odr);
  // z, then y, then x: the transverse sweeps leave dt/dz Dz and dt/dy Dy in
  // the accumulators, the x sweep applies the stage sum
  for (int d = 2; d >= 0; d--) {
    if (!f[d]) return hipErrorInvalidValue;
    const int nseg = (dims[d] - 5 + SEG - 1) / SEG;
    if (d == 0) {
      const long long nline = g.nd / dims[d];
      QAS_LAUNCH(c, QAS_K_SWEEP_X, f[d], nblk3(nline * nseg * WAVE, 256), 256, 0, st, p, g, c->bc, gam, nseg);
    } else {
      // one 256-thread workgroup per (4-column group, segment, other transverse index)
      const long long nwg = (long long)((g.nx + YC - 1) / YC) * nseg * (d == 1 ? g.nz : g.ny);
      QAS_LAUNCH(c, QAS_K_SWEEP_X + d, f[d], (unsigned)nwg, 256, 0, st, p, g, c->bc, gam, nseg);
    }
  }
  const long long ne = (long long)(g.nx - 5) * (g.ny - 5) * (g.nz - 5);
  const long long nc = (long long)(g.nx - 7) * (g.ny - 7) * (g.nz - 7);
  // (the face parts are stored in the sweep precision)
  const long long nemf = (long long)((g.nx - 5 + EMF_TX - 1) / EMF_TX) * ((g.ny - 5 + EMF_TY - 1) / EMF_TY) *
                         ((g.nz - 5 + EMF_KC - 1) / EMF_KC);
  (void)ne;
  (void)nc;
  const long long nct = (long long)((g.nx - 7 + EMF_TX - 1) / EMF_TX) * ((g.ny - 7 + EMF_TY - 1) / EMF_TY) *
                        ((g.nz - 7 + CT_KC - 1) / CT_KC);
#define QAS_EMF_CT(O)                                                                                      \
  if (f32) QAS_LAUNCH(c, QAS_K_EMF, (k_emf3d_z<O, float>), (unsigned)nemf, 256, 0, st, p, g, c->s.ctw);   \
  else QAS_LAUNCH(c, QAS_K_EMF, (k_emf3d_z<O, double>), (unsigned)nemf, 256, 0, st, p, g, c->s.ctw);       \
  if (f32) QAS_LAUNCH(c, QAS_K_CT, (k_ct3d<O, float>), (unsigned)nct, 256, 0, st, p, g);                    \
  else QAS_LAUNCH(c, QAS_K_CT, (k_ct3d<O, double>), (unsigned)nct, 256, 0, st, p, g);
  switch (odr) {
    case 1: QAS_EMF_CT(1) break;
    case 2: QAS_EMF_CT(2) break;
    case 3: QAS_EMF_CT(3) break;
    default: QAS_EMF_CT(4) break;
  }
#undef QAS_EMF_CT
  return hipGetLastError();
}

}  // namespace qas
