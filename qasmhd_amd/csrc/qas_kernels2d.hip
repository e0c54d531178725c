// qas_kernels2d.hip -- one SSP-RK stage of the 2D ideal-MHD step on gfx950.
//
// Replaces the stage body of MHD2D::ideal (2D/SERIAL/OTvortex/
// mhd2d_solve.cpp:44-386).  A stage is four launches:
//   sweep<Y> cell-centred B and primitives of its LDS tile from the state
//            (stored for the x sweep)                        (:51-89)
//            then characteristic WCNS + Riemann flux per y face, for every
//            column; writes dt/dy times the y flux divergence of the
//            interior and the upwind/central parts of the Bx flux
//                                                           (:234-306)
//   sweep<X> same along x, for every row; applies the RK stage sum
//            a u0 + b (u + ((-dt/dx Dx) - dt/dy Dy)) to ro, mx, my, mz, bz,
//            en (the reference's expression)               (:127-199, 341-358)
//   emf2d    CUCT corner EMF from both sweeps' face parts   (:91-103, 201-232, 308-339)
//   ct2d     staggered bx, by update                        (:359-380)
//            and the bound() of the stage output             (:384-385)
// A sweep assigns one 64-lane wavefront to 64 consecutive cells of a line;
// lane l reconstructs cell f0-2+l, neighbour states and fluxes move between
// lanes by cross-lane shuffles, and the wavefront owns the 60 faces/cells
// [f0, f0+60).  Everything runs in FP64 with the reference's operation
// order (the library is built with -ffp-contract=off).
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_stage2d.hpp"
#include "qas_sweep2d.hpp"

namespace qas {

// ------------------------------------------------------------------ kernels
// bodies in qas_stage2d.hpp (shared with the batched kernels).  Blocks are
// renumbered XCD-aware (xcd_remap): the stencils of neighbouring rows then
// meet in the same L2.
// corner EMF: CJ corners of a column per thread (CJ = 1: one corner).
// Measured at 1024^2: 20.7 us with one corner per thread, 18.3 with 2, 18.8
// with 4 (156 VGPRs); at 256^2 the 2-corner form made the step 0.8 % slower
// (half the wavefronts for a domain that fills the GPU once), so it is
// used from EMF_COL_MIN corners on.
constexpr long long EMF_COL_MIN = 512LL * 512;
template <int ODR, typename FT, int CJ>
__global__ __launch_bounds__(256) void k_emf2d(StagePtrs p, Geo g, int ctw)
{
  const long long t = (long long)xcd_remap((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
  if constexpr (CJ == 1) emf2d_at<ODR, FT>(p, g, ctw, t);
  else emf2d_col<ODR, FT, CJ>(p, g, ctw, t);
}

// CT and the stage output's ghost fill in one launch: blocks [0, nbct) do
// the CT faces (renumbered XCD-aware among themselves), the rest the ghosts
template <int ODR>
__global__ __launch_bounds__(256) void k_ctb2d(StagePtrs p, Geo g, BcFlags bc, const double* phi, double gfac,
                                               long long nghost, int nbct)
{
  const int b = (int)blockIdx.x;
  if (b < nbct)
    ctb2d_at<ODR>(p, g, bc, phi, gfac, nghost, (long long)xcd_remap(b, nbct) * blockDim.x + threadIdx.x, true);
  else
    ctb2d_at<ODR>(p, g, bc, phi, gfac, nghost, (long long)(b - nbct) * blockDim.x + threadIdx.x, false);
}

// ------------------------------------------------------------------ launch
static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_stage2d(qas_ctx* c, const StagePtrs& p, bool faces_only)
{
  const Geo& g = c->geo;
  const int odr = c->s.odr;
  const double gam = c->s.gam;
  hipStream_t st = c->stream;
  const bool f32 = c->s.precision == 1;
  // 1) y sweep (with the primitives), 2) x sweep
  const int nsx = (g.nx - 5 + SEG - 1) / SEG;
  const int nsy = (g.ny - 5 + SEG - 1) / SEG;
  const int prec = c->s.precision;
  xsweep2d_fn fx = prec == 1 ? pick_xsweep2d_f32(c->s.rmn, odr, false)
                             : (prec == 2 ? pick_xsweep2d_fma(c->s.rmn, odr, false) : pick_xsweep2d(c->s.rmn, odr, false));
  xsweep2d_fn fy = prec == 1 ? pick_ysweep2d_f32(c->s.rmn, odr, false)
                             : (prec == 2 ? pick_ysweep2d_fma(c->s.rmn, odr, false) : pick_ysweep2d(c->s.rmn, odr, false));
  // with gravity the x sweep takes the round-tripped stage energy, which
  // the y sweep computes from the state's own energy and stores
  StagePtrs ps = p;
  if (p.phi) {
    fx = pick_grav_xsweep(c->s.rmn, odr);
    fy = pick_grav_ysweep(c->s.rmn, odr);
    ps.in[7] = p.enrt;
  }
  if (!fx || !fy) return hipErrorInvalidValue;
  // y sweep first (one 256-thread workgroup per (4-column group, segment)):
  // it stores the primitives and leaves dt/dy * Dy in the accumulators; the
  // x sweep applies the stage sum
  QAS_LAUNCH(c, QAS_K_SWEEP_Y, fy, (unsigned)(((g.nx + YC - 1) / YC) * nsy), 256, 0, st, p, g, c->bc, gam, nsy,
             (const BatchInst*)nullptr, (const int*)nullptr, 0, 0);
  // x sweep: one wavefront per (line, segment), c->xwpb per workgroup
  const unsigned xb = (unsigned)(WAVE * c->xwpb);
  QAS_LAUNCH(c, QAS_K_SWEEP_X, fx, nblk((long long)nsx * g.ny * WAVE, (int)xb), xb, 0, st, ps, g, c->bc, gam, nsx,
             (const BatchInst*)nullptr, (const int*)nullptr, 0, 0);
  if (faces_only) return hipGetLastError();
  // 3) corner EMF, 4) CT with the ghost fill of the stage output
  const bool col = (long long)(g.nx - 5) * (g.ny - 5) >= EMF_COL_MIN;
  const long long ne = (long long)(g.nx - 5) * (col ? (g.ny - 5 + 1) / 2 : g.ny - 5);
  const long long nc = (long long)(g.nx - 7) * (g.ny - 7);
  const long long nghost = (long long)g.nx * g.ny - (long long)(g.nx - 2 * NG) * (g.ny - 2 * NG);
  const double gfac = c->d_phi ? (2.0 - gam) / (gam - 1.0) : 0.0;
  // (the face parts are stored in the sweep precision)
  const int eb = c->ewb;
  const unsigned nbct = nblk(nc, eb);

#define QAS_EMF_CT(O)                                                                                      \
  if (f32 && col) QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<O, float, 2>), nblk(ne, eb), eb, 0, st, p, g, c->s.ctw); \
  else if (f32) QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<O, float, 1>), nblk(ne, eb), eb, 0, st, p, g, c->s.ctw);  \
  else if (col) QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<O, double, 2>), nblk(ne, eb), eb, 0, st, p, g, c->s.ctw); \
  else QAS_LAUNCH(c, QAS_K_EMF, (k_emf2d<O, double, 1>), nblk(ne, eb), eb, 0, st, p, g, c->s.ctw);         \
  QAS_LAUNCH(c, QAS_K_CT, (k_ctb2d<O>), nbct + nblk(nghost * 8, eb), eb, 0, st, p, g, c->bc, c->d_phi, gfac,   \
             nghost, (int)nbct);
  switch (odr) {
    case 1: QAS_EMF_CT(1) break;
    case 2: QAS_EMF_CT(2) break;
    case 3: QAS_EMF_CT(3) break;
    default: QAS_EMF_CT(4) break;
  }
#undef QAS_EMF_CT
  return hipGetLastError();
}

xsweep2d_fn pick_ysweep2d(int rmn, int odr, bool batch) { return pick_ysweep2d_t<double>(rmn, odr, batch); }
xsweep2d_fn pick_xsweep2d(int rmn, int odr, bool batch) { return pick_xsweep2d_t<double>(rmn, odr, batch); }

}  // namespace qas
