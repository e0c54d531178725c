// qas_mixed.hip -- FP32 instantiation of the face sweeps (the mixed path).
//
// Reconstruction (characteristic projection + WCNS), the Riemann solver and
// the B-flux face states run in FP32; the conserved state, the primitive /
// pressure recovery (prim kernels), the flux differencing, the RK stage sum
// and the CT update stay FP64.  This translation unit is compiled with
// -fno-hip-fp32-correctly-rounded-divide-sqrt (FP32 divide and sqrt by the
// hardware reciprocal / square-root estimates with their refinement) and
// with FMA contraction; its error against the FP64 path is measured, not
// assumed.
#include "qas_sweep2d.hpp"
#include "qas_sweep3d.hpp"

namespace qas {

typedef void (*sweep_fn_t)(StagePtrs, Geo, BcFlags, double, int);

template <int DIR>
static sweep_fn_t pick3(int rmn, int odr)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return DIR == 0 ? k_sweep3d<0, R, O, float> : k_sweep3d_t<DIR == 0 ? 1 : DIR, R, O, float>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

sweep_fn_t pick_sweep_f32(int ndim, int dir, int rmn, int odr)
{
  if (ndim == 3) return dir == 0 ? pick3<0>(rmn, odr) : (dir == 1 ? pick3<1>(rmn, odr) : pick3<2>(rmn, odr));
  return nullptr;
}

xsweep2d_fn pick_ysweep2d_f32(int rmn, int odr, bool batch) { return pick_ysweep2d_t<float>(rmn, odr, batch); }
xsweep2d_fn pick_xsweep2d_f32(int rmn, int odr, bool batch) { return pick_xsweep2d_t<float>(rmn, odr, batch); }

}  // namespace qas
