// qas_fma64.hip -- FP64 face sweeps built with FMA contraction
// (qas_scheme.precision = 2).
//
// The same sweep templates as the bit-exact path (qas_sweep2d.hpp,
// qas_sweep3d.hpp), instantiated under the build tag FMA = 1 in this
// translation unit, which is compiled with -ffp-contract=fast: a multiply
// feeding an add becomes one fused multiply-add (one rounding instead of
// two).  Every operation is still FP64; divisions and square roots stay
// IEEE.  The primitive recovery in the y sweep keeps its own contraction
// off.  The result differs from the reference's rounding sequence by a few
// units in the last place per operation, which the tests bound as a
// relative L1 distance from the bit-exact path after many steps; it is not
// bit-identical to the reference, so parity tests use precision 0.
#include "qas_sweep2d.hpp"
#include "qas_sweep3d.hpp"

namespace qas {

xsweep2d_fn pick_ysweep2d_fma(int rmn, int odr, bool batch) { return pick_ysweep2d_t<double, 1>(rmn, odr, batch); }
xsweep2d_fn pick_xsweep2d_fma(int rmn, int odr, bool batch) { return pick_xsweep2d_t<double, 1>(rmn, odr, batch); }

typedef void (*sweep_fn_t)(StagePtrs, Geo, BcFlags, double, int);

template <int DIR>
static sweep_fn_t pick3(int rmn, int odr)
{
#define QAS_S(R, O)                                                                                  \
  if (rmn == R && odr == O)                                                                          \
    return DIR == 0 ? k_sweep3d<0, R, O, double, 1> : k_sweep3d_t<DIR == 0 ? 1 : DIR, R, O, double, 1>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

sweep_fn_t pick_sweep_fma(int ndim, int dir, int rmn, int odr)
{
  if (ndim == 3) return dir == 0 ? pick3<0>(rmn, odr) : (dir == 1 ? pick3<1>(rmn, odr) : pick3<2>(rmn, odr));
  return nullptr;
}

}  // namespace qas
