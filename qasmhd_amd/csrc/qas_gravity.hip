// qas_gravity.hip -- the 2D step with a gravitational potential: the
// reference's RTI driver (2D/SERIAL/RTI/mhd2d_solve.cpp, the OTvortex
// ideal() plus five lines; mhd2d_class.cpp:10-48).
//
// The gravity terms live in the kernels they touch, switched on by the
// context's potential:
//   prim     the potential energy ro phi is taken out of en for the
//            pressure and put back; the round-tripped en is the stage input
//            energy the x sweep's stage sum reads (the reference writes it
//            back in place), kept in a scratch array so the step base is
//            the untouched state (qas_stage2d.hpp)
//   sweeps   work done by gravity in the energy flux, F_en += F_ro (phi_L +
//            phi_R)/2; the x sweep adds -b ro (phi_{i+1} - phi_{i-1})/2 dt/dx
//            to mx and the same along y to my with the updated density
//            (GRAV instantiations of qas_sweep2d.hpp, made here)
//   bound    hydrostatic energy ghosts along y (qas_common.hpp)
//   CFL      the potential round trip of setdt, in place (qas_common.hpp)
#include <string>

#include "qas_sweep2d.hpp"

namespace qas {

xsweep2d_fn pick_grav_ysweep(int rmn, int odr)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return k_sweep2d_y<R, O, double, false, true>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

xsweep2d_fn pick_grav_xsweep(int rmn, int odr)
{
#define QAS_S(R, O) \
  if (rmn == R && odr == O) return k_sweep2d<0, R, O, double, false, true>;
  QAS_S(0, 1) QAS_S(0, 2) QAS_S(0, 3) QAS_S(0, 4)
  QAS_S(1, 1) QAS_S(1, 2) QAS_S(1, 3) QAS_S(1, 4)
  QAS_S(2, 1) QAS_S(2, 2) QAS_S(2, 3) QAS_S(2, 4)
  QAS_S(3, 1) QAS_S(3, 2) QAS_S(3, 3) QAS_S(3, 4)
#undef QAS_S
  return nullptr;
}

void grav_release(qas_ctx* c)
{
  if (c && c->d_phi) {
    (void)hipFree(c->d_phi);
    c->d_phi = nullptr;
    c->d_enrt = nullptr;
  }
}

}  // namespace qas

using namespace qas;

static int gfail(qas_ctx* c, int code, const std::string& m)
{
  if (c) c->err = m;
  return code;
}

extern "C" int qas_set_gravity(qas_ctx* c, const double* phi)
{
  if (!c || !c->slab) return gfail(c, QAS_E_ARG, "qas_set_gravity: null context");
  c->graph_flg = -1;  // the captured step graph holds the old kernels
  if (!phi) {
    grav_release(c);
    return QAS_OK;
  }
  if (c->g.ndim != 2) return gfail(c, QAS_E_ARG, "qas_set_gravity: gravity covers 2D domains");
  if (c->s.precision != 0) return gfail(c, QAS_E_ARG, "qas_set_gravity: gravity is FP64 only");
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  if (!c->d_phi) {
    double* d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * bytes);
    if (e != hipSuccess) return gfail(c, QAS_E_NOMEM, std::string("qas_set_gravity: ") + hipGetErrorName(e));
    c->d_phi = d;
    c->d_enrt = d + c->geo.nd;
  }
  hipError_t e = hipMemcpyAsync(c->d_phi, phi, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return gfail(c, QAS_E_HIP, std::string("qas_set_gravity: ") + hipGetErrorName(e));
  c->bounded = false;  // the energy ghosts now follow the hydrostatic rule
  return QAS_OK;
}
