/*
 * harness2d.cpp -- runs the reference's own MHD2D::ideal/bound/setdt
 * (2D/SERIAL/OTvortex/mhd2d_{solve,class}.cpp, built from source) for an
 * exact number of steps on a state read from disk (TEST INFRASTRUCTURE).
 * Grid size and solver switches are the reference's compile-time macros
 * (XMESH, YMESH, RMN, ODR, R_K, CTW, CFL), supplied by oracle/build.py.
 *
 * usage: harness2d <param-file>
 *   in/out      raw state: 8 SoA arrays of nx*ny doubles (ghosts included)
 *   dthist      raw doubles: dt used by each step
 *   dx dy gam   hex floats
 *   nsteps      step count
 *   flg         1: dt from setdt(1) then setdt(n%2==0) after each step
 *               (exec_(1) schedule); 0: fixed dt
 *   dt          fixed dt (flg 0)
 *   dnx dny     8 boundary flags each (ro,mx,my,mz,bx,by,bz,en)
 */
#include "mhd2d_class.hpp"
#include "harness_common.hpp"
#include <chrono>

class RefRun2D : public MHD2D {
 public:
  int run(const HParams &p) {
    dx = p.d("dx");
    dy = p.d("dy");
    dr = min(dx, dy);
    setgam(p.d("gam"));
    p.ints("dnx", dnxs, nm);
    p.ints("dny", dnys, nm);
    if (!read_raw(p.s("in").c_str(), val, nm, nd)) return 2;
#ifdef QAS_HARNESS_GRAVITY
    {  // the RTI driver's gravitational potential (allocated by its MHD2D())
      double *const pg[1] = {phi_g};
      if (!read_raw(p.s("phi").c_str(), pg, 1, nd)) return 2;
    }
#endif
    const int nsteps = p.i("nsteps");
    const int flg = p.i("flg");
    std::vector<double> hist;
    if (flg) setdt(1); else dt = p.d("dt");
    const auto t0 = std::chrono::steady_clock::now();
    for (n = 1; n <= nsteps; n++) {
      hist.push_back(dt);
      bound(val, nm, stxs, dnxs, stys, dnys);
      ideal(dt);
      setdt(flg * (n % 2 == 0));
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("elapsed_s %.9f\n", el);
    if (!write_raw(p.s("out").c_str(), val, nm, nd)) return 3;
    if (!p.s("dthist").empty()) write_vec(p.s("dthist").c_str(), hist);
    return 0;
  }
};

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s params\n", argv[0]); return 1; }
  HParams p;
  if (!p.load(argv[1])) return 1;
  RefRun2D r;
  return r.run(p);
}
