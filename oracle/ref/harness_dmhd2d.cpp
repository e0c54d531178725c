/*
 * harness_dmhd2d.cpp -- runs the reference's own DMHD2D::dsptv
 * (2D/SERIAL/MRX/dmhd2d_solve.cpp, built from source) on a state and
 * per-cell viscosity / resistivity read from disk (TEST INFRASTRUCTURE).
 * Grid size and solver switches are the reference's compile-time macros,
 * supplied by oracle/build.py.
 *
 * usage: harness_dmhd2d <param-file>
 *   in/out        raw state: 8 SoA arrays of nx*ny doubles (ghosts included)
 *   nu/eta        raw nx*ny doubles each (the arrays DMHD2D::setdc fills)
 *   dx dy gam dt  hex floats
 *   nsub          number of {bound; dsptv(dt)} calls            (mode 0)
 *   mode          0: dsptv sub-steps only
 *                 1: nsteps of the exec_ body of DMHD2D (dmhd2d_class.cpp:15-33)
 *                    at fixed dt: bound; ideal(dt); ncmax sub-steps of
 *                    {bound; dsptv(dt/ncmax)} with ncmax from dcmax
 *   nsteps dcmax  mode 1
 *   dnx dny       8 boundary flags each
 */
#include "dmhd2d_class.hpp"
#include "harness_common.hpp"
#include <chrono>

class RefDsp2D : public DMHD2D {
 public:
  int run(const HParams &p) {
    dx = p.d("dx");
    dy = p.d("dy");
    dr = min(dx, dy);
    setgam(p.d("gam"));
    p.ints("dnx", dnxs, nm);
    p.ints("dny", dnys, nm);
    if (!read_raw(p.s("in").c_str(), val, nm, nd)) return 2;
    double *const pn[1] = {nu};
    double *const pe[1] = {eta};
    if (!read_raw(p.s("nu").c_str(), pn, 1, nd)) return 2;
    if (!read_raw(p.s("eta").c_str(), pe, 1, nd)) return 2;
    dt = p.d("dt");
    const auto t0 = std::chrono::steady_clock::now();
    if (p.i("mode") == 0) {
      const int nsub = p.i("nsub");
      for (int k = 0; k < nsub; k++) {
        bound(val, nm, stxs, dnxs, stys, dnys);
        dsptv(dt);
      }
    } else {
      const int nsteps = p.i("nsteps");
      const double dcmax = p.d("dcmax");
      for (n = 1; n <= nsteps; n++) {
        bound(val, nm, stxs, dnxs, stys, dnys);
        ideal(dt);
        if (dcmax > 1e-15) {
          int nc, ncmax = 1 + (int)(2 * (dcmax * dt) / (dr * dr));
          for (nc = 0; nc < ncmax; nc++) {
            bound(val, nm, stxs, dnxs, stys, dnys);
            dsptv(dt / ncmax);
          }
        }
      }
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("elapsed_s %.9f\n", el);
    if (!write_raw(p.s("out").c_str(), val, nm, nd)) return 3;
    return 0;
  }
};

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s params\n", argv[0]); return 1; }
  HParams p;
  if (!p.load(argv[1])) return 1;
  RefDsp2D r;
  return r.run(p);
}
