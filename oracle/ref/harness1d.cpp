/*
 * harness1d.cpp -- runs the reference's own MHD1D::ideal/bound/setdt
 * (1D/SERIAL/shock/mhd1d_{solve,class}.cpp, built from source) for an exact
 * number of steps on a state read from disk (TEST INFRASTRUCTURE).
 * Parameters as in harness2d.cpp (dx, gam, nsteps, flg, dt, dnx).
 */
#include "mhd1d_class.hpp"
#include "harness_common.hpp"
#include <chrono>

class RefRun1D : public MHD1D {
 public:
  int run(const HParams &p) {
    dx = p.d("dx");
    setgam(p.d("gam"));
    p.ints("dnx", dnxs, nm);
    if (!read_raw(p.s("in").c_str(), val, nm, nx)) return 2;
    const int nsteps = p.i("nsteps");
    const int flg = p.i("flg");
    std::vector<double> hist;
    if (flg) setdt(1); else dt = p.d("dt");
    const auto t0 = std::chrono::steady_clock::now();
    for (n = 1; n <= nsteps; n++) {
      hist.push_back(dt);
      bound(val, nm, dnxs);
      ideal(dt);
      setdt(flg * (n % 2 == 0));
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("elapsed_s %.9f\n", el);
    if (!write_raw(p.s("out").c_str(), val, nm, nx)) return 3;
    if (!p.s("dthist").empty()) write_vec(p.s("dthist").c_str(), hist);
    return 0;
  }
};

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s params\n", argv[0]); return 1; }
  HParams p;
  if (!p.load(argv[1])) return 1;
  RefRun1D r;
  return r.run(p);
}
