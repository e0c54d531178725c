// ot2d_driver.cpp -- a 2D problem written against qasmhd_driver.hpp the way
// a reference problem directory is written against MHD2D: paras() sets the
// domain and boundary flags (cf. 2D/SERIAL/OTvortex/mhd2d_paras.cpp), init_()
// the Orszag-Tang state (cf. mhd2d_init_.cpp), main() calls setdt(1) and
// exec_(1).  It dumps the initial and final FP64 states so a test can rerun
// the oracle on the same initial data.
//
// usage: ot2d_driver <xmesh> <ymesh> <nsteps> <rmn> <init.bin> <final.bin> [outdir [restart|mpi]]
// With outdir, the initial output (dout_) and a backup at the end go there;
// with "restart", the run instead resumes from that backup (bkup_load) and
// advances nsteps more steps.  With "mpi" it runs as the MPI driver does
// (mpi_files): an output and a backup after every step, and a run that
// finds a backup in outdir resumes from it up to step nsteps.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "qasmhd_driver.hpp"

class OTVortex : public qasmhd::MHD2DDevice {
 public:
  // every_step: dtrec = 1e-9 (an output after every step) and n_out large
  // enough that tmax (= dtrec * n_out) never ends the run
  OTVortex(int xm, int ym, int nsteps, bool every_step = false)
      : MHD2DDevice(xm, ym, every_step ? 1000000000 : 1, every_step ? 1e-9 : 1e6, 0.4), steps_(nsteps) {}
  void paras()
  {
    xmin = 0.0;
    xmax = 2.0 * M_PI;
    ymin = 0.0;
    ymax = 2.0 * M_PI;
    dx = (xmax - xmin) / (nx - 2 * xoff);
    dy = (ymax - ymin) / (ny - 2 * yoff);
    dr = dx < dy ? dx : dy;
  }
  void init_()
  {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int ss = nx * j + i;
        const double x = (i - xoff) * dx + xmin, y = (j - yoff) * dy + ymin;
        const double r = gam * gam, vx = -std::sin(y), vy = std::sin(x), p = gam;
        const double cx = -std::sin(y), cy = std::sin(2 * x);
        ro[ss] = r;
        mx[ss] = r * vx;
        my[ss] = r * vy;
        mz[ss] = 0.0;
        bx[ss] = cx;
        by[ss] = cy;
        bz[ss] = 0.0;
        en[ss] = p / (gam - 1) + 0.5 * (r * (vx * vx + vy * vy) + (cx * cx + cy * cy));
      }
  }
  void run_steps(int flg)
  {
    nmax = steps_;
    exec_(flg);
    download();
  }
  void save(const char* path)
  {
    FILE* f = std::fopen(path, "wb");
    for (int m = 0; m < nm; m++) std::fwrite(val[m], sizeof(double), nd, f);
    std::fclose(f);
  }
  void set_scheme(int r) { rmn = r; }
  // exec_ leaves n at nmax + 1 when it stops on the step count (its loop
  // test increments n once more); the reference saves its backup inside the
  // loop, after step n.  Save the state after step nmax the same way.
  void backup_after_run()
  {
    n = nmax;
    bkup_save();
  }
  // resume from the backup in fildir and advance `steps_` more steps
  void restart_steps(int flg)
  {
    if (!bkup_load()) throw qasmhd::Error("no backup in " + fildir);
    nflg_ = 1;
    nmax = n + steps_;
    exec_(flg);
    download();
  }

 private:
  int steps_;
};

int main(int argc, char** argv)
{
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s xmesh ymesh nsteps rmn init.bin final.bin\n", argv[0]);
    return 1;
  }
  try {
    const std::string mode = argc > 8 ? argv[8] : "";
    OTVortex ot(std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]), mode == "mpi");
    ot.set_scheme(std::atoi(argv[4]));
    ot.paras();
    if (argc > 7) ot.fildir = argv[7];
    if (mode == "mpi") {
      ot.mpi_files = true;
      ot.init_();
      ot.save(argv[5]);
      ot.setdt(1);
      ot.run_steps(1);
      ot.save(argv[6]);
    } else if (mode == "restart") {
      ot.restart_steps(1);
      ot.save(argv[6]);
    } else {
      ot.init_();
      ot.save(argv[5]);
      ot.setdt(1);
      ot.run_steps(1);
      ot.save(argv[6]);
      if (argc > 7) ot.backup_after_run();
    }
    std::printf("dt %.17g\n", ot.getdt());
  } catch (const qasmhd::Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
