/*
 * unit.cpp -- evaluates the reference's own per-function routines
 * (common/*.c, built from source) on batches of inputs read from a raw
 * double file, to write golden vectors (TEST INFRASTRUCTURE).  The
 * selection tables are taken from the reference's class MHD
 * (common/mhd_class.hpp:102-115) through a subclass.
 *
 * usage: unit <func> <sel> <n> <gam-hex> <in.bin> <out.bin>
 *   interp  sel=ODR-1 (interpol) or 10+ODR-1 (l_interp): in n x 5 -> out n x 2
 *   fcen/df sel=ODR:  in n x 4 -> out n
 *   lrstate sel=ODR:  in n x 36 (ro,vn,vt,vu,bt,bu,pr stencils, bn) -> out n x 14
 *   flux    sel=RMN:  in n x 17 (L[7], R[7], bn, dv[2]) -> out n x 7
 *   cuct    sel=0:    in n x 16 (ro[4], vx[4], vy[4], bx(i,j-1), bx(i,j),
 *                     by(i-1,j), by(i,j)) -> out n
 */
#include "mhd_class.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

struct Tabs : public MHD {
  void (*interp(int odr))(const double *, double *, double *) { return interpol[odr - 1]; }
  void (*lin(int odr))(const double *, double *, double *) { return l_interp[odr - 1]; }
  double (*fc(int odr))(const double *) { return fcen[odr - 1]; }
  double (*d1(int odr))(const double *) { return df1[odr - 1]; }
  void (*rs(int rmn))(double, double, double, double, double, double, double, double, double, double, double,
                      double, double, double, double, double, const double *, double *, double *, double *,
                      double *, double *, double *, double *) { return riemann[rmn]; }
};

int main(int argc, char **argv) {
  if (argc < 7) { std::fprintf(stderr, "usage\n"); return 1; }
  const char *fn = argv[1];
  const int sel = std::atoi(argv[2]);
  const long n = std::atol(argv[3]);
  const double gam = std::strtod(argv[4], nullptr);
  int w_in = 0, w_out = 0;
  if (!std::strcmp(fn, "interp")) { w_in = 5; w_out = 2; }
  else if (!std::strcmp(fn, "fcen") || !std::strcmp(fn, "df")) { w_in = 4; w_out = 1; }
  else if (!std::strcmp(fn, "lrstate")) { w_in = 36; w_out = 14; }
  else if (!std::strcmp(fn, "flux")) { w_in = 17; w_out = 7; }
  else if (!std::strcmp(fn, "cuct")) { w_in = 16; w_out = 1; }
  else return 1;
  std::vector<double> in(n * w_in), out(n * w_out);
  FILE *f = std::fopen(argv[5], "rb");
  if (!f || std::fread(in.data(), sizeof(double), in.size(), f) != in.size()) return 2;
  std::fclose(f);
  Tabs t;
  for (long q = 0; q < n; q++) {
    const double *a = &in[q * w_in];
    double *o = &out[q * w_out];
    if (!std::strcmp(fn, "interp")) {
      auto fp = (sel >= 10) ? t.lin(sel - 10 + 1) : t.interp(sel + 1);
      fp(&a[2], &o[0], &o[1]);
    } else if (!std::strcmp(fn, "fcen")) {
      o[0] = t.fc(sel)(&a[1]);
    } else if (!std::strcmp(fn, "df")) {
      o[0] = t.d1(sel)(&a[1]);
    } else if (!std::strcmp(fn, "lrstate")) {
      mhd_lrstate(&a[2], &a[7], &a[12], &a[17], &a[22], &a[27], &a[32], a[35], gam, 1, t.interp(sel), &o[0], &o[7]);
    } else if (!std::strcmp(fn, "flux")) {
      double F[7];
      t.rs(sel)(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], a[12], a[13], a[14], gam,
                &a[15], &F[0], &F[1], &F[2], &F[3], &F[4], &F[5], &F[6]);
      std::memcpy(o, F, sizeof(F));
    } else if (!std::strcmp(fn, "cuct")) {
      /* 2x2 cells at strides (1, 2); pointer at cell (i,j) = index 3 */
      double ro[4], vx[4], vy[4], bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
      for (int k = 0; k < 4; k++) { ro[k] = a[k]; vx[k] = a[4 + k]; vy[k] = a[8 + k]; }
      bx[1] = a[12]; bx[3] = a[13]; by[2] = a[14]; by[3] = a[15];
      o[0] = mhd_cuct_weight(&ro[3], &vx[3], &vy[3], &bx[3], &by[3], 1, 2);
    }
  }
  f = std::fopen(argv[6], "wb");
  if (!f) return 3;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  return 0;
}
