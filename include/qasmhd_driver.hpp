// qasmhd_driver.hpp -- C++ host driver over the C ABI of qasmhd.h.
//
// Mirrors the reference's per-dimension driver classes (MHD1D
// 1D/SERIAL/shock/mhd1d_class.hpp:7-38, MHD2D 2D/SERIAL/OTvortex/
// mhd2d_class.hpp:7-47, MHD3D 3D/MPI/OTvortex/mhd3d_class.hpp:8-70): the same
// public members (nx, ny, nz with ghosts, xoff.., gam, cfl), host SoA arrays
// with the reference's names (ro, mx, my, mz, bx, by, bz, en, val[8]),
// boundary/staggering flags (dnxs.., stxs..), and the same methods --
// setdt(int), exec_(int), and the protected bound() / ideal(double) -- whose
// bodies run on the device.  A problem fills the host arrays in its init_()
// exactly as it does for the reference and calls setdt(1); exec_(1).  The
// state lives on the device between steps; the host arrays are refreshed
// only for output (dout_) and on request (download()).
//
// Output and restart files follow the reference: dout_ writes params.dat,
// offsets.dat, t.dat, x/y/z.dat and outdat_%05d.dat (FP64, 8 arrays of nd
// doubles, ghosts included) into fildir; bkup_save/bkup_load the
// bkup_stamp/time/data files of common_func.c:281-344.  With mpi_files set
// (the default of the 3D driver, whose reference is MPI only) they follow
// the MPI drivers run as one rank instead (2D/MPI/OTvortex/mhd2d_class.cpp,
// 3D/MPI/OTvortex/mhd3d_class.cpp): exec_ resumes from the backup in fildir
// when there is one and saves a backup at every output, and dout_ adds
// mpinum.dat, names the coordinate files x_00000.dat.. and writes
// outdat_%05d_00000.dat in FP32.
//
// Header-only; link with libqasmhd_amd.so.  Errors throw qasmhd::Error
// (the C ABI itself never throws).
#pragma once
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "qasmhd.h"

namespace qasmhd {

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// One device-resident domain (RAII over qas_ctx).
class Domain {
 public:
  Domain(const qas_grid& g, const qas_scheme& s)
  {
    int r = qas_create(&c_, &g, &s);
    if (r != QAS_OK) {
      std::string m = c_ ? qas_last_error(c_) : "qas_create failed";
      qas_destroy(c_);
      c_ = nullptr;
      throw Error("qas_create: " + m);
    }
  }
  ~Domain() { qas_destroy(c_); }
  Domain(const Domain&) = delete;
  Domain& operator=(const Domain&) = delete;
  void upload(double* const val[8]) { check(qas_upload(c_, val), "qas_upload"); }
  void download(double* const val[8]) { check(qas_download(c_, val), "qas_download"); }
  void bound() { check(qas_bound(c_), "qas_bound"); }
  void step(double dt) { check(qas_step(c_, dt), "qas_step"); }
  double cfl_dt()
  {
    double d = 0.0;
    check(qas_cfl_dt(c_, &d), "qas_cfl_dt");
    return d;
  }
  // nsteps of exec_'s loop body on the device; returns the current dt
  double run(int nsteps, int flg, double dt, std::vector<double>* hist = nullptr)
  {
    double d = dt;
    if (hist) hist->resize(nsteps);
    check(qas_run(c_, nsteps, flg, &d, hist ? hist->data() : nullptr), "qas_run");
    return d;
  }
  qas_ctx* get() { return c_; }

 private:
  void check(int r, const char* what)
  {
    if (r != QAS_OK) throw Error(std::string(what) + ": " + qas_last_error(c_));
  }
  qas_ctx* c_ = nullptr;
};

// Common state of the drivers: field names and flags of common/mhd_class.hpp:32-42.
class MHDDevice {
 public:
  static const int nm = 8;
  void setgam(double g) { gam = g; }
  double getgam() const { return gam; }
  double getdt() const { return dt; }
  // solver switches (RMN, ODR, R_K, CTW of mymacros.hpp) and precision
  int rmn = 1, odr = 4, r_k = 3, ctw = 1, precision = 0;

 protected:
  double gam = 5.0 / 3.0;
  double *ro = nullptr, *mx = nullptr, *my = nullptr, *mz = nullptr, *en = nullptr;
  double *bx = nullptr, *by = nullptr, *bz = nullptr;
  double* val[nm] = {};
  int dnxs[nm] = {0}, dnys[nm] = {0}, dnzs[nm] = {0};
  int stxs[nm] = {0}, stys[nm] = {0}, stzs[nm] = {0};
  double dt = 0.0, tim = 0.0, trec = 0.0;
  int n = 0, cnt = 0, nmax = 1;
  std::vector<double> store_;
  Domain* dom_ = nullptr;

  void alloc(long long nd)
  {
    store_.assign(8 * (size_t)nd, 0.0);
    for (int m = 0; m < nm; m++) val[m] = store_.data() + (size_t)m * nd;
    ro = val[0]; mx = val[1]; my = val[2]; mz = val[3];
    bx = val[4]; by = val[5]; bz = val[6]; en = val[7];
    stxs[4] = 1;
    stys[5] = 1;
    stzs[6] = 1;
  }
  qas_scheme scheme(double cfl) const
  {
    qas_scheme s;
    qas_scheme_default(&s);
    s.gam = gam;
    s.cfl = cfl;
    s.rmn = rmn;
    s.odr = odr;
    s.r_k = r_k;
    s.ctw = ctw;
    s.precision = precision;
    for (int m = 0; m < nm; m++) {
      s.dn[0][m] = dnxs[m]; s.dn[1][m] = dnys[m]; s.dn[2][m] = dnzs[m];
      s.st[0][m] = stxs[m]; s.st[1][m] = stys[m]; s.st[2][m] = stzs[m];
    }
    return s;
  }
  virtual qas_grid grid() const = 0;
  virtual double cfl_value() const = 0;
  // the device context is created on first use, after paras()/init_() set
  // spacing, flags and the initial state
  Domain& dom()
  {
    if (!dom_) {
      dom_ = new Domain(grid(), scheme(cfl_value()));
      dom_->upload(val);
    }
    return *dom_;
  }

 public:
  virtual ~MHDDevice() { delete dom_; }
  // copy the device state into the host arrays (ro..en)
  void download() { if (dom_) dom_->download(val); }
  // re-send the host arrays (after the caller changed them)
  void upload() { dom().upload(val); }

  // Output directory of dout_ and the backup files (empty: no files).
  std::string fildir;
  // the file set and backup cycle of the MPI drivers, as rank 0 of 1
  bool mpi_files = false;

  // Backup files of common_func.c:318-344 (bkup_save, rank 0):
  // bkup_stamp.dat "n cnt", bkup_time.dat {tim, dt, trec} FP64,
  // bkup_data_00000.dat the 8 arrays (nd doubles each, ghosts included).
  void bkup_save()
  {
    download();
    write_backup();
  }
  // bkup_load (common_func.c:281-316): restores n, cnt, tim, dt, trec and
  // the state if bkup_stamp.dat exists; returns whether it did.
  bool bkup_load()
  {
    FILE* f = std::fopen(path("bkup_stamp.dat").c_str(), "r");
    if (!f) return false;
    int vi[2] = {0, 0};
    const int got = std::fscanf(f, "%d %d", &vi[0], &vi[1]);
    std::fclose(f);
    if (got != 2) throw Error("bkup_load: malformed bkup_stamp.dat");
    double vd[3] = {0, 0, 0};
    f = open_or_throw(path("bkup_time.dat"), "rb");
    const size_t nt = std::fread(vd, sizeof(double), 3, f);
    std::fclose(f);
    if (nt != 3) throw Error("bkup_load: short bkup_time.dat");
    f = open_or_throw(path("bkup_data_00000.dat"), "rb");
    for (int m = 0; m < nm; m++)
      if (std::fread(val[m], sizeof(double), store_.size() / nm, f) != store_.size() / nm) {
        std::fclose(f);
        throw Error("bkup_load: short bkup_data_00000.dat");
      }
    std::fclose(f);
    n = vi[0];
    cnt = vi[1];
    tim = vd[0];
    dt = vd[1];
    trec = vd[2];
    upload();
    return true;
  }

 protected:
  std::string path(const char* name) const { return fildir + "/" + name; }
  static FILE* open_or_throw(const std::string& p, const char* mode)
  {
    FILE* f = std::fopen(p.c_str(), mode);
    if (!f) throw Error("cannot open " + p);
    return f;
  }
  static void write_text(const std::string& p, const char* mode, const std::string& text)
  {
    FILE* f = open_or_throw(p, mode);
    std::fputs(text.c_str(), f);
    std::fclose(f);
  }
  // the 8 host arrays, nd doubles each, in val[] order (the dout_ format)
  void write_state(const std::string& p)
  {
    FILE* f = open_or_throw(p, "wb");
    for (int m = 0; m < nm; m++) std::fwrite(val[m], sizeof(double), store_.size() / nm, f);
    std::fclose(f);
  }
  // the same in FP32 (conv_d2f, common_func.c:22-26: a plain cast per value)
  void write_state_f32(const std::string& p)
  {
    FILE* f = open_or_throw(p, "wb");
    const size_t nd = store_.size() / nm;
    std::vector<float> buf(nd);
    for (int m = 0; m < nm; m++) {
      for (size_t i = 0; i < nd; i++) buf[i] = (float)val[m][i];
      std::fwrite(buf.data(), sizeof(float), nd, f);
    }
    std::fclose(f);
  }
  // bkup_save's files from the host arrays as they are
  void write_backup()
  {
    const double vald[3] = {tim, dt, trec};
    write_text(path("bkup_stamp.dat"), "w", std::to_string(n) + " " + std::to_string(cnt) + "\n");
    FILE* f = open_or_throw(path("bkup_time.dat"), "wb");
    std::fwrite(vald, sizeof(double), 3, f);
    std::fclose(f);
    write_state(path("bkup_data_00000.dat"));
  }
  // exec_ of the MPI drivers starts with bkup_(0) (mhd2d/mhd3d_class.cpp
  // of 2D/MPI and 3D/MPI): resume from the backup in fildir if one is there
  void resume_if_mpi()
  {
    if (mpi_files && !fildir.empty()) bkup_load();
  }
  // the output step of exec_ (state already downloaded): the MPI drivers
  // save a backup (bkup_(1)) before dout_
  void backup_if_mpi()
  {
    if (mpi_files && !fildir.empty()) write_backup();
  }
  // dout_ of the 2D/1D/3D serial drivers (e.g. 2D/SERIAL/OTvortex/
  // mhd2d_class.cpp:62-112): params.dat, offsets.dat, t.dat and the
  // coordinate files at n == 0, t.dat appended afterwards, and
  // outdat_%05d.dat (FP64 SoA incl. ghosts) at every call.
  // 1D (mhd1d_class.cpp:55-95) prints gam and t with "%.12f" and the
  // offset into xoff.dat; 2D with "%f" into offsets.dat.
  void dout_files(int msg, const std::vector<std::vector<double>>& coords, const std::vector<int>& offs)
  {
    if (fildir.empty()) return;
    if (mpi_files) return dout_files_mpi(msg, coords, offs);
    const bool one = offs.size() == 1;
    const char* tf = one ? "%.12f\n" : "%f\n";
    char buf[64];
    if (n == 0) {
      std::snprintf(buf, sizeof(buf), tf, gam);
      write_text(path("params.dat"), "w", buf);
      std::string o;
      for (size_t a = 0; a < offs.size(); a++) o += (a ? " " : "") + std::to_string(offs[a]);
      write_text(path(one ? "xoff.dat" : "offsets.dat"), "w", o + "\n");
      std::snprintf(buf, sizeof(buf), tf, tim);
      write_text(path("t.dat"), "w", buf);
      const char* names[3] = {"x.dat", "y.dat", "z.dat"};
      for (size_t a = 0; a < coords.size(); a++) {
        std::string t;
        for (double v : coords[a]) {
          std::snprintf(buf, sizeof(buf), "%.12f\n", v);
          t += buf;
        }
        write_text(path(names[a]), "w", t);
      }
    } else {
      std::snprintf(buf, sizeof(buf), tf, tim);
      write_text(path("t.dat"), "a", buf);
    }
    std::snprintf(buf, sizeof(buf), "outdat_%05d.dat", cnt);
    write_state(path(buf));
    if (msg) std::printf("Output data at t=%.4f (%d iterations).\n", tim, n);
  }
  // dout_ of the MPI drivers as rank 0 of a 1-rank decomposition
  // (2D/MPI/OTvortex/mhd2d_class.cpp dout_, 3D/MPI/OTvortex/
  // mhd3d_class.cpp:78-145): params.dat and t.dat with "%f", offsets.dat,
  // mpinum.dat "1 1[ 1]", x_00000.dat.. with "%.12f", and
  // outdat_%05d_00000.dat, the 8 arrays in FP32 (ghosts included).
  void dout_files_mpi(int msg, const std::vector<std::vector<double>>& coords, const std::vector<int>& offs)
  {
    char buf[64];
    if (n == 0) {
      std::snprintf(buf, sizeof(buf), "%f\n", gam);
      write_text(path("params.dat"), "w", buf);
      std::string o, r;
      for (size_t a = 0; a < offs.size(); a++) {
        o += (a ? " " : "") + std::to_string(offs[a]);
        r += a ? " 1" : "1";
      }
      write_text(path("offsets.dat"), "w", o + "\n");
      write_text(path("mpinum.dat"), "w", r + "\n");
      std::snprintf(buf, sizeof(buf), "%f\n", tim);
      write_text(path("t.dat"), "w", buf);
      const char* names[3] = {"x_00000.dat", "y_00000.dat", "z_00000.dat"};
      for (size_t a = 0; a < coords.size(); a++) {
        std::string t;
        for (double v : coords[a]) {
          std::snprintf(buf, sizeof(buf), "%.12f\n", v);
          t += buf;
        }
        write_text(path(names[a]), "w", t);
      }
    } else {
      std::snprintf(buf, sizeof(buf), "%f\n", tim);
      write_text(path("t.dat"), "a", buf);
    }
    std::snprintf(buf, sizeof(buf), "outdat_%05d_00000.dat", cnt);
    write_state_f32(path(buf));
    if (msg) std::printf("Output data at t=%.4f (%d iterations).\n", tim, n);
  }

 public:

 protected:
  // bound(val) of the reference: ghost fill on the device
  void bound() { dom().bound(); }
  // ideal(dt) of the reference: one RK step on the device
  void ideal(double d) { dom().step(d); }
};

// 2D driver: the members and methods of MHD2D (mhd2d_class.hpp:7-47).
class MHD2DDevice : public MHDDevice {
 public:
  const int xoff = 4, yoff = 4;
  const int nx, ny, nd;
  const int nout;
  const double dtrec, tmax, cfl;
  MHD2DDevice(int xmesh, int ymesh, int n_out, double dt_rec, double cfl_)
      : nx(xmesh + 8), ny(ymesh + 8), nd((xmesh + 8) * (ymesh + 8)), nout(n_out), dtrec(dt_rec),
        tmax(dt_rec * n_out), cfl(cfl_)
  {
    alloc(nd);
    trec = dtrec;
    nmax = nout;
  }
  // setdt (mhd2d_class.cpp:11-33): CFL dt on the device; nmax fixed by the
  // first call
  void setdt(int flg)
  {
    if (flg) dt = dom().cfl_dt();
    if (!nflg_) {
      nmax = (int)(tmax / dt + 0.5);
      nflg_ = 1;
    }
  }
  // exec_ (mhd2d_class.cpp:35-61): the step loop; output via dout_ (with
  // mpi_files, the MPI driver's exec_: backup resume and a backup per output)
  void exec_(int flg)
  {
    resume_if_mpi();
    if (n == 0) dout_(0);
    while (n++ < nmax && tim < tmax) {
      tim += dt;
      bound();
      ideal(dt);
      setdt(flg * (n % 2 == 0));
      if (tim >= trec) {
        cnt++;
        trec += dtrec;
        download();
        backup_if_mpi();
        dout_(1);
      }
    }
  }

 protected:
  double xmin = 0, xmax = 1, dx = 1, ymin = 0, ymax = 1, dy = 1, dr = 1;
  int nflg_ = 0;
  qas_grid grid() const override
  {
    qas_grid g{2, nx, ny, 1, xoff, yoff, 0, dx, dy, 1.0};
    return g;
  }
  double cfl_value() const override { return cfl; }
  // output (the reference's dout_ file set, when fildir is set)
  virtual void dout_(int msg)
  {
    std::vector<std::vector<double>> c(2);
    for (int i = 0; i < nx; i++) c[0].push_back((i - xoff) * dx + xmin);
    for (int j = 0; j < ny; j++) c[1].push_back((j - yoff) * dy + ymin);
    dout_files(msg, c, {xoff, yoff});
  }
};

// 1D driver (mhd1d_class.hpp:7-38): fixed-dt exec_(0) with nrec output
// steps, as the shock-tube drivers run.
class MHD1DDevice : public MHDDevice {
 public:
  const int xoff = 4;
  const int nx;
  const int nout;
  const double dtrec, tmax, cfl;
  MHD1DDevice(int xmesh, int n_out, double dt_rec, double cfl_)
      : nx(xmesh + 8), nout(n_out), dtrec(dt_rec), tmax(dt_rec * n_out), cfl(cfl_)
  {
    alloc(nx);
    trec = dtrec;
  }
  void setdt(int flg)
  {
    if (flg) dt = dom().cfl_dt();
    if (!nflg_) {
      nrec = (int)(dtrec / dt + 0.5);
      nmax = nrec * nout;
      nflg_ = 1;
    }
  }
  void exec_(int flg)
  {
    if (n == 0) dout_();
    while ((flg) ? (n++ < nmax && tim < tmax) : (n++ < nmax)) {
      tim += dt;
      bound();
      ideal(dt);
      setdt(flg * (n % 2 == 0));
      if ((flg) ? (tim >= trec) : ((n % nrec) == 0)) {
        cnt++;
        trec += dtrec;
        download();
        dout_();
      }
    }
  }

 protected:
  double xmin = 0, xmax = 1, dx = 1;
  int nrec = 1, nflg_ = 0;
  qas_grid grid() const override
  {
    qas_grid g{1, nx, 1, 1, xoff, 0, 0, dx, 1.0, 1.0};
    return g;
  }
  double cfl_value() const override { return cfl; }
  virtual void dout_()
  {
    std::vector<std::vector<double>> c(1);  // cell centres (mhd1d_paras.cpp:24)
    for (int i = 0; i < nx; i++) c[0].push_back((i - xoff + 0.5) * dx + xmin);
    dout_files(0, c, {xoff});
  }
};

// 3D driver (mhd3d_class.hpp:8-70) on one device-resident domain.
class MHD3DDevice : public MHDDevice {
 public:
  const int xoff = 4, yoff = 4, zoff = 4;
  const int nx, ny, nz;
  const long long nd;
  const int nout;
  const double dtrec, tmax, cfl;
  MHD3DDevice(int xmesh, int ymesh, int zmesh, int n_out, double dt_rec, double cfl_)
      : nx(xmesh + 8), ny(ymesh + 8), nz(zmesh + 8), nd((long long)(xmesh + 8) * (ymesh + 8) * (zmesh + 8)),
        nout(n_out), dtrec(dt_rec), tmax(dt_rec * n_out), cfl(cfl_)
  {
    alloc(nd);
    trec = dtrec;
    nmax = nout;
    mpi_files = true;  // the 3D reference driver is MPI only
  }
  void setdt(int flg)
  {
    if (flg) dt = dom().cfl_dt();
    if (!nflg_) {
      nmax = (int)(tmax / dt + 0.5);
      nflg_ = 1;
    }
  }
  void exec_(int flg)
  {
    resume_if_mpi();
    if (n == 0) dout_(0);
    while (n++ < nmax && tim < tmax) {
      tim += dt;
      bound();
      ideal(dt);
      setdt(flg * (n % 2 == 0));
      if (tim >= trec) {
        cnt++;
        trec += dtrec;
        download();
        backup_if_mpi();
        dout_(1);
      }
    }
  }

 protected:
  double xmin = 0, xmax = 1, dx = 1, ymin = 0, ymax = 1, dy = 1, zmin = 0, zmax = 1, dz = 1, dr = 1;
  int nflg_ = 0;
  qas_grid grid() const override
  {
    qas_grid g{3, nx, ny, nz, xoff, yoff, zoff, dx, dy, dz};
    return g;
  }
  double cfl_value() const override { return cfl; }
  virtual void dout_(int msg)
  {
    std::vector<std::vector<double>> c(3);
    for (int i = 0; i < nx; i++) c[0].push_back((i - xoff) * dx + xmin);
    for (int j = 0; j < ny; j++) c[1].push_back((j - yoff) * dy + ymin);
    for (int k = 0; k < nz; k++) c[2].push_back((k - zoff) * dz + zmin);
    dout_files(msg, c, {xoff, yoff, zoff});
  }
};

}  // namespace qasmhd
