// qas_internal.hpp -- device context and kernel-launch interface shared by
// the HIP translation units of libqasmhd_amd.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "../../include/qasmhd.h"

namespace qas {

constexpr int NG = 4;       // ghost layers (mhd2d_class.hpp:10)
constexpr int SEG = 60;     // faces owned by one wavefront segment (64 lanes - 4 halo)
constexpr int WAVE = 64;
constexpr int CFL_MAX_BLOCKS = 1024;
// accumulator slots: 2D uses 6 (dt/dy Dy of ro mx my mz bz en); 3D uses
// 5 + 5 (dt/dy Dy and dt/dz Dz of ro mx my mz en)
constexpr int NACC = 12;  // blocks of the CFL reduction (grid-stride)

// Boundary flags as kernel arguments (dnxs/dnys/dnzs, stxs/stys/stzs).
struct BcFlags {
  signed char dn[3][8];
  signed char st[3][8];
};

// Geometry as kernel arguments.
struct Geo {
  int ndim, nx, ny, nz;
  long long sy, sz, nd;  // strides of y and z, cells per array
  double dx, dy, dz;
};

// Pointers of one stage, passed by value.
struct StagePtrs {
  const double* in[8];   // stage input state (ghosts filled)
  double* out[8];        // stage output state
  const double* u0[8];   // state at the start of the step (RK base)
  double *vx, *vy, *vz, *pr, *cx, *cy, *cz;  // primitive scratch
  double* acc[NACC];     // flux divergences of the sweeps run before the x sweep
  double* emf[12];       // face parts of the B fluxes: [sweep][comp][upwind/central]
  double *ex, *ey, *ez;  // edge EMFs
  const double* dt;      // device dt
  double a, b;           // RK stage coefficients (rk_fac)
  const double* phi;     // 2D gravitational potential per cell (gravity runs only)
  double* enrt;          // the stage energy after the potential round trip (gravity)
};

// The primitive scratch as the sweeps read it: element type PT = the
// reconstruction precision (double: FP64 path; float: mixed path, where the
// prim kernels round the FP64 primitives once on store).
template <typename PT>
struct PrimView {
  const PT *vx, *vy, *vz, *pr, *cx, *cy, *cz;
};
template <typename PT>
__host__ __device__ inline PrimView<PT> prim_view(const StagePtrs& p)
{
  return PrimView<PT>{reinterpret_cast<const PT*>(p.vx), reinterpret_cast<const PT*>(p.vy),
                      reinterpret_cast<const PT*>(p.vz), reinterpret_cast<const PT*>(p.pr),
                      reinterpret_cast<const PT*>(p.cx), reinterpret_cast<const PT*>(p.cy),
                      reinterpret_cast<const PT*>(p.cz)};
}
// XCD-aware tile order: the hardware hands consecutive workgroups to the 8
// XCDs in turn; remap block b of nb so that each XCD runs one contiguous
// range of logical tiles (neighbouring column groups then share that XCD's
// L2 lines).
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
  const int x = b % 8, slot = b / 8;
  const int q = nb / 8, r = nb % 8;
  return (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// The face parts of the B fluxes (emf) and the stored transverse
// divergences (acc) in element type ST = the sweep precision: double on the
// FP64 path; float on the mixed path, where the face parts are FP32 values
// already and the divergences are rounded once on store.
// G(p): p as a global-memory pointer.  Every array of a stage is in global
// memory, but a pointer read from device memory (a batch's instance table)
// is generic to the compiler, whose flat loads and stores also count
// against the LDS wait counter; accesses through G(p) are global ones.
template <typename T>
__host__ __device__ __forceinline__ auto G(T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return (__attribute__((address_space(1))) T*)p;
#else
  return p;  // host runs of the same bodies (tests)
#endif
}
template <typename ST>
__device__ __forceinline__ void store_as(double* a, long long i, double v)
{
  G(reinterpret_cast<ST*>(a))[i] = (ST)v;
}
template <typename ST>
__device__ __forceinline__ double load_as(const double* a, long long i)
{
  return (double)G(reinterpret_cast<const ST*>(a))[i];
}

template <typename PT>
__device__ __forceinline__ void store_prim(double* a, long long t, double v)
{
  reinterpret_cast<PT*>(a)[t] = (PT)v;
}

// Ghost-index resolution: bound() writes ghost cells as (factor x) a kept
// cell of the same array.  Applying x then y then z (bc2d/bc3d order) to a
// cell gives value = kz*(ky*(kx*f[src])) with the sources below.
struct Src {
  int i;
  double k;
  bool mul;
};

__host__ __device__ __forceinline__ Src resolve(int i, int n, int st, int dn)
{
  Src r{i, 1.0, false};
  if (dn == 0) {  // periodic: plain copy (common_func.c:67-74)
    if (i < NG) r.i = n - 2 * NG + i;
    else if (i >= n - NG) r.i = i - (n - 2 * NG);
  } else if (dn == 1 || dn == -1) {  // mirror x dn (common_func.c:75-86)
    if (i < NG) {
      r.i = (2 * NG - 1 + st) - i;
      r.k = (double)dn;
      r.mul = true;
    } else if (i >= n - NG + st) {
      r.i = (n - 2 * NG + st) + (n - 1 - i);
      r.k = (double)dn;
      r.mul = true;
    }
  } else if (dn == 2 || dn == -2) {  // open / zero-fix (common_func.c:87-98)
    if (i < NG) {
      r.i = NG;
      r.k = 0.25 * (double)(2 + dn);
      r.mul = true;
    } else if (i >= n - NG + st) {
      r.i = n - 1 - NG + st;
      r.k = 0.25 * (double)(2 + dn);
      r.mul = true;
    }
  }
  return r;
}

__host__ __device__ __forceinline__ double apply(const Src& s, double v) { return s.mul ? s.k * v : v; }

}  // namespace qas

struct qas_ctx {
  qas_grid g;
  qas_scheme s;
  qas::Geo geo;
  qas::BcFlags bc;
  int device = 0;
  hipStream_t stream = nullptr;
  void* slab = nullptr;
  size_t slab_bytes = 0;
  double* U[3][8] = {};   // U[cur] holds the current state
  int cur = 0;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr, *pr = nullptr, *cx = nullptr, *cy = nullptr,
         *cz = nullptr;
  double* acc[qas::NACC] = {};
  double* emf[12] = {};
  double *ex = nullptr, *ey = nullptr, *ez = nullptr;
  double* d_dt = nullptr;             // current dt
  double* d_hist = nullptr;           // dt history of qas_run
  int hist_cap = 0;
  double* d_cflpart = nullptr;        // CFL_MAX_BLOCKS per-block maxima
  int cfl_blocks = 1;                 // blocks of the last CFL launch
  long long* d_hcount = nullptr;      // dt-history slot counter (graph replays)
  // hipGraph of two consecutive steps of qas_run (qas_use_graphs(c, 1));
  // off by default: host launches measured as fast (the host enqueues
  // ahead of the GPU, so replay saves nothing)
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_flg = -1, graph_hist = -1, graph_cur = -1;
  bool use_graph = false;
  // 2D launch shape (qas_tune): x-sweep wavefronts per workgroup and the
  // workgroup size of the element-wise kernels (EMF, CT); every shape gives
  // the same results
  int xwpb = 4, ewb = 256;
  bool tuned = false;
  double* h_pinned = nullptr;         // small pinned staging
  bool uploaded = false;
  bool bounded = false;
  long long nstep = 0;                // running step counter (exec_'s n)
  std::string err;
  // timing hooks (qas_prof_enable)
  unsigned prof_mask = 0;
  std::vector<hipEvent_t> prof_ev;    // start/stop pairs
  std::vector<int> prof_kid;
  size_t prof_n = 0;
  // discrete-adjoint workspace (qas_adjoint2d.hip), created on first use;
  // adj_keep: the next qas_adj_forward keeps every stage's intermediates
  struct qas_adj_work* adj = nullptr;
  int adj_keep = 0;
  // dissipative step (qas_dissipative.hip): per-cell viscosity and
  // resistivity (one allocation, eta follows nu) and max(2 nu, eta)
  double *d_nu = nullptr, *d_eta = nullptr;
  double dcmax = 0.0;
  // gravity (2D): potential per cell and the stage energy after its round
  // trip (one allocation, d_enrt follows d_phi)
  double *d_phi = nullptr, *d_enrt = nullptr;
};

namespace qas {
// The event pair of the next timed launch of kernel id kid, or false when
// that kernel is not timed (qas_prof_enable).  The events are passed to
// hipExtLaunchKernel, which stamps them with the start and end of the
// kernel's own dispatch (not with markers enqueued around it).
inline bool prof_slot(qas_ctx* c, int kid, hipEvent_t* e0, hipEvent_t* e1)
{
  if (!((c->prof_mask >> kid) & 1u) || c->prof_n >= c->prof_kid.size()) return false;
  *e0 = c->prof_ev[2 * c->prof_n];
  *e1 = c->prof_ev[2 * c->prof_n + 1];
  c->prof_kid[c->prof_n++] = kid;
  return true;
}
}  // namespace qas

// launch KERNEL; a kernel id the timing hooks select goes through
// hipExtLaunchKernelGGL with the next event pair
#define QAS_LAUNCH(c, KID, KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)                                      \
  do {                                                                                                   \
    hipEvent_t qas_e0_, qas_e1_;                                                                         \
    if (qas::prof_slot((c), (KID), &qas_e0_, &qas_e1_))                                                 \
      hipExtLaunchKernelGGL(KERNEL, dim3(GRID), dim3(BLOCK), (SHMEM), (STREAM), qas_e0_, qas_e1_, 0u,     \
                            __VA_ARGS__);                                                                \
    else                                                                                                 \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, __VA_ARGS__);                               \
  } while (0)

// ---- launchers (defined in the per-dimension translation units) ----
namespace qas {
hipError_t launch_bound(qas_ctx* c, double* const f[8], int nvar);
hipError_t launch_cfl(qas_ctx* c, const double* const U[8]);
hipError_t launch_cfl_finalize(qas_ctx* c, double* hist_slot);
hipError_t launch_set_dt(qas_ctx* c, double dt);
hipError_t launch_record_counted(qas_ctx* c);
hipError_t launch_zero_counter(qas_ctx* c);
hipError_t launch_stage1d(qas_ctx* c, const StagePtrs& p);
// faces_only: primitives and both sweeps (face parts, stage sum) but no
// corner EMF / CT (the adjoint's recompute of a step's last stage)
// (the CT launch also fills the ghosts of p.out: bound() of the stage
// output, so no launch_bound follows)
hipError_t launch_stage2d(qas_ctx* c, const StagePtrs& p, bool faces_only = false);
hipError_t launch_stage3d(qas_ctx* c, const StagePtrs& p);
void adj_release(qas_ctx* c);  // frees the adjoint workspace (qas_adjoint2d.hip)
// context set-up shared by qas_create and the allocation-free step
// (qas_capi.hip): configuration check ("" when valid), geometry, buffer
// layout over one device block, and ideal(dt) with dt in device memory
std::string validate_config(const qas_grid* g, const qas_scheme* s);
void ctx_init(qas_ctx* c, const qas_grid* g, const qas_scheme* s);
size_t ctx_layout(qas_ctx* c, char* base, double* state);
int ctx_step(qas_ctx* c);
int ctx_tune(qas_ctx* c);  // picks the 2D launch shape (qas_tune)
// dissipative step (qas_dissipative.hip)
void dsp_release(qas_ctx* c);
// gravity (qas_gravity.hip)
void grav_release(qas_ctx* c);
int dsp_step(qas_ctx* c, double dt);
int dsp_after_ideal(qas_ctx* c);
// FP32 face sweeps of the 3D mixed path (qas_mixed.hip) and the contracted
// FP64 ones (qas_fma64.hip)
void (*pick_sweep_f32(int ndim, int dir, int rmn, int odr))(StagePtrs, Geo, BcFlags, double, int);
void (*pick_sweep_fma(int ndim, int dir, int rmn, int odr))(StagePtrs, Geo, BcFlags, double, int);
}  // namespace qas
