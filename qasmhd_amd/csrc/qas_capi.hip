// qas_capi.hip -- extern "C" entry points of libqasmhd_amd (include/qasmhd.h).
//
// Owns the device state of one domain: three state buffers A, B, C for the
// RK stages (stage k reads one and writes the next; the step base U^n is A
// itself, read and rewritten element-wise by the last stage, so no U^n
// snapshot copy is made as cpy_array does in mhd2d_solve.cpp:38-41), the
// primitive scratch, the sweep accumulators and face EMF parts, and dt in
// device memory so that the exec_ loop runs without host round trips.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qas_internal.hpp"

namespace qas {
hipError_t launch_record_dt(qas_ctx* c, double* slot);
}

using namespace qas;

static int fail(qas_ctx* c, int code, const std::string& msg)
{
  if (c) c->err = msg;
  return code;
}

static int hipfail(qas_ctx* c, hipError_t e, const char* where)
{
  return fail(c, QAS_E_HIP, std::string(where) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

#define QAS_CHECK(c, expr)                              \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hipfail((c), _e, #expr); \
  } while (0)

extern "C" void qas_scheme_default(qas_scheme* s)
{
  std::memset(s, 0, sizeof(*s));
  s->gam = 5.0 / 3.0;
  s->cfl = 0.4;
  s->rmn = 1;
  s->odr = 4;
  s->r_k = 3;
  s->ctw = 1;
  s->st[0][4] = 1;
  s->st[1][5] = 1;
  s->st[2][6] = 1;
  s->precision = 0;
}

static std::string validate(const qas_grid* g, const qas_scheme* s)
{
  if (!g || !s) return "null grid or scheme";
  if (g->ndim < 1 || g->ndim > 3) return "ndim must be 1, 2 or 3";
  const int n[3] = {g->nx, g->ny, g->nz};
  const int off[3] = {g->xoff, g->yoff, g->zoff};
  const double d[3] = {g->dx, g->dy, g->dz};
  for (int a = 0; a < 3; a++) {
    if (a < g->ndim) {
      if (off[a] != NG) return "ghost width must be 4 on active axes";
      if (n[a] < 2 * NG + 8) return "each active axis needs at least 8 interior cells";
      if (!(d[a] > 0.0)) return "grid spacing must be positive";
    } else {
      if (n[a] != 1 || off[a] != 0) return "inactive axes must have n = 1 and no ghosts";
    }
  }
  if ((long long)g->nx * g->ny * g->nz > (1LL << 40)) return "grid too large";
  if (s->rmn < 0 || s->rmn > 3) return "rmn must be 0..3";
  if (s->odr < 1 || s->odr > 4) return "odr must be 1..4";
  if (s->r_k < 1 || s->r_k > 3) return "r_k must be 1..3";
  if (s->ctw != 0 && s->ctw != 1) return "ctw must be 0 or 1";
  if (s->precision < 0 || s->precision > 2)
    return "precision must be 0 (FP64), 1 (mixed) or 2 (FP64 with FMA contraction)";
  if (!(s->gam > 1.0)) return "gam must exceed 1";
  for (int a = 0; a < 3; a++)
    for (int m = 0; m < 8; m++) {
      const int dn = s->dn[a][m], st = s->st[a][m];
      if (dn < -2 || dn > 2) return "boundary flags must be in -2..2";
      if (st != 0 && st != 1) return "staggering flags must be 0 or 1";
    }
  // one device domain: an axis is periodic for every variable or for none
  // (the reference's MPI bound applies the periodic copy of dnxs[0] to all
  // variables, common_mpi.c:214-269)
  for (int a = 0; a < g->ndim; a++) {
    const bool p0 = s->dn[a][0] == 0;
    for (int m = 1; m < 8; m++)
      if ((s->dn[a][m] == 0) != p0) return "an axis must be periodic for all variables or for none";
  }
  if (g->ndim == 1 && s->precision != 0) return "the 1D path is FP64 only";
  return "";
}

namespace qas {
std::string validate_config(const qas_grid* g, const qas_scheme* s) { return validate(g, s); }

// Geometry and boundary flags of a context from a validated grid / scheme.
void ctx_init(qas_ctx* c, const qas_grid* g, const qas_scheme* s)
{
  c->g = *g;
  c->s = *s;
  Geo& geo = c->geo;
  geo.ndim = g->ndim;
  geo.nx = g->nx;
  geo.ny = g->ny;
  geo.nz = g->nz;
  geo.sy = g->nx;
  geo.sz = (long long)g->nx * g->ny;
  geo.nd = (long long)g->nx * g->ny * g->nz;
  geo.dx = g->dx;
  geo.dy = g->dy;
  geo.dz = g->dz;
  for (int a = 0; a < 3; a++)
    for (int m = 0; m < 8; m++) {
      c->bc.dn[a][m] = (signed char)s->dn[a][m];
      c->bc.st[a][m] = (signed char)s->st[a][m];
    }
}

// Device buffers of a context carved from one block: the three state
// buffers of the RK stages (the first is `state`, 8 x nd doubles, when
// given: the caller's own array), primitives, sweep accumulators, face
// parts of the B fluxes, edge EMFs, and a tail of small words (dt,
// dt-history counter, CFL partial maxima).  Arrays are padded to 512
// bytes.  Returns the bytes the block needs; assigns the pointers when
// base is not null.
size_t ctx_layout(qas_ctx* c, char* base, double* state)
{
  const int ndim = c->geo.ndim;
  const long long nd = c->geo.nd;
  const size_t stride = (size_t)((nd + 63) / 64 * 64);
  const int nprim = ndim == 3 ? 7 : (ndim == 2 ? 6 : 4);
  const int nacc = ndim == 1 ? 0 : (ndim == 2 ? 6 : 10);
  const int nemf = ndim == 3 ? 12 : (ndim == 2 ? 8 : 0);
  const int nedge = ndim == 3 ? 3 : (ndim == 2 ? 1 : 0);
  const size_t narr = (state ? 16 : 24) + nprim + nacc + nemf + nedge;
  const size_t bytes = narr * stride * sizeof(double) + 4096 + sizeof(double) * CFL_MAX_BLOCKS;
  if (!base) return bytes;
  double* p = (double*)base;
  for (int b = 0; b < 3; b++)
    for (int m = 0; m < 8; m++) {
      if (b == 0 && state) {
        c->U[0][m] = state + (size_t)m * nd;
        continue;
      }
      c->U[b][m] = p;
      p += stride;
    }
  c->vx = p; p += stride;
  c->vy = p; p += stride;
  c->vz = p; p += stride;
  c->pr = p; p += stride;
  if (ndim >= 2) { c->cx = p; p += stride; c->cy = p; p += stride; }
  if (ndim == 3) { c->cz = p; p += stride; }
  for (int k = 0; k < nacc; k++) { c->acc[k] = p; p += stride; }
  for (int k = 0; k < nemf; k++) { c->emf[k] = p; p += stride; }
  if (ndim == 2) { c->ez = p; p += stride; }
  if (ndim == 3) { c->ex = p; p += stride; c->ey = p; p += stride; c->ez = p; p += stride; }
  char* tail = (char*)p;
  c->d_dt = (double*)tail;
  c->d_hcount = (long long*)(tail + 128);
  c->d_cflpart = (double*)(tail + 4096);  // per-block maxima of the CFL reduction
  return bytes;
}
}  // namespace qas

extern "C" int qas_create(qas_ctx** out, const qas_grid* g, const qas_scheme* s)
{
  if (!out) return QAS_E_ARG;
  *out = nullptr;
  qas_ctx* c = new qas_ctx();
  std::string v = validate(g, s);
  if (!v.empty()) {
    c->err = v;
    *out = c;  // return the context so qas_last_error can report
    return QAS_E_ARG;
  }
  ctx_init(c, g, s);
  *out = c;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return hipfail(c, e, "hipGetDevice");
  QAS_CHECK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->slab_bytes = ctx_layout(c, nullptr, nullptr);
  e = hipMalloc(&c->slab, c->slab_bytes);
  if (e != hipSuccess) return fail(c, QAS_E_NOMEM, std::string("hipMalloc of the domain failed: ") + hipGetErrorName(e));
  ctx_layout(c, (char*)c->slab, nullptr);
  QAS_CHECK(c, hipMemsetAsync(c->slab, 0, c->slab_bytes, c->stream));
  QAS_CHECK(c, hipHostMalloc((void**)&c->h_pinned, 64 * sizeof(double), hipHostMallocDefault));
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  return QAS_OK;
}

extern "C" void qas_destroy(qas_ctx* c)
{
  if (!c) return;
  // teardown: errors here have nowhere to go
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  adj_release(c);
  dsp_release(c);
  grav_release(c);
  if (c->slab) (void)hipFree(c->slab);
  if (c->d_hist) (void)hipFree(c->d_hist);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* qas_last_error(const qas_ctx* c)
{
  return c ? c->err.c_str() : "null context";
}

static bool ready(qas_ctx* c) { return c && c->slab; }

extern "C" int qas_upload(qas_ctx* c, const double* const val[8])
{
  if (!ready(c) || !val) return fail(c, QAS_E_ARG, "qas_upload: null context or arrays");
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  for (int m = 0; m < 8; m++) {
    if (!val[m]) return fail(c, QAS_E_ARG, "qas_upload: null array");
    QAS_CHECK(c, hipMemcpyAsync(c->U[c->cur][m], val[m], bytes, hipMemcpyHostToDevice, c->stream));
  }
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->bounded = false;
  return QAS_OK;
}

extern "C" int qas_upload_packed(qas_ctx* c, const double* state)
{
  if (!ready(c) || !state) return fail(c, QAS_E_ARG, "qas_upload_packed: null argument");
  const double* v[8];
  for (int m = 0; m < 8; m++) v[m] = state + (size_t)m * c->geo.nd;
  return qas_upload(c, v);
}

extern "C" int qas_upload_device(qas_ctx* c, const double* state)
{
  if (!ready(c) || !state) return fail(c, QAS_E_ARG, "qas_upload_device: null argument");
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  for (int m = 0; m < 8; m++)
    QAS_CHECK(c, hipMemcpyAsync(c->U[c->cur][m], state + (size_t)m * c->geo.nd, bytes, hipMemcpyDeviceToDevice,
                                c->stream));
  c->uploaded = true;
  c->bounded = false;
  return QAS_OK;
}

extern "C" int qas_download(qas_ctx* c, double* const val[8])
{
  if (!ready(c) || !val) return fail(c, QAS_E_ARG, "qas_download: null context or arrays");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_download: no state uploaded");
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  for (int m = 0; m < 8; m++)
    if (!val[m]) return fail(c, QAS_E_ARG, "qas_download: null array");
  // on the context's stream (a null-stream copy would also wait for every
  // other context's work)
  for (int m = 0; m < 8; m++)
    QAS_CHECK(c, hipMemcpyAsync(val[m], c->U[c->cur][m], bytes, hipMemcpyDeviceToHost, c->stream));
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  return QAS_OK;
}

extern "C" int qas_download_packed(qas_ctx* c, double* state)
{
  if (!ready(c) || !state) return fail(c, QAS_E_ARG, "qas_download_packed: null argument");
  double* v[8];
  for (int m = 0; m < 8; m++) v[m] = state + (size_t)m * c->geo.nd;
  return qas_download(c, v);
}

extern "C" int qas_download_device(qas_ctx* c, double* state)
{
  if (!ready(c) || !state) return fail(c, QAS_E_ARG, "qas_download_device: null argument");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_download_device: no state uploaded");
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  for (int m = 0; m < 8; m++)
    QAS_CHECK(c, hipMemcpyAsync(state + (size_t)m * c->geo.nd, c->U[c->cur][m], bytes, hipMemcpyDeviceToDevice,
                                c->stream));
  return QAS_OK;
}

extern "C" int qas_bound(qas_ctx* c)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_bound: null context");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_bound: no state uploaded");
  QAS_CHECK(c, launch_bound(c, c->U[c->cur], 8));
  c->bounded = true;
  return QAS_OK;
}

// rk_fac of mhd2d_solve.cpp:7
static void rk_coef(int r_k, int rk, double* a, double* b)
{
  if (rk == 0) { *a = 0.0; *b = 1.0; }
  else if (rk == 1) { *a = 0.5 + (r_k - 2) * 0.25; *b = 0.5 - (r_k - 2) * 0.25; }
  else { *a = 1. / 3.; *b = 2. / 3.; }
}

// the pointers of RK stage rk reading buffer `in` and writing `out` (the
// step's input in buffer A), with the context's scratch and device dt
static StagePtrs stage_ptrs(const qas_ctx* c, int rk, int in, int out, int A)
{
  StagePtrs p;
  std::memset(&p, 0, sizeof(p));
  for (int m = 0; m < 8; m++) {
    p.in[m] = c->U[in][m];
    p.out[m] = c->U[out][m];
    p.u0[m] = c->U[A][m];
  }
  p.vx = c->vx; p.vy = c->vy; p.vz = c->vz; p.pr = c->pr;
  p.cx = c->cx; p.cy = c->cy; p.cz = c->cz;
  for (int k = 0; k < NACC; k++) p.acc[k] = c->acc[k];
  for (int k = 0; k < 12; k++) p.emf[k] = c->emf[k];
  p.ex = c->ex; p.ey = c->ey; p.ez = c->ez;
  p.dt = c->d_dt;
  p.phi = c->d_phi;
  p.enrt = c->d_enrt;
  rk_coef(c->s.r_k, rk, &p.a, &p.b);
  return p;
}

// ideal(dt) with dt already in device memory
int qas::ctx_step(qas_ctx* c)
{
  const int R = c->s.r_k;
  const int A = c->cur, B = (c->cur + 1) % 3, C = (c->cur + 2) % 3;
  int seq_in[3], seq_out[3];
  if (R == 1) { seq_in[0] = A; seq_out[0] = B; }
  else if (R == 2) { seq_in[0] = A; seq_out[0] = B; seq_in[1] = B; seq_out[1] = A; }
  else { seq_in[0] = A; seq_out[0] = B; seq_in[1] = B; seq_out[1] = C; seq_in[2] = C; seq_out[2] = A; }
  for (int rk = 0; rk < R; rk++) {
    const StagePtrs p = stage_ptrs(c, rk, seq_in[rk], seq_out[rk], A);
    hipError_t e;
    if (c->g.ndim == 1) e = launch_stage1d(c, p);
    else if (c->g.ndim == 2) e = launch_stage2d(c, p);  // ghost fill inside the CT launch
    else e = launch_stage3d(c, p);
    if (e != hipSuccess) return hipfail(c, e, "stage launch");
    if (c->g.ndim != 2) {
      e = launch_bound(c, c->U[seq_out[rk]], 8);
      if (e != hipSuccess) return hipfail(c, e, "bound launch");
    }
  }
  c->cur = seq_out[R - 1];
  c->bounded = true;
  return QAS_OK;
}

extern "C" int qas_step(qas_ctx* c, double dt)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_step: null context");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_step: no state uploaded");
  if (!(dt > 0.0)) return fail(c, QAS_E_ARG, "qas_step: dt must be positive");
  QAS_CHECK(c, launch_set_dt(c, dt));
  return ctx_step(c);
}

extern "C" int qas_cfl_dt(qas_ctx* c, double* dt_out)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_cfl_dt: null context");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_cfl_dt: no state uploaded");
  QAS_CHECK(c, launch_cfl(c, c->U[c->cur]));
  QAS_CHECK(c, launch_cfl_finalize(c, nullptr));
  QAS_CHECK(c, hipMemcpyAsync(c->h_pinned, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  if (dt_out) *dt_out = c->h_pinned[0];
  return QAS_OK;
}

// 2D launch-shape autotuner: x-sweep wavefronts per workgroup (1, 2, 4) and
// element-wise workgroup size (64, 128, 256).  Every shape gives the same
// arithmetic per cell (bit-identical results); they differ in how the
// workgroups fill the GPU, by about 1 %, so each timing brackets 3 steps'
// worth of RK stages with their ghost fills, the shapes are visited in turn
// twice after a warm-up, and the shape with the lowest time is kept.  The
// stages read the current state A (ghosts filled, dt on the device) and
// run A -> B -> C -> B: the last stage writes the buffer of the first
// instead of A, so the state is not changed (the run overwrites B, C and
// the scratch).  Timing stage 1 alone ranked the shapes differently from
// whole runs.
int qas::ctx_tune(qas_ctx* c)
{
  c->tuned = true;
  if (c->g.ndim != 2) return QAS_OK;  // one shape in 1D and 3D
  const int cand[9][2] = {{4, 256}, {2, 256}, {1, 256}, {4, 128}, {2, 128}, {1, 128}, {4, 64}, {2, 64}, {1, 64}};
  const int A = c->cur, B = (c->cur + 1) % 3;
  constexpr int NSHAPE = 9, NREP = 3, NROUND = 2;
  const int R = c->s.r_k;
  const int C = (c->cur + 2) % 3;
  // stage rk reads s_in[rk] and writes s_out[rk] (R_K 1: A -> B; 2: A -> B -> C;
  // 3: A -> B -> C -> B)
  const int s_in[3] = {A, B, C}, s_out[3] = {B, C, B};
  StagePtrs ps[3];
  for (int rk = 0; rk < R; rk++) ps[rk] = stage_ptrs(c, rk, s_in[rk], s_out[rk], A);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  QAS_CHECK(c, hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  float tmin[NSHAPE];
  for (int k = 0; k < NSHAPE; k++) tmin[k] = -1.0f;
  for (int round = -1; round < NROUND && e == hipSuccess; round++) {  // round -1: warm-up
    for (int k = 0; k < NSHAPE && e == hipSuccess; k++) {
      c->xwpb = cand[k][0];
      c->ewb = cand[k][1];
      e = hipEventRecord(e0, c->stream);
      for (int rep = 0; rep < NREP && e == hipSuccess; rep++)
        for (int rk = 0; rk < R && e == hipSuccess; rk++) {
          e = launch_stage2d(c, ps[rk]);
        }
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess && round >= 0 && (tmin[k] < 0.0f || ms < tmin[k])) tmin[k] = ms;
    }
  }
  int bi = 0;
  for (int k = 1; k < NSHAPE; k++)
    if (tmin[k] >= 0.0f && tmin[k] < tmin[bi]) bi = k;
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  c->xwpb = cand[bi][0];
  c->ewb = cand[bi][1];
  c->graph_flg = -1;  // a captured graph has the old grid sizes
  if (e != hipSuccess) return hipfail(c, e, "qas_tune");
  return QAS_OK;
}

extern "C" int qas_tune(qas_ctx* c)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_tune: null context");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_tune: no state uploaded");
  if (c->prof_mask) return fail(c, QAS_E_STATE, "qas_tune: timing hooks are on");
  if (!c->bounded) {
    QAS_CHECK(c, launch_bound(c, c->U[c->cur], 8));
    c->bounded = true;
  }
  return ctx_tune(c);
}

extern "C" int qas_set_shape(qas_ctx* c, int xwpb, int ewb)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_set_shape: null context");
  if ((xwpb != 1 && xwpb != 2 && xwpb != 4) || (ewb != 64 && ewb != 128 && ewb != 256))
    return fail(c, QAS_E_ARG, "qas_set_shape: xwpb 1/2/4, ewb 64/128/256");
  c->xwpb = xwpb;
  c->ewb = ewb;
  c->tuned = true;
  c->graph_flg = -1;
  return QAS_OK;
}

extern "C" int qas_shape(const qas_ctx* c, int* xwpb, int* ewb)
{
  if (!c || !xwpb || !ewb) return QAS_E_ARG;
  *xwpb = c->xwpb;
  *ewb = c->ewb;
  return QAS_OK;
}

// One step of exec_'s loop body on the device: [record dt] bound (if the
// state is not known to be bounded) ideal(dt) [setdt when flg and n even].
static int loop_step(qas_ctx* c, int flg, bool hist, bool even)
{
  if (hist) QAS_CHECK(c, launch_record_counted(c));
  // exec_ bounds every step; the bound is skipped when nothing touched the
  // state since the last one -- not with gravity, whose setdt rewrites the
  // interior energy (so a captured step pair always carries it)
  if (!c->bounded || c->d_phi) QAS_CHECK(c, launch_bound(c, c->U[c->cur], 8));
  int r = ctx_step(c);
  if (r) return r;
  // DMHD2D::exec_: the sub-cycled dissipative step after ideal(dt)
  r = dsp_after_ideal(c);
  if (r) return r;
  if (flg && even) {
    QAS_CHECK(c, launch_cfl(c, c->U[c->cur]));
    QAS_CHECK(c, launch_cfl_finalize(c, nullptr));
  }
  return QAS_OK;
}

// Capture two consecutive steps (odd, even) as one graph.  The kernels take
// dt from device memory and the buffers rotate back to the same current
// buffer after R_K = 2 or 3 stages, so one graph serves every pair.
static int build_pair_graph(qas_ctx* c, int flg, bool hist)
{
  // an earlier replay of the old graph may still run on the stream: wait
  // for it before its executable is destroyed
  if (c->gexec || c->graph) QAS_CHECK(c, hipStreamSynchronize(c->stream));
  if (c->gexec) { QAS_CHECK(c, hipGraphExecDestroy(c->gexec)); c->gexec = nullptr; }
  if (c->graph) { QAS_CHECK(c, hipGraphDestroy(c->graph)); c->graph = nullptr; }
  const int cur0 = c->cur;
  const bool b0 = c->bounded;
  QAS_CHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int r = loop_step(c, flg, hist, false);
  if (!r) r = loop_step(c, flg, hist, true);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->cur = cur0;  // capture does not execute; restore the host view
  c->bounded = b0;
  if (r) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
  c->graph = g;
  QAS_CHECK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  c->graph_flg = flg;
  c->graph_hist = hist ? 1 : 0;
  c->graph_cur = cur0;
  return QAS_OK;
}

extern "C" int qas_run(qas_ctx* c, int nsteps, int flg, double* dt_inout, double* dt_hist)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_run: null context");
  if (!c->uploaded) return fail(c, QAS_E_STATE, "qas_run: no state uploaded");
  if (nsteps < 0 || !dt_inout) return fail(c, QAS_E_ARG, "qas_run: bad arguments");
  const bool hist = dt_hist != nullptr;
  if (hist && nsteps > c->hist_cap) {
    if (c->d_hist) QAS_CHECK(c, hipFree(c->d_hist));
    c->d_hist = nullptr;
    QAS_CHECK(c, hipMalloc(&c->d_hist, sizeof(double) * (size_t)nsteps));
    c->hist_cap = nsteps;
    c->graph_hist = -1;  // the graph holds the old history pointer
  }
  if (hist) QAS_CHECK(c, launch_zero_counter(c));
  if (*dt_inout > 0.0) {
    QAS_CHECK(c, launch_set_dt(c, *dt_inout));
  } else {
    QAS_CHECK(c, launch_cfl(c, c->U[c->cur]));
    QAS_CHECK(c, launch_cfl_finalize(c, nullptr));
  }
  // the first run of a 2D context picks its launch shape (qas_tune)
  if (!c->tuned && c->g.ndim == 2 && c->prof_mask == 0 && nsteps > 0) {
    if (!c->bounded) {
      QAS_CHECK(c, launch_bound(c, c->U[c->cur], 8));
      c->bounded = true;
    }
    const int rt = ctx_tune(c);
    if (rt) return rt;
  }
  int n = 0;
  // bring the running counter to an even count before pairing steps
  if (n < nsteps && (c->nstep % 2) == 1) {
    c->nstep++;
    int r = loop_step(c, flg, hist, true);
    if (r) return r;
    n++;
  }
  // (the dissipative sub-cycle count depends on dt: no graph then)
  const bool graphs = c->use_graph && c->s.r_k >= 2 && c->prof_mask == 0 && !c->d_nu;
  if (graphs && nsteps - n >= 2) {
    if (!c->bounded) {
      QAS_CHECK(c, launch_bound(c, c->U[c->cur], 8));
      c->bounded = true;
    }
    if (!c->gexec || c->graph_flg != flg || c->graph_hist != (hist ? 1 : 0) || c->graph_cur != c->cur) {
      int r = build_pair_graph(c, flg, hist);
      if (r) return r;
    }
    for (; n + 2 <= nsteps; n += 2) {
      QAS_CHECK(c, hipGraphLaunch(c->gexec, c->stream));
      c->nstep += 2;
    }
  }
  for (; n < nsteps; n++) {
    c->nstep++;
    int r = loop_step(c, flg, hist, c->nstep % 2 == 0);
    if (r) return r;
  }
  QAS_CHECK(c, hipMemcpyAsync(c->h_pinned, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  *dt_inout = c->h_pinned[0];
  if (hist && nsteps > 0) {
    QAS_CHECK(c, hipMemcpyAsync(dt_hist, c->d_hist, sizeof(double) * (size_t)nsteps, hipMemcpyDeviceToHost,
                                c->stream));
    QAS_CHECK(c, hipStreamSynchronize(c->stream));
  }
  return QAS_OK;
}

extern "C" int qas_use_graphs(qas_ctx* c, int on)
{
  if (!c) return QAS_E_ARG;
  c->use_graph = on != 0;
  return QAS_OK;
}

extern "C" int qas_reset_counter(qas_ctx* c)
{
  if (!c) return QAS_E_ARG;
  c->nstep = 0;
  return QAS_OK;
}

extern "C" int qas_sync(qas_ctx* c)
{
  if (!ready(c)) return fail(c, QAS_E_ARG, "qas_sync: null context");
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  return QAS_OK;
}

extern "C" long long qas_device_bytes(const qas_ctx* c) { return c ? (long long)c->slab_bytes : 0; }

extern "C" int qas_prof_enable(qas_ctx* c, unsigned mask, int capacity)
{
  if (!ready(c) || capacity < 0) return fail(c, QAS_E_ARG, "qas_prof_enable: bad arguments");
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  c->prof_mask = 0;
  c->prof_n = 0;
  if (mask == 0) return QAS_OK;
  while ((int)c->prof_kid.size() < capacity) {
    hipEvent_t a, b;
    QAS_CHECK(c, hipEventCreate(&a));
    QAS_CHECK(c, hipEventCreate(&b));
    c->prof_ev.push_back(a);
    c->prof_ev.push_back(b);
    c->prof_kid.push_back(-1);
  }
  c->prof_mask = mask;
  return QAS_OK;
}

extern "C" int qas_prof_query(qas_ctx* c, int kid, double* total_ms, long long* launches)
{
  if (!ready(c) || !total_ms || !launches) return fail(c, QAS_E_ARG, "qas_prof_query: bad arguments");
  QAS_CHECK(c, hipStreamSynchronize(c->stream));
  double tot = 0.0;
  long long n = 0;
  for (size_t r = 0; r < c->prof_n; r++) {
    if (c->prof_kid[r] != kid) continue;
    float ms = 0.0f;
    QAS_CHECK(c, hipEventElapsedTime(&ms, c->prof_ev[2 * r], c->prof_ev[2 * r + 1]));
    tot += ms;
    n++;
  }
  *total_ms = tot;
  *launches = n;
  return QAS_OK;
}
