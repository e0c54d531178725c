// qas_batch.hip -- batched 2D domains: qas_batch_create / qas_batch_run.
//
// A batch advances n independent 2D domains (each a qas_ctx with its own
// grid size, state and dt) with one launch per stage kernel for all of them
// (qas_batch.hpp), and captures two consecutive steps of those launches in
// one hipGraph that the run replays.  Per instance the result is the one
// qas_run would give: the same kernels bodies, the same dt schedule (CFL
// every second step when flg), the same buffer rotation.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "qas_adjbatch.hpp"
#include "qas_batch.hpp"
#include "qas_common.hpp"
#include "qas_device.hpp"
#include "qas_internal.hpp"
#include "qas_stage2d.hpp"
#include "qas_sweep2d.hpp"

namespace qas {

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ long long batch_thread(const int* off, int i, int b)
{
  return (long long)(b - off[i]) * blockDim.x + threadIdx.x;
}

// corner EMF, two corners of a column per thread (the batch's launch is
// one large grid however small its domains)
template <int ODR, typename FT>
__global__ __launch_bounds__(256) void kb_emf2d(const BatchInst* I, const int* off, int n, int stage, int ctw)
{
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x), i = batch_find(off, n, b);
  emf2d_col<ODR, FT, 2>(I[i].st[stage], I[i].g, ctw, batch_thread(off, i, b));
}

template <int ODR>
__global__ __launch_bounds__(256) void kb_ct2d(const BatchInst* I, const int* off, int n, int stage)
{
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x), i = batch_find(off, n, b);
  ct2d_at<ODR>(I[i].st[stage], I[i].g, batch_thread(off, i, b));
}

// ghost fill of stage `stage`'s output, or of the step's state (stage < 0)
__global__ __launch_bounds__(256) void kb_bound(const BatchInst* I, const int* off, int n, int stage)
{
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x), i = batch_find(off, n, b);
  const BatchInst& B = I[i];
  bound_at_f([&](int m) { return stage < 0 ? B.cur[m] : B.st[stage].out[m]; }, 8, nullptr, 0.0, B.g, B.bc, B.nghost,
             NG, NG, 0, batch_thread(off, i, b));
}

__global__ __launch_bounds__(256) void kb_cfl(const BatchInst* I, const int* off, int n, double gam)
{
  const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x), i = batch_find(off, n, b);
  const BatchInst& B = I[i];
  CflArgs a;
  for (int m = 0; m < 8; m++) a.u[m] = B.cur[m];
  a.part = B.cflpart;
  a.gam = gam;
  a.stx4 = B.bc.st[0][4];
  a.sty5 = B.bc.st[1][5];
  a.stz6 = B.bc.st[2][6];
  cfl_block(a, B.g, b - off[i], off[i + 1] - off[i]);
}

// one block per instance
__global__ __launch_bounds__(256) void kb_cfl_final(const BatchInst* I, int n)
{
  (void)n;  // the grid has one block per instance
  const BatchInst& B = I[blockIdx.x];
  cfl_final_block(B.cflpart, B.cflblocks, B.dt, nullptr, B.cfl_dr);
}

// one thread per instance
__global__ void kb_set_dt(const BatchInst* I, int n, const double* dts)
{
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n && dts[i] > 0.0) *I[i].dt = dts[i];
}
__global__ void kb_record(const BatchInst* I, int n)
{
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const long long k = *I[i].hcount;
  if (k < I[i].hcap) I[i].hist[k] = *I[i].dt;
  *I[i].hcount = k + 1;
}
__global__ void kb_zero_counts(const BatchInst* I, int n)
{
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) *I[i].hcount = 0;
}
__global__ void kb_get_dt(const BatchInst* I, int n, double* out)
{
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) out[i] = *I[i].dt;
}

// ---- batched adjoint: per-instance buffers of a recorded forward
struct AdjStepInst {
  double* ck;           // checkpoints: step n, variable m at ck + (8 n + m) * stride
  long long stride, nd;
  const double* dts;    // recorded step sizes
  double* dslot;        // the step size the backward's kernels read
  double* dt;           // the context's dt (read by the forward stages)
  double* cur[8];       // the context's state (forward)
  double* ub[8];        // the backward's step input buffer
};

// forward of step `step`: checkpoint of every instance's state and its dt
__global__ __launch_bounds__(256) void kb_adj_ckpt(const AdjStepInst* A, const int* off, int n, int step)
{
  const int b = (int)blockIdx.x, i = batch_find(off, n, b);
  const AdjStepInst& a = A[i];
  const long long e = (long long)(b - off[i]) * blockDim.x + threadIdx.x;
  if (e == 0) *a.dt = a.dts[step];
  if (e >= 8 * a.nd) return;
  const int m = (int)(e / a.nd);
  const long long q = e - (long long)m * a.nd;
  a.ck[(8LL * step + m) * a.stride + q] = a.cur[m][q];
}

// backward of step `step`: its input state and dt into the fixed buffers
__global__ __launch_bounds__(256) void kb_adj_load(const AdjStepInst* A, const int* off, int n, int step)
{
  const int b = (int)blockIdx.x, i = batch_find(off, n, b);
  const AdjStepInst& a = A[i];
  const long long e = (long long)(b - off[i]) * blockDim.x + threadIdx.x;
  if (e == 0) *a.dslot = a.dts[step];
  if (e >= 8 * a.nd) return;
  const int m = (int)(e / a.nd);
  const long long q = e - (long long)m * a.nd;
  a.ub[m][q] = a.ck[(8LL * step + m) * a.stride + q];
}

typedef void (*emf_fn)(const BatchInst*, const int*, int, int, int);
typedef void (*ct_fn)(const BatchInst*, const int*, int, int);

}  // namespace qas

using namespace qas;

struct qas_batch {
  std::vector<qas_ctx*> c;
  int n = 0;
  qas_scheme s;
  hipStream_t stream = nullptr;
  BatchInst* d_inst = nullptr;  // n descriptors
  int* d_off = nullptr;         // BK_COUNT tables of n + 1 offsets
  double* d_dts = nullptr;      // n step sizes (in / out)
  int nblk[BK_COUNT] = {};
  std::vector<BatchInst> inst;
  xsweep2d_fn sx = nullptr;
  xsweep2d_fn sy = nullptr;
  emf_fn emf = nullptr;
  ct_fn ct = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_flg = -1, graph_hist = -1, graph_pairs = 0;
  bool use_graph = true;  // qas_batch_use_graphs
  long long nstep = 0;
  // launch shape (qas_batch_tune): block order of the instances (0 as given,
  // 1 largest first), wavefronts per x-sweep workgroup, threads per
  // workgroup of the element-wise stage kernels
  int order = 0, xwpb = 4, ewb = 256;
  bool tuned = false;
  std::vector<int> perm;  // table position -> context index
  std::string err;
};

static int bfail(qas_batch* b, int code, const std::string& m)
{
  if (b) b->err = m;
  return code;
}

#define QAS_BCHECK(b, expr)                                                                            \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) return bfail((b), QAS_E_HIP, std::string(#expr) + ": " + hipGetErrorName(_e)); \
  } while (0)

static void rk_ab(int r_k, int rk, double* a, double* b)
{
  // rk_fac of mhd2d_solve.cpp:7
  if (rk == 0) { *a = 0.0; *b = 1.0; }
  else if (rk == 1) { *a = 0.5 + (r_k - 2) * 0.25; *b = 0.5 - (r_k - 2) * 0.25; }
  else { *a = 1. / 3.; *b = 2. / 3.; }
}

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// descriptors from the contexts' current buffers (rebuilt every run)
static int fill_table(qas_batch* b)
{
  const int n = b->n;
  const int R = b->s.r_k;
  b->inst.assign(n, BatchInst());
  b->perm.resize(n);
  for (int i = 0; i < n; i++) b->perm[i] = i;
  if (b->order == 1)  // largest domains first: the last workgroups of a launch are the small ones
    std::stable_sort(b->perm.begin(), b->perm.end(),
                     [b](int x, int y) { return b->c[x]->geo.nd > b->c[y]->geo.nd; });
  std::vector<int> off((size_t)BK_COUNT * (n + 1), 0);
  for (int i = 0; i < n; i++) {
    qas_ctx* c = b->c[b->perm[i]];
    BatchInst& B = b->inst[i];
    std::memset(&B, 0, sizeof(B));
    const Geo& g = c->geo;
    const int A = c->cur, Bb = (c->cur + 1) % 3, Cc = (c->cur + 2) % 3;
    int in_[3], out_[3];
    if (R == 2) { in_[0] = A; out_[0] = Bb; in_[1] = Bb; out_[1] = A; }
    else { in_[0] = A; out_[0] = Bb; in_[1] = Bb; out_[1] = Cc; in_[2] = Cc; out_[2] = A; }
    for (int rk = 0; rk < R; rk++) {
      StagePtrs& p = B.st[rk];
      for (int m = 0; m < 8; m++) {
        p.in[m] = c->U[in_[rk]][m];
        p.out[m] = c->U[out_[rk]][m];
        p.u0[m] = c->U[A][m];
      }
      p.vx = c->vx; p.vy = c->vy; p.vz = c->vz; p.pr = c->pr; p.cx = c->cx; p.cy = c->cy;
      for (int k = 0; k < NACC; k++) p.acc[k] = c->acc[k];
      for (int k = 0; k < 8; k++) p.emf[k] = c->emf[k];
      p.ez = c->ez;
      p.dt = c->d_dt;
      rk_ab(R, rk, &p.a, &p.b);
    }
    B.g = g;
    B.bc = c->bc;
    for (int m = 0; m < 8; m++) B.cur[m] = c->U[A][m];
    B.cflpart = c->d_cflpart;
    B.dt = c->d_dt;
    B.hist = c->d_hist;
    B.hcount = c->d_hcount;
    B.hcap = c->hist_cap;
    B.nghost = (long long)g.nx * g.ny - (long long)(g.nx - 2 * NG) * (g.ny - 2 * NG);
    B.cfl_dr = c->s.cfl * ((g.dx >= g.dy) ? g.dy : g.dx);
    B.nsx = (g.nx - 5 + SEG - 1) / SEG;
    B.nsy = (g.ny - 5 + SEG - 1) / SEG;
    const long long ninter = (long long)(g.nx - 2 * NG) * (g.ny - 2 * NG);
    B.cflblocks = std::min(cdiv(ninter, 256), CFL_MAX_BLOCKS);
    const int cnt[BK_COUNT] = {cdiv((long long)B.nsx * g.ny * WAVE, (long long)WAVE * b->xwpb),
                               ((g.nx + YC - 1) / YC) * B.nsy,
                               cdiv((long long)(g.nx - 5) * ((g.ny - 5 + 1) / 2), b->ewb),
                               cdiv((long long)(g.nx - 7) * (g.ny - 7), b->ewb),
                               cdiv(B.nghost * 8, 256),
                               B.cflblocks};
    for (int k = 0; k < BK_COUNT; k++) off[(size_t)k * (n + 1) + i + 1] = off[(size_t)k * (n + 1) + i] + cnt[k];
  }
  for (int k = 0; k < BK_COUNT; k++) b->nblk[k] = off[(size_t)k * (n + 1) + n];
  QAS_BCHECK(b, hipMemcpyAsync(b->d_inst, b->inst.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice, b->stream));
  QAS_BCHECK(b, hipMemcpyAsync(b->d_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, b->stream));
  QAS_BCHECK(b, hipStreamSynchronize(b->stream));
  return QAS_OK;
}

static const int* boff(const qas_batch* b, int k) { return b->d_off + (size_t)k * (b->n + 1); }

static hipError_t launch_bound_all(qas_batch* b, int stage, const BatchInst* tbl = nullptr)
{
  hipLaunchKernelGGL(kb_bound, (unsigned)b->nblk[BK_BOUND], 256, 0, b->stream, tbl ? tbl : b->d_inst,
                     boff(b, BK_BOUND), b->n, stage);
  return hipGetLastError();
}

static hipError_t launch_cfl_all(qas_batch* b)
{
  hipLaunchKernelGGL(kb_cfl, (unsigned)b->nblk[BK_CFL], 256, 0, b->stream, b->d_inst, boff(b, BK_CFL), b->n,
                     b->s.gam);
  hipLaunchKernelGGL(kb_cfl_final, (unsigned)b->n, 256, 0, b->stream, b->d_inst, b->n);
  return hipGetLastError();
}

// the four kernels of RK stage rk for every instance, in the batch's shape,
// on the instance table `tbl` (the batch's own by default; faces_only: the
// two sweeps without corner EMF and CT)
static hipError_t launch_stage_all(qas_batch* b, int rk, const BatchInst* tbl = nullptr, bool faces_only = false)
{
  hipStream_t st = b->stream;
  const int n = b->n;
  const BatchInst* I = tbl ? tbl : b->d_inst;
  const unsigned eb = (unsigned)b->ewb, xb = (unsigned)(WAVE * b->xwpb);
  hipLaunchKernelGGL(b->sy, (unsigned)b->nblk[BK_SY], 256, 0, st, StagePtrs{}, Geo{}, BcFlags{}, b->s.gam, 0, I,
                     boff(b, BK_SY), n, rk);
  hipLaunchKernelGGL(b->sx, (unsigned)b->nblk[BK_SX], xb, 0, st, StagePtrs{}, Geo{}, BcFlags{}, b->s.gam, 0, I,
                     boff(b, BK_SX), n, rk);
  if (faces_only) return hipGetLastError();
  hipLaunchKernelGGL(b->emf, (unsigned)b->nblk[BK_EMF], eb, 0, st, I, boff(b, BK_EMF), n, rk, b->s.ctw);
  hipLaunchKernelGGL(b->ct, (unsigned)b->nblk[BK_CT], eb, 0, st, I, boff(b, BK_CT), n, rk);
  return hipGetLastError();
}

// one step of exec_'s loop body for every instance
static hipError_t batch_step(qas_batch* b, int flg, bool hist, bool even)
{
  hipStream_t st = b->stream;
  const int n = b->n;
  if (hist) hipLaunchKernelGGL(kb_record, (unsigned)cdiv(n, 64), 64, 0, st, b->d_inst, n);
  for (int rk = 0; rk < b->s.r_k; rk++) {
    hipError_t e = launch_stage_all(b, rk);
    if (e == hipSuccess) e = launch_bound_all(b, rk);
    if (e != hipSuccess) return e;
  }
  if (flg && even) return launch_cfl_all(b);
  return hipGetLastError();
}

template <int ODR>
static void pick_stage(qas_batch* b, bool f32)
{
  b->emf = f32 ? (emf_fn)kb_emf2d<ODR, float> : (emf_fn)kb_emf2d<ODR, double>;
  b->ct = kb_ct2d<ODR>;
}

extern "C" int qas_batch_create(qas_batch** out, qas_ctx* const* ctxs, int n)
{
  if (!out) return QAS_E_ARG;
  *out = nullptr;
  qas_batch* b = new qas_batch();
  *out = b;
  if (!ctxs || n <= 0) return bfail(b, QAS_E_ARG, "qas_batch_create: no contexts");
  b->n = n;
  b->c.assign(ctxs, ctxs + n);
  for (int i = 0; i < n; i++)
    if (!b->c[i] || !b->c[i]->slab) return bfail(b, QAS_E_ARG, "qas_batch_create: invalid context");
  b->s = b->c[0]->s;
  int dev0 = b->c[0]->device;
  for (int i = 0; i < n; i++) {
    const qas_ctx* c = b->c[i];
    const qas_scheme& s = c->s;
    if (c->g.ndim != 2) return bfail(b, QAS_E_ARG, "qas_batch_create: batches hold 2D domains");
    if (c->device != dev0) return bfail(b, QAS_E_ARG, "qas_batch_create: all domains on one device");
    if (c->d_phi || c->d_nu)
      return bfail(b, QAS_E_ARG, "qas_batch_create: batches run the ideal step (no gravity, no dissipation)");
    if (s.rmn != b->s.rmn || s.odr != b->s.odr || s.r_k != b->s.r_k || s.ctw != b->s.ctw ||
        s.precision != b->s.precision || s.gam != b->s.gam)
      return bfail(b, QAS_E_ARG, "qas_batch_create: every domain needs the same rmn, odr, r_k, ctw, precision, gam");
    for (int j = 0; j < i; j++)
      if (b->c[j] == c) return bfail(b, QAS_E_ARG, "qas_batch_create: a context appears twice");
  }
  if (b->s.r_k < 2) return bfail(b, QAS_E_ARG, "qas_batch_create: r_k must be 2 or 3");
  const bool f32 = b->s.precision == 1;
  const bool fma = b->s.precision == 2;
  b->sx = f32 ? pick_xsweep2d_f32(b->s.rmn, b->s.odr, true)
              : (fma ? pick_xsweep2d_fma(b->s.rmn, b->s.odr, true) : pick_xsweep2d(b->s.rmn, b->s.odr, true));
  b->sy = f32 ? pick_ysweep2d_f32(b->s.rmn, b->s.odr, true)
              : (fma ? pick_ysweep2d_fma(b->s.rmn, b->s.odr, true) : pick_ysweep2d(b->s.rmn, b->s.odr, true));
  if (!b->sx || !b->sy) return bfail(b, QAS_E_ARG, "qas_batch_create: no kernel for this scheme");
  switch (b->s.odr) {
    case 1: pick_stage<1>(b, f32); break;
    case 2: pick_stage<2>(b, f32); break;
    case 3: pick_stage<3>(b, f32); break;
    default: pick_stage<4>(b, f32); break;
  }
  QAS_BCHECK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  QAS_BCHECK(b, hipMalloc(&b->d_inst, sizeof(BatchInst) * n));
  QAS_BCHECK(b, hipMalloc(&b->d_off, sizeof(int) * (size_t)BK_COUNT * (n + 1)));
  QAS_BCHECK(b, hipMalloc(&b->d_dts, sizeof(double) * n));
  return QAS_OK;
}

extern "C" void qas_batch_destroy(qas_batch* b)
{
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_off) (void)hipFree(b->d_off);
  if (b->d_dts) (void)hipFree(b->d_dts);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" const char* qas_batch_last_error(const qas_batch* b) { return b ? b->err.c_str() : "null batch"; }

// Capture `pairs` consecutive step pairs (odd, even) as one graph: a run of
// up to 2 * QAS_BATCH_GRAPH_PAIRS steps is one graph launch.  The kernels
// take dt and the dt-history slot from device memory, so one graph serves
// every run of the same length.
constexpr int QAS_BATCH_GRAPH_PAIRS = 128;
static int build_batch_graph(qas_batch* b, int flg, bool hist, int pairs)
{
  // wait for replays of the old graph before its executable is destroyed
  if (b->gexec || b->graph) QAS_BCHECK(b, hipStreamSynchronize(b->stream));
  if (b->gexec) { QAS_BCHECK(b, hipGraphExecDestroy(b->gexec)); b->gexec = nullptr; }
  if (b->graph) { QAS_BCHECK(b, hipGraphDestroy(b->graph)); b->graph = nullptr; }
  QAS_BCHECK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e1 = hipSuccess;
  for (int q = 0; q < pairs && e1 == hipSuccess; q++) {
    e1 = batch_step(b, flg, hist, false);
    if (e1 == hipSuccess) e1 = batch_step(b, flg, hist, true);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e1 != hipSuccess || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, QAS_E_HIP, std::string("batch graph capture: ") + hipGetErrorName(e1 != hipSuccess ? e1 : e));
  }
  b->graph = g;
  QAS_BCHECK(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->graph_flg = flg;
  b->graph_hist = hist ? 1 : 0;
  b->graph_pairs = pairs;
  return QAS_OK;
}

extern "C" int qas_batch_run(qas_batch* b, int nsteps, int flg, double* dt_inout, double* dt_hist)
{
  if (!b || !b->stream) return bfail(b, QAS_E_ARG, "qas_batch_run: invalid batch");
  if (nsteps < 0 || !dt_inout) return bfail(b, QAS_E_ARG, "qas_batch_run: bad arguments");
  const int n = b->n;
  const long long n0 = b->c[0]->nstep;
  for (int i = 0; i < n; i++) {
    qas_ctx* c = b->c[i];
    if (!c->uploaded) return bfail(b, QAS_E_STATE, "qas_batch_run: a domain has no state uploaded");
    if (c->nstep != n0) return bfail(b, QAS_E_STATE, "qas_batch_run: the domains' step counters differ");
    QAS_BCHECK(b, hipStreamSynchronize(c->stream));
  }
  const bool hist = dt_hist != nullptr;
  if (hist) {
    for (int i = 0; i < n; i++) {
      qas_ctx* c = b->c[i];
      if (nsteps > c->hist_cap) {
        if (c->d_hist) QAS_BCHECK(b, hipFree(c->d_hist));
        c->d_hist = nullptr;
        QAS_BCHECK(b, hipMalloc(&c->d_hist, sizeof(double) * (size_t)nsteps));
        c->hist_cap = nsteps;
        // the context's own step-pair graph (qas_run) holds the old history
        // pointer and capacity in its record node: rebuild it on next use
        // (the batch graph reads the pointers from the device table, which
        // fill_table below rewrites)
        c->graph_hist = -1;
      }
    }
  }
  if (!b->tuned) {
    const int rt = qas_batch_tune(b);
    if (rt) return rt;
  }
  int r = fill_table(b);
  if (r) return r;
  if (hist) hipLaunchKernelGGL(kb_zero_counts, (unsigned)cdiv(n, 64), 64, 0, b->stream, b->d_inst, n);
  // dt: instances with dt_inout[i] > 0 take it, the others the CFL value
  bool any_cfl = false;
  for (int i = 0; i < n; i++) any_cfl = any_cfl || !(dt_inout[i] > 0.0);
  if (any_cfl) QAS_BCHECK(b, launch_cfl_all(b));
  // the device table is in the batch's block order (perm)
  std::vector<double> dts(n);
  for (int k = 0; k < n; k++) dts[k] = dt_inout[b->perm[k]];
  QAS_BCHECK(b, hipMemcpyAsync(b->d_dts, dts.data(), sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(kb_set_dt, (unsigned)cdiv(n, 64), 64, 0, b->stream, b->d_inst, n, b->d_dts);
  bool bounded = true;
  for (int i = 0; i < n; i++) bounded = bounded && b->c[i]->bounded;
  if (!bounded) QAS_BCHECK(b, launch_bound_all(b, -1));
  long long ns = n0;
  int k = 0;
  if (k < nsteps && (ns % 2) == 1) {
    ns++;
    QAS_BCHECK(b, batch_step(b, flg, hist, true));
    k++;
  }
  if (b->use_graph && nsteps - k >= 2) {
    // the remaining step pairs as one graph launch (runs longer than the
    // cap: the cap's graph launched repeatedly, the rest step by step)
    const int pairs = (nsteps - k) / 2;
    const int gp = pairs < QAS_BATCH_GRAPH_PAIRS ? pairs : QAS_BATCH_GRAPH_PAIRS;
    if (!b->gexec || b->graph_flg != flg || b->graph_hist != (hist ? 1 : 0) || b->graph_pairs != gp) {
      r = build_batch_graph(b, flg, hist, gp);
      if (r) return r;
    }
    for (; k + 2 * gp <= nsteps; k += 2 * gp) {
      QAS_BCHECK(b, hipGraphLaunch(b->gexec, b->stream));
      ns += 2 * gp;
    }
  }
  for (; k < nsteps; k++) {
    ns++;
    QAS_BCHECK(b, batch_step(b, flg, hist, ns % 2 == 0));
  }
  hipLaunchKernelGGL(kb_get_dt, (unsigned)cdiv(n, 64), 64, 0, b->stream, b->d_inst, n, b->d_dts);
  QAS_BCHECK(b, hipMemcpyAsync(dts.data(), b->d_dts, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  QAS_BCHECK(b, hipStreamSynchronize(b->stream));
  for (int k = 0; k < n; k++) dt_inout[b->perm[k]] = dts[k];
  for (int i = 0; i < n; i++) {
    qas_ctx* c = b->c[i];
    c->nstep = ns;
    c->bounded = true;
    if (hist && nsteps > 0)
      QAS_BCHECK(b, hipMemcpy(dt_hist + (size_t)i * nsteps, c->d_hist, sizeof(double) * (size_t)nsteps,
                              hipMemcpyDeviceToHost));
  }
  return QAS_OK;
}

// ---------------------------------------------------------------- autotuner
// Launch shapes of the batched stage: block order of the instances, x-sweep
// wavefronts per workgroup, element-wise (EMF, CT) workgroup size.  Every
// shape gives the same arithmetic per cell (bit-identical results); they
// differ in how the workgroups of uneven domains fill the GPU.  The tuner
// times RK stage 1 (both sweeps, EMF, CT, ghost fill of its output) in each
// shape, three times each after one warm-up, and keeps the fastest.  Stage
// 1 reads the current states and writes only the next-stage buffers and
// scratch, which the run overwrites.
extern "C" int qas_batch_tune(qas_batch* b)
{
  if (!b || !b->stream) return bfail(b, QAS_E_ARG, "qas_batch_tune: invalid batch");
  const int cand[8][3] = {{0, 4, 256}, {1, 4, 256}, {0, 2, 256}, {1, 2, 256},
                          {0, 4, 128}, {1, 4, 128}, {0, 2, 128}, {1, 2, 128}};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  QAS_BCHECK(b, hipEventCreate(&e0));
  QAS_BCHECK(b, hipEventCreate(&e1));
  float best = -1.0f;
  int bi = 0;
  int r = QAS_OK;
  for (int k = 0; k < 8 && !r; k++) {
    b->order = cand[k][0];
    b->xwpb = cand[k][1];
    b->ewb = cand[k][2];
    r = fill_table(b);
    if (r) break;
    if (launch_stage_all(b, 0) != hipSuccess || launch_bound_all(b, 0) != hipSuccess) {
      r = bfail(b, QAS_E_HIP, "qas_batch_tune: launch");
      break;
    }
    float tmin = -1.0f;
    for (int rep = 0; rep < 3; rep++) {
      hipError_t e = hipEventRecord(e0, b->stream);
      if (e == hipSuccess) e = launch_stage_all(b, 0);
      if (e == hipSuccess) e = launch_bound_all(b, 0);
      if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) { r = bfail(b, QAS_E_HIP, std::string("qas_batch_tune: ") + hipGetErrorName(e)); break; }
      if (tmin < 0.0f || ms < tmin) tmin = ms;
    }
    if (!r && (best < 0.0f || tmin < best)) {
      best = tmin;
      bi = k;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (r) return r;
  b->order = cand[bi][0];
  b->xwpb = cand[bi][1];
  b->ewb = cand[bi][2];
  b->tuned = true;
  b->graph_flg = -1;  // the captured graph has the old grid sizes
  return fill_table(b);
}

extern "C" int qas_batch_use_graphs(qas_batch* b, int on)
{
  if (!b) return QAS_E_ARG;
  b->use_graph = on != 0;
  return QAS_OK;
}

extern "C" int qas_batch_set_shape(qas_batch* b, int order, int xwpb, int ewb)
{
  if (!b) return QAS_E_ARG;
  if ((order != 0 && order != 1) || (xwpb != 1 && xwpb != 2 && xwpb != 4) || (ewb != 64 && ewb != 128 && ewb != 256))
    return bfail(b, QAS_E_ARG, "qas_batch_set_shape: order 0/1, xwpb 1/2/4, ewb 64/128/256");
  b->order = order;
  b->xwpb = xwpb;
  b->ewb = ewb;
  b->tuned = true;
  b->graph_flg = -1;
  return QAS_OK;
}

extern "C" int qas_batch_shape(const qas_batch* b, int* order, int* xwpb, int* ewb)
{
  if (!b || !order || !xwpb || !ewb) return QAS_E_ARG;
  *order = b->order;
  *xwpb = b->xwpb;
  *ewb = b->ewb;
  return b->tuned ? 1 : 0;
}

// ---------------------------------------------------------------- batched adjoint
// qas_adj_forward / qas_adj_backward for every domain of the batch, with one
// launch per kernel for all of them (the adjoint kernels of
// qas_adjoint2d.hip in their batched form, the forward stage kernels of the
// batch for the recorded forward and the recompute).  Per domain the
// operations are those of the single-domain entry points.
static int adj_step_table(qas_batch* b, AdjStepInst** d_tbl, int** d_off, int* nblocks)
{
  const int n = b->n;
  std::vector<AdjStepInst> t(n);
  std::vector<int> off(n + 1, 0);
  for (int k = 0; k < n; k++) {
    qas_ctx* c = b->c[b->perm[k]];
    AdjBufs a;
    adj_bufs(c, &a);
    AdjStepInst& s = t[k];
    s.ck = a.ck;
    s.stride = (long long)a.stride;
    s.nd = c->geo.nd;
    s.dts = a.d_dts;
    s.dslot = a.dslot;
    s.dt = c->d_dt;
    // the backward's step input buffer: a state buffer the context does not
    // hold its state in (its layout has the workspace stride)
    const int ubi = (c->cur + 1) % 3;
    for (int m = 0; m < 8; m++) {
      s.cur[m] = c->U[c->cur][m];
      s.ub[m] = c->U[ubi][m];
    }
    off[k + 1] = off[k] + cdiv(8 * c->geo.nd, 256);
  }
  *nblocks = off[n];
  QAS_BCHECK(b, hipMalloc(d_tbl, sizeof(AdjStepInst) * n));
  QAS_BCHECK(b, hipMalloc(d_off, sizeof(int) * (n + 1)));
  QAS_BCHECK(b, hipMemcpyAsync(*d_tbl, t.data(), sizeof(AdjStepInst) * n, hipMemcpyHostToDevice, b->stream));
  QAS_BCHECK(b, hipMemcpyAsync(*d_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream));
  QAS_BCHECK(b, hipStreamSynchronize(b->stream));
  return QAS_OK;
}

static int adj_batch_scope(qas_batch* b, const char* who)
{
  if (!b || !b->stream) return bfail(b, QAS_E_ARG, std::string(who) + ": invalid batch");
  for (int i = 0; i < b->n; i++) {
    const std::string why = adj_scope(b->c[i]);
    if (!why.empty()) return bfail(b, QAS_E_ARG, std::string(who) + ": " + why);
  }
  return QAS_OK;
}

extern "C" int qas_batch_adj_forward(qas_batch* b, int nsteps, const double* dts)
{
  int r = adj_batch_scope(b, "qas_batch_adj_forward");
  if (r) return r;
  if (nsteps <= 0 || !dts) return bfail(b, QAS_E_ARG, "qas_batch_adj_forward: bad arguments");
  const int n = b->n;
  for (int i = 0; i < n; i++) {
    qas_ctx* c = b->c[i];
    if (!c->uploaded) return bfail(b, QAS_E_STATE, "qas_batch_adj_forward: a domain has no state uploaded");
    r = adj_prepare(c, nsteps, dts + (size_t)i * nsteps);
    if (r) return bfail(b, r, std::string("qas_batch_adj_forward: ") + c->err);
    QAS_BCHECK(b, hipStreamSynchronize(c->stream));
  }
  if (!b->tuned) {
    r = qas_batch_tune(b);
    if (r) return r;
  }
  r = fill_table(b);
  if (r) return r;
  AdjStepInst* d_tbl = nullptr;
  int* d_toff = nullptr;
  int nbk = 0;
  r = adj_step_table(b, &d_tbl, &d_toff, &nbk);
  bool bounded = true;
  for (int i = 0; i < n; i++) bounded = bounded && b->c[i]->bounded;
  hipError_t e = hipSuccess;
  if (!r && !bounded) e = launch_bound_all(b, -1);
  for (int k = 0; !r && e == hipSuccess && k < nsteps; k++) {
    hipLaunchKernelGGL(kb_adj_ckpt, (unsigned)nbk, 256, 0, b->stream, d_tbl, d_toff, n, k);
    for (int rk = 0; rk < 3 && e == hipSuccess; rk++) {
      e = launch_stage_all(b, rk);
      if (e == hipSuccess) e = launch_bound_all(b, rk);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (d_tbl) (void)hipFree(d_tbl);
  if (d_toff) (void)hipFree(d_toff);
  if (r) return r;
  if (e != hipSuccess) return bfail(b, QAS_E_HIP, std::string("qas_batch_adj_forward: ") + hipGetErrorName(e));
  for (int i = 0; i < n; i++) {
    adj_set_recorded(b->c[i], nsteps);
    b->c[i]->bounded = true;
  }
  return QAS_OK;
}

extern "C" int qas_batch_adj_backward(qas_batch* b, const double* const* g_final, double* const* g_init)
{
  int r = adj_batch_scope(b, "qas_batch_adj_backward");
  if (r) return r;
  if (!g_final || !g_init) return bfail(b, QAS_E_ARG, "qas_batch_adj_backward: bad arguments");
  const int n = b->n;
  const int nrec = adj_recorded(b->c[0]);
  for (int i = 0; i < n; i++) {
    if (!g_final[i] || !g_init[i]) return bfail(b, QAS_E_ARG, "qas_batch_adj_backward: null gradient array");
    if (adj_recorded(b->c[i]) == 0 || adj_recorded(b->c[i]) != nrec)
      return bfail(b, QAS_E_STATE, "qas_batch_adj_backward: every domain needs the same recorded forward");
  }
  r = fill_table(b);  // block order (perm) and shape of the batch
  if (r) return r;
  // the final-state adjoints into each context's staging
  for (int i = 0; i < n; i++) {
    qas_ctx* c = b->c[i];
    AdjBufs a;
    adj_bufs(c, &a);
    QAS_BCHECK(b, hipMemcpyAsync(a.stage, g_final[i], 8 * (size_t)c->geo.nd * sizeof(double),
                                 hipMemcpyHostToDevice, b->stream));
  }
  // recompute table: the batch's descriptors with the three stages reading
  // the step input buffer (ub) and writing the adjoint workspace
  std::vector<BatchInst> rec = b->inst;
  std::vector<qas_ctx*> ord(n);
  std::vector<double*> ub(8 * (size_t)n);
  const double ra[3] = {0.0, 0.75, 1.0 / 3.0}, rbc[3] = {1.0, 0.25, 2.0 / 3.0};
  for (int k = 0; k < n; k++) {
    qas_ctx* c = b->c[b->perm[k]];
    ord[k] = c;
    AdjBufs a;
    adj_bufs(c, &a);
    double* const* U = c->U[(c->cur + 1) % 3];
    for (int m = 0; m < 8; m++) ub[8 * (size_t)k + m] = U[m];
    for (int rk = 0; rk < 3; rk++) {
      StagePtrs& p = rec[k].st[rk];
      std::memset(&p, 0, sizeof(p));
      for (int m = 0; m < 8; m++) {
        p.in[m] = rk == 0 ? U[m] : a.S[rk - 1][m];
        p.out[m] = a.S[rk][m];
        p.u0[m] = U[m];
      }
      p.vx = a.prim[rk][0]; p.vy = a.prim[rk][1]; p.vz = a.prim[rk][2]; p.pr = a.prim[rk][3];
      p.cx = a.prim[rk][4]; p.cy = a.prim[rk][5];
      for (int q = 0; q < 6; q++) p.acc[q] = c->acc[q];
      for (int q = 0; q < 8; q++) p.emf[q] = a.emf[rk][q];
      p.ez = a.ez[rk];
      p.dt = a.dslot;
      p.a = ra[rk];
      p.b = rbc[rk];
    }
  }
  BatchInst* d_rec = nullptr;
  AdjStepInst* d_tbl = nullptr;
  int* d_toff = nullptr;
  int nbk = 0;
  AdjBatchDev* ad = nullptr;
  std::string aerr;
  hipError_t e = hipMalloc(&d_rec, sizeof(BatchInst) * n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_rec, rec.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) r = adj_step_table(b, &d_tbl, &d_toff, &nbk);
  if (e == hipSuccess && !r) r = adjb_create(&ad, ord.data(), ub.data(), n, b->stream, &aerr);
  if (e == hipSuccess && !r) e = adjb_fold_final(ad, b->stream);
  for (int step = nrec - 1; e == hipSuccess && !r && step >= 0; step--) {
    hipLaunchKernelGGL(kb_adj_load, (unsigned)nbk, 256, 0, b->stream, d_tbl, d_toff, n, step);
    // forward recompute of the step: stage inputs and intermediates
    for (int rk = 0; rk < 3 && e == hipSuccess; rk++) {
      e = launch_stage_all(b, rk, d_rec, rk == 2);
      if (e == hipSuccess && rk < 2) e = launch_bound_all(b, rk, d_rec);
    }
    for (int rk = 2; rk >= 0 && e == hipSuccess; rk--) e = adjb_stage(ad, rk, rk == 0, b->stream);
  }
  // the initial-state adjoints back to the host
  for (int i = 0; e == hipSuccess && !r && i < n; i++) {
    qas_ctx* c = b->c[i];
    AdjBufs a;
    adj_bufs(c, &a);
    for (int m = 0; m < 8 && e == hipSuccess; m++)
      e = hipMemcpyAsync(g_init[i] + (size_t)m * c->geo.nd, a.gO[m], (size_t)c->geo.nd * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  adjb_destroy(ad);
  if (d_rec) (void)hipFree(d_rec);
  if (d_tbl) (void)hipFree(d_tbl);
  if (d_toff) (void)hipFree(d_toff);
  if (r) return aerr.empty() ? r : bfail(b, r, "qas_batch_adj_backward: " + aerr);
  if (e != hipSuccess) return bfail(b, QAS_E_HIP, std::string("qas_batch_adj_backward: ") + hipGetErrorName(e));
  return QAS_OK;
}
