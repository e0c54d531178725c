// qas_adjoint2d.hip -- hand-written discrete adjoint of the 2D step.
//
// Scope: FP64, HLLD, LHLLD or MLAU (RMN 1-3), ODR 4 (WCNS-4 / 4th-order linear /
// fcen / difference), SSP-RK3, CUCT on or off, periodic on both axes, dt frozen
// (the step sizes are inputs, as a recorded dt history is).  The adjoint is
// that of exactly the forward kernels of qas_kernels2d.hip / qas_sweep2d.hpp,
// which reproduce MHD2D::ideal (2D/SERIAL/OTvortex/mhd2d_solve.cpp:3-400)
// bit for bit; non-smooth operations are differentiated at the branch the
// forward takes (qas_adjoint.hpp).  This unit is built with FMA contraction
// inside each source expression and with FP64 division as a refined
// reciprocal (no operand scaling or special-case fix-up): the forward
// quantities it recomputes (face states, branch predicates) round
// differently from the forward kernels at the last bit, so a branch can
// differ only within one rounding of its threshold.
//
// Checkpointing: qas_adj_forward keeps every step's input state on the
// device.  qas_adj_backward walks the steps in reverse; for each step it
// recomputes the three forward stages from the checkpoint, keeping each
// stage's primitives and face EMF parts, and then runs the stage adjoints
// 3, 2, 1.  With qas_adj_keep_stages the forward also writes every stage's
// input state, primitives and face parts into a per-step store (the same
// kernels on other buffers) and the backward reads them instead of
// recomputing.  A stage adjoint is these kernels, all in gather form (every
// output element is written by one thread; no global atomics):
//   corner    adjoint of the CT update -> g_ez at corners; forward CUCT
//             weight and the adjoint of its inputs (once per corner)
//   faceparts adjoint of the corner EMF interpolation -> g of the four face
//             parts (x/y sweep, upwind/central)
//   sweep     per direction one pass over line segments (one wavefront
//             each): it reconstructs the face states, runs the flux adjoint
//             (HLLD) per face, hands the face-state adjoints to the lanes of
//             the two cells by cross-lane shuffles and runs the adjoint of
//             the reconstruction per cell, gathering the stencil adjoints
//             onto cells through a per-wavefront LDS row (58 owned cells per
//             wavefront) -> g of ro, vx, vy, vz, cx, cy, bz, pr and face B.
//             The x pass sets these arrays; the y pass runs on transposed
//             copies of its inputs and its outputs are added back (two
//             transpose kernels)
//   prim      adjoint of the primitive recovery + CUCT corner terms of
//             ro, vx, vy + the RK pass-through b * g_out -> g_in of
//             ro, mx, my, mz, bz, en; a * g_out into the step-base adjoint
//   face      adjoint of fcen (with the periodic ghost copies of cx, cy
//             folded onto their sources) + CUCT and sweep terms -> g_in of
//             bx, by
//   fold      adjoint of bound(): ghost entries added onto the cells they
//             copy, ghosts zeroed
#include <cstring>
#include <string>
#include <vector>

#include "qas_adjbatch.hpp"
#include "qas_adjoint.hpp"
#include "qas_batch.hpp"
#include "qas_device.hpp"
#include "qas_internal.hpp"

namespace qas {

constexpr int SEGA = 58;  // owned cells per wavefront segment of the adjoint sweeps (lanes 3..60)

struct AdjPtrs {
  const double* in[8];   // stage input state (ghosts filled)
  const double *vx, *vy, *vz, *pr, *cx, *cy;  // forward primitives of `in`
  const double* emf[8];  // forward face parts ([0] x upwind, [1] x central, [4] y upwind, [5] y central)
  double* gO[8];         // adjoint of the stage output (ghost entries zero); k_adj_foldstage
                         // overwrites it with the previous stage's
  double* gI[8];         // adjoint of the stage input (before the ghost fold)
  double* gU0[8];        // adjoint of the step base (accumulated over the stages)
  double* gp[8];         // adjoint of ro, vx, vy, vz, cx, cy, bz, pr as the sweeps read them
  double* gbf[2];        // adjoint of face bx (x sweep), by (y sweep)
  double* gemf[4];       // adjoint of the face parts: x upwind, x central, y upwind, y central
  double* gdv[2];        // MLAU: adjoint of the shock sensor's transverse part dv[1] per x / y face
  double *gez, *ctv;     // corner adjoint of ez; forward CUCT weight
  double* gcc[5];        // per corner, the CUCT weight's input adjoints: ro, vx, vy (each of the
                         // four cells), bx, by (each of the two faces)
  const double* dt;
  double a, b;
  BcFlags bc;            // boundary flags: the adjoint of bound() in the folds
};

// ---------------------------------------------------------------- helpers
// periodic images of a kept index q in [4, n-5]: q itself and the ghost that
// bound() fills from it (common_func.c:67-74)
struct Img {
  int i[2];
  int n;
};
__device__ __forceinline__ Img images(int q, int n)
{
  Img r;
  r.i[0] = q;
  r.n = 1;
  const int p = n - 2 * NG;
  if (q + p < n) r.i[r.n++] = q + p;
  else if (q - p >= 0) r.i[r.n++] = q - p;
  return r;
}

// f at (i, j) plus f at its ghost images.  The four loads are issued
// unconditionally (a missing image re-reads the kept cell) and masked by
// selects: with per-image branches, successive calls could not overlap
// their memory latencies.
template <typename F>
__device__ __forceinline__ double fold_with(const F& f, int i, int j, const Geo& g)
{
  const Img ix = images(i, g.nx), iy = images(j, g.ny);
  const long long nx = g.nx;
  const int i1 = ix.n > 1 ? ix.i[1] : ix.i[0];
  const int j1 = iy.n > 1 ? iy.i[1] : iy.i[0];
  const double a = f(nx * iy.i[0] + ix.i[0]), b = f(nx * iy.i[0] + i1);
  const double c = f(nx * j1 + ix.i[0]), d = f(nx * j1 + i1);
  double v = a;
  v += ix.n > 1 ? b : 0.0;
  v += iy.n > 1 ? c : 0.0;
  v += (ix.n > 1 && iy.n > 1) ? d : 0.0;
  return v;
}
__device__ __forceinline__ double fold_at(const double* f, int i, int j, const Geo& g)
{
  return fold_with([f](long long q) { return f[q]; }, i, j, g);
}

__device__ __forceinline__ bool kept(int i, int n) { return i >= NG && i <= n - NG - 1; }

// The bound() map of any boundary kind (common_func.c:60-138): a ghost
// index holds (factor x) the value at its source index (resolve).  Its
// adjoint gathers, at a source index q, q itself and every ghost whose
// source is q, with that ghost's factor (an open boundary copies its edge
// cell into all NG ghosts).  Candidate c of an axis: c = 0 is q itself,
// c = 1 + 2 g + s the ghost g (s = 0) or n - 1 - g (s = 1); its index and
// factor are recomputed on the fly (a per-thread array indexed at run time
// would live in scratch memory).
__device__ __forceinline__ bool is_source(int i, int n, int st, int dn)
{
  const Src s = resolve(i, n, st, dn);
  return s.i == i && !s.mul;
}
__device__ __forceinline__ double candidate(int c, int q, int n, int st, int dn, int* idx)
{
  if (c == 0) {
    *idx = q;
    return 1.0;
  }
  const int g = (c - 1) >> 1;
  const int gi = ((c - 1) & 1) ? n - 1 - g : g;
  *idx = gi;
  const Src s = resolve(gi, n, st, dn);
  return ((s.i != gi || s.mul) && s.i == q) ? (s.mul ? s.k : 1.0) : 0.0;
}
// f at (i, j) plus (factor x) f at every ghost bound() fills from it, the
// x map of (stx, dnx) and the y map of (sty, dny) composed as bound()
// composes them (y ghosts of x ghosts); any boundary kinds.  No ghost maps
// to an index more than 2 NG + 1 away from both ends.
template <typename F>
__device__ __forceinline__ double fold_general(const F& f, int i, int j, const Geo& g, int stx, int dnx, int sty,
                                               int dny)
{
  const long long nx = g.nx;
  const int ncx = (i > 2 * NG && i < g.nx - 2 * NG - 1) ? 1 : 1 + 2 * NG;
  const int ncy = (j > 2 * NG && j < g.ny - 2 * NG - 1) ? 1 : 1 + 2 * NG;
  double v = 0.0;
#pragma unroll 1
  for (int b = 0; b < ncy; b++) {
    int jy;
    const double kb = candidate(b, j, g.ny, sty, dny, &jy);
    if (kb == 0.0) continue;
    double row = 0.0;
#pragma unroll 1
    for (int a = 0; a < ncx; a++) {
      int ix;
      const double ka = candidate(a, i, g.nx, stx, dnx, &ix);
      if (ka != 0.0) row += ka * f(nx * jy + ix);
    }
    v += kb * row;
  }
  return v;
}
// variable m's fold: the periodic images (fold_with) when m is periodic on
// both axes, else the general gather.  cell: a cell-centred quantity under
// m's boundary kinds (the cell-centred B of bx, by, whose map has no
// stagger: prim2d_cell resolves it with st = 0)
template <typename F>
__device__ __forceinline__ double fold_var(const F& f, int i, int j, const Geo& g, const BcFlags& bc, int m,
                                           bool cell = false)
{
  if (bc.dn[0][m] == 0 && bc.dn[1][m] == 0) return fold_with(f, i, j, g);
  return fold_general(f, i, j, g, cell ? 0 : bc.st[0][m], bc.dn[0][m], cell ? 0 : bc.st[1][m], bc.dn[1][m]);
}
__device__ __forceinline__ bool source_cell(int i, int j, const Geo& g, const BcFlags& bc, int m)
{
  return is_source(i, g.nx, bc.st[0][m], bc.dn[0][m]) && is_source(j, g.ny, bc.st[1][m], bc.dn[1][m]);
}

// inputs of the CUCT weight at corner ss (k_emf2d): ro, vx, vy of the cells
// (i-1,j-1), (i,j-1), (i-1,j), (i,j); bx(i,j-1), bx(i,j), by(i-1,j), by(i,j)
__device__ __forceinline__ void cuct_inputs(const AdjPtrs& p, long long ss, long long nx, double* x)
{
  const long long c[4] = {ss - 1 - nx, ss - nx, ss - 1, ss};
  for (int k = 0; k < 4; k++) {
    x[k] = p.in[0][c[k]];
    x[4 + k] = p.vx[c[k]];
    x[8 + k] = p.vy[c[k]];
  }
  x[12] = p.in[4][ss - nx];
  x[13] = p.in[4][ss];
  x[14] = p.in[5][ss - 1];
  x[15] = p.in[5][ss];
}

__device__ __forceinline__ bool corner_ok(int i, int j, const Geo& g)
{
  return i >= 3 && i <= g.nx - 3 && j >= 3 && j <= g.ny - 3;
}

// ---------------------------------------------------------------- corner
// k_ct2d: bx(i,j) uses dfc(ez(i,j-1..j+2)) scaled by -dt/dy; by(i,j) uses
// dfc(ez(i-1..i+2,j)) scaled by +dt/dx; out = a u0 + b (in + ...).  With
// CUCT the corner also runs the adjoint of its weight once: the weight's
// inputs enter through four-cell and two-face means, so one value per input
// kind (gcc) serves every cell and face of the corner.
template <int CTW>
__device__ __forceinline__ void corner_at(const AdjPtrs& p, const Geo& g, long long t)
{
  if (t >= g.nd) return;
  const int i = (int)(t % g.nx), j = (int)(t / g.nx);
  const long long nx = g.nx;
  double gez = 0.0, ctv = 0.5;
  double gc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (corner_ok(i, j, g)) {
    const double dtv = *p.dt;
    const double fy = -p.b * (dtv / g.dy), fx = p.b * (dtv / g.dx);
    const double w[4] = {1.0 / 24.0, -27.0 / 24.0, 27.0 / 24.0, -1.0 / 24.0};  // offsets -1..2
    for (int kk = 0; kk < 4; kk++) {
      const int k = kk - 1;
      const int jf = j - k, iff = i - k;  // in [1, n-2] at a corner in [3, n-3]
      gez += w[kk] * fy * p.gO[4][nx * jf + i];
      gez += w[kk] * fx * p.gO[5][nx * j + iff];
    }
    if (CTW) {
      double x[16];
      cuct_inputs(p, t, nx, x);
      ctv = cuct_weight(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12], x[13],
                        x[14], x[15]);
      double a[5], bq[5], l0, r0, l1, r1, dum;
      for (int m = 0; m < 5; m++) {
        a[m] = p.emf[1][t + (m - 3) * nx];
        bq[m] = p.emf[1][t + (m - 2) * nx];
      }
      lin5<4>(a, l0, dum);
      lin5<4>(bq, dum, r0);
      for (int m = 0; m < 5; m++) {
        a[m] = p.emf[5][t + (m - 3)];
        bq[m] = p.emf[5][t + (m - 2)];
      }
      lin5<4>(a, l1, dum);
      lin5<4>(bq, dum, r1);
      const double gct = gez * (0.5 * (l0 + r0) + 0.5 * (l1 + r1));
      if (gct != 0.0) {
        double gx[16] = {};
        cuct_weight_adj(x, gct, gx);
        gc[0] = gx[0];   // = gx[1..3]
        gc[1] = gx[4];   // = gx[5..7]
        gc[2] = gx[8];   // = gx[9..11]
        gc[3] = gx[12];  // = gx[13]
        gc[4] = gx[14];  // = gx[15]
      }
    }
  }
  p.gez[t] = gez;
  p.ctv[t] = ctv;
  if (CTW)
#pragma unroll
    for (int k = 0; k < 5; k++) p.gcc[k][t] = gc[k];
}

// ---------------------------------------------------------------- face parts
// ez(c) = -1/2 [(uL+uR) + (1-w)(cL+cR)]_x + 1/2 [(uL+uR) + w (cL+cR)]_y with
// the one-sided linear interpolations of k_emf2d
__device__ __forceinline__ void faceparts_at(const AdjPtrs& p, const Geo& g, long long t)
{
  if (t >= g.nd) return;
  const int i = (int)(t % g.nx), j = (int)(t / g.nx);
  const long long nx = g.nx;
  const double cl[5] = {0.003125 * 9.0, 0.003125 * -56.0, 0.003125 * 234.0, 0.003125 * 144.0, 0.003125 * -11.0};
  const double cr[5] = {0.003125 * -11.0, 0.003125 * 144.0, 0.003125 * 234.0, 0.003125 * -56.0, 0.003125 * 9.0};
  double xu = 0.0, xc = 0.0, yu = 0.0, yc = 0.0;
#pragma unroll
  for (int s = -3; s <= 2; s++) {
    const double coef = (s <= 1 ? cl[s + 3] : 0.0) + (s >= -2 ? cr[s + 2] : 0.0);
    const int jc = j - s, ic = i - s;
    // loads unconditional (an absent corner re-reads cell t), terms selected
    const bool okx = corner_ok(i, jc, g), oky = corner_ok(ic, j, g);
    const long long cx = okx ? nx * jc + i : t, cy = oky ? nx * j + ic : t;
    const double gex = p.gez[cx], ctx = p.ctv[cx], gey = p.gez[cy], cty = p.ctv[cy];
    xu += okx ? -0.5 * coef * gex : 0.0;
    xc += okx ? -0.5 * (1.0 - ctx) * coef * gex : 0.0;
    yu += oky ? 0.5 * coef * gey : 0.0;
    yc += oky ? 0.5 * cty * coef * gey : 0.0;
  }
  p.gemf[0][t] = xu;
  p.gemf[1][t] = xc;
  p.gemf[2][t] = yu;
  p.gemf[3][t] = yc;
}

__global__ __launch_bounds__(256) void k_adj_faceparts(AdjPtrs p, Geo g)
{
  faceparts_at(p, g, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- sweeps
// Addressing of the line passes.  x lines are rows (stride 1); y lines are
// columns, read either in place (stride nx, 64 cache lines per wave-wide
// load) or from transposed copies (TR: element (i, j) at i * ny + j, so a
// column is contiguous and the y pass loads like the x pass).
template <int DIR, bool TR>
__device__ __forceinline__ long long line_stride(const Geo& g)
{
  return (DIR == 0 || TR) ? 1 : g.sy;
}
template <int DIR, bool TR>
__device__ __forceinline__ long long line_base(const Geo& g, int line)
{
  return DIR == 0 ? g.sy * line : (TR ? (long long)g.ny * line : (long long)line);
}

// The shock sensor's per-cell part (dv_at of the sweeps) with periodic
// ghost resolution: cell (a, b) of an na x nb plane resolved to its kept
// source, then min(v[Q+so] - v[Q], v[Q] - v[Q-so]) at Q = sa ra + sb rb.
// The same subtractions on the same values as dv_at: bit-identical.
__device__ __forceinline__ int per_res(int i, int n)
{
  return i < NG ? n - 2 * NG + i : (i >= n - NG ? i - (n - 2 * NG) : i);
}
__device__ __forceinline__ double dv_per(const double* v, int a, int b, int na, int nb, long long sa, long long sb,
                                         long long so)
{
  const long long Q = sa * per_res(a, na) + sb * per_res(b, nb);
  return tmin(v[Q + so] - v[Q], v[Q] - v[Q - so]);
}

// Forward of k_sweep2d (ODR 4) up to the face states, for the cell of this
// lane.
template <int DIR>
struct SweepFwd {
  const double *vn, *vt, *vu, *bt, *bu, *cn, *bf;
  double s[7][5];
  double pp[5], qq[5];
  double bnc, bnm, bnp;
  __device__ __forceinline__ void load(const AdjPtrs& p, long long sc, long long o)
  {
    vn = DIR == 0 ? p.vx : p.vy;
    vt = DIR == 0 ? p.vy : p.vz;
    vu = DIR == 0 ? p.vz : p.vx;
    bt = DIR == 0 ? p.cy : p.in[6];
    bu = DIR == 0 ? p.in[6] : p.cx;
    cn = DIR == 0 ? p.cx : p.cy;
    bf = DIR == 0 ? p.in[4] : p.in[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const long long a = sc + (m - 2) * o;
      s[0][m] = p.in[0][a];
      s[1][m] = vn[a];
      s[2][m] = vt[a];
      s[3][m] = vu[a];
      s[4][m] = bt[a];
      s[5][m] = bu[a];
      s[6][m] = p.pr[a];
    }
    bnc = cn[sc];
    // lr_fb inputs: p = b_pair * v_n, q = v_pair (x: cy, vy; y: cx, vx)
#pragma unroll
    for (int m = 0; m < 5; m++) {
      pp[m] = (DIR == 0 ? s[4][m] : s[5][m]) * s[1][m];
      qq[m] = DIR == 0 ? s[2][m] : s[3][m];
    }
    bnm = bf[sc];
    bnp = bf[sc + o];
  }
};

// Line pass (one per direction): one wavefront per segment of a line, lane
// l on cell q = seg*SEGA - 3 + l.
//   1. forward reconstruction of cell q (face states, positivity-fallback
//      flags), as k_sweep2d computes it;
//   2. lanes l >= 1: flux adjoint (HLLD) of face q (between cells q-1 and
//      q; the left state comes from lane l-1) from the divergence and
//      EMF-part adjoints -> adjoints of the two face states and of bn;
//   3. reconstruction adjoint of cell q from faces q (right state, this
//      lane) and q+1 (left state, lane l+1).  Its stencil adjoints are
//      added into a per-wavefront LDS row indexed by cell (lane l's stencil
//      point m is cell slot l + m), which is the gather onto cells; the
//      stencil is read from memory where it is used (loads re-issued after
//      each fence rather than held in registers across the reverse pass).
// Lanes 1..62 have both faces; lanes 3..60 own their cell and face (all
// five of their stencil contributors are complete).  Lanes 0 and 63 only
// add into slots of cells no lane owns.
template <int DIR>
struct StencilLd {
  const double* a[7];  // ro vn vt vu bt bu pr
  long long sc, o;
  __device__ __forceinline__ double operator()(int j, int m) const { return a[j][sc + (m - 2) * o]; }
  // a compiler barrier: loads after it are issued again instead of reusing
  // registers that would otherwise stay live across the whole reverse pass
  __device__ __forceinline__ void fence() const { asm volatile("" ::: "memory"); }
};
struct SlotAcc {
  double* row;  // this wavefront's [7][WAVE + 4] accumulators
  int lane;
  __device__ __forceinline__ void operator()(int j, int m, double v) const
  {
    atomicAdd(&row[j * (WAVE + 4) + lane + m], v);  // one add per lane and (j, m): no two lanes share a slot
  }
};

// blk: this workgroup's index within the pass
template <int RMN, int DIR, bool TR>
__device__ __forceinline__ void adj_sweep_body(const AdjPtrs& p, const Geo& g, double gam, int nseg, int blk)
{
  static_assert(RMN >= 1 && RMN <= 3, "HLLD, LHLLD or MLAU");
  static_assert(TR == (DIR == 1), "the y pass runs on transposed arrays, the x pass in place");
  __shared__ double gacc[256 / WAVE][7 * (WAVE + 4)];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = (int)(((long long)blk * blockDim.x + threadIdx.x) / WAVE);
  const int nline = DIR == 0 ? g.ny : g.nx;
  if (wave >= nseg * nline) return;  // whole wavefront exits together
  const int line = wave / nseg;
  const int seg = wave - line * nseg;
  const int n = DIR == 0 ? g.nx : g.ny;
  const long long o = line_stride<DIR, TR>(g);
  const long long base = line_base<DIR, TR>(g, line);
  const int q = seg * SEGA - 3 + lane;
  const int qc = q < 2 ? 2 : (q > n - 3 ? n - 3 : q);
  const long long fq = base + o * qc;
  double* row = gacc[threadIdx.x / WAVE];
#pragma unroll
  for (int j = 0; j < 7; j++) {
    row[j * (WAVE + 4) + lane] = 0.0;
    if (lane < 4) row[j * (WAVE + 4) + WAVE + lane] = 0.0;
  }
  // ---- 1, 2: forward reconstruction, flux adjoint of face q
  double gvl[7], gvr[7], gbn = 0.0, gql, gqr = 0.0;
  int fbc;
  {
    SweepFwd<DIR> f;
    f.load(p, fq, o);
    double vl[7], vr[7];
    lrstate<4, double, 0, false>(f.s, f.bnc, gam, vl, vr, &fbc);
    double L[7];
#pragma unroll
    for (int k = 0; k < 7; k++) L[k] = __shfl_up(vl[k], 1, WAVE);
    double gL[7] = {};
    double gqlf = 0.0, gdv1 = 0.0;
#pragma unroll
    for (int k = 0; k < 7; k++) gvr[k] = 0.0;
    if (lane >= 1 && q >= 3 && q <= n - 3) {
      const int ntl = DIR == 0 ? g.ny : g.nx;
      double gF[7] = {};
      if (line >= NG && line <= ntl - NG - 1) {
        const double sfac = -p.b * (*p.dt / (DIR == 0 ? g.dx : g.dy));
        // flux component fed into each updated variable (k_sweep2d fv[])
        const int fk[6] = {0, DIR == 0 ? 1 : 3, DIR == 0 ? 2 : 1, DIR == 0 ? 3 : 2, DIR == 0 ? 5 : 4, 6};
        const int vm[6] = {0, 1, 2, 3, 6, 7};
        const double w[4] = {1.0 / 24.0, -27.0 / 24.0, 27.0 / 24.0, -1.0 / 24.0};  // F(c+k), k = -1..2
#pragma unroll
        for (int v = 0; v < 6; v++) {
          double acc = 0.0;
#pragma unroll
          for (int kk = 0; kk < 4; kk++) {
            const int c = q - (kk - 1);
            if (c >= NG && c <= n - NG - 1) acc += w[kk] * p.gO[vm[v]][base + o * c];
          }
          gF[fk[v]] = sfac * acc;
        }
      }
      // face parts of the B flux: fbu = F_B - fc, fc = (ql(q-1) + qr(q)) / 2
      // (their adjoints are read in place also by the y pass: two loads)
      const long long fe = TR ? (long long)g.nx * qc + line : fq;
      const double gu = p.gemf[DIR * 2 + 0][fe], gc = p.gemf[DIR * 2 + 1][fe];
      gF[DIR == 0 ? 4 : 5] += gu;
      const double gfc = gc - gu;
      gqlf = 0.5 * gfc;
      gqr = 0.5 * gfc;
      const double bnf = 0.5 * (f.bf[fq] + f.bf[fq]);
      if constexpr (RMN >= 2) {
        // shock sensor of the face (k_sweep2d / k_sweep2d_y): dv0 along the
        // line, dv1 from the transverse differences of the two cells
        const double* vn = DIR == 0 ? p.vx : p.vy;
        const double dv0 = vn[fq] - vn[fq - o];
        double dv1;
        if (DIR == 0)  // dvy of cells (q-1, line), (q, line)
          dv1 = tmin(dv_per(p.vy, q - 1, line, g.nx, g.ny, 1, g.nx, g.nx),
                     dv_per(p.vy, q, line, g.nx, g.ny, 1, g.nx, g.nx));
        else  // dvx of cells (line, q-1), (line, q), on the transposed vx
          dv1 = tmin(dv_per(p.vx, line, q - 1, g.nx, g.ny, g.ny, 1, g.ny),
                     dv_per(p.vx, line, q, g.nx, g.ny, g.ny, 1, g.ny));
        double gdv0 = 0.0;
        if constexpr (RMN == 3) mlau_adj(L, vr, bnf, gam, dv0, dv1, gF, gL, gvr, gbn, gdv0, gdv1);
        else hlld_adj_g<true>(L, vr, bnf, gam, dv0, dv1, gF, gL, gvr, gbn, gdv0, gdv1);
        // dv0 = vn(q) - vn(q-1): slots of cells q (lane + 2) and q-1 (lane + 1)
        atomicAdd(&row[1 * (WAVE + 4) + lane + 2], gdv0);
        atomicAdd(&row[1 * (WAVE + 4) + lane + 1], -gdv0);
      } else {
        hlld_adj(L, vr, bnf, gam, gF, gL, gvr, gbn);  // d(0.5 (b + b))/db = 1
      }
    }
    if constexpr (RMN >= 2) {
      // dv1's adjoint per face, natural layout, one writer per face (the
      // owning lanes); gin adds it onto the transverse velocity
      if (lane >= 3 && lane < 3 + SEGA && q >= 0 && q < n)
        p.gdv[DIR][DIR == 0 ? base + q : (long long)g.nx * q + line] = gdv1;
    }
    // this cell's left state is the left state of face q+1 (lane l+1)
#pragma unroll
    for (int k = 0; k < 7; k++) gvl[k] = __shfl_down(gL[k], 1, WAVE);
    gql = __shfl_down(gqlf, 1, WAVE);
  }
  // ---- 3: reconstruction adjoint of cell q
  SlotAcc acc{row, lane};
  double gcn = 0.0, gbm = 0.0, gbp = 0.0;
  if (q >= 2 && q <= n - 3) {
    StencilLd<DIR> ld;
    ld.a[0] = p.in[0];
    ld.a[1] = DIR == 0 ? p.vx : p.vy;
    ld.a[2] = DIR == 0 ? p.vy : p.vz;
    ld.a[3] = DIR == 0 ? p.vz : p.vx;
    ld.a[4] = DIR == 0 ? p.cy : p.in[6];
    ld.a[5] = DIR == 0 ? p.in[6] : p.cx;
    ld.a[6] = p.pr;
    ld.sc = fq;
    ld.o = o;
    const double* cn = DIR == 0 ? p.cx : p.cy;
    const double* bf = DIR == 0 ? p.in[4] : p.in[5];
    lrstate_adj_g<true>(ld, cn[fq], gam, gvl, gvr, acc, gcn, fbc);
    // lr_fb: p = b_pair * v_n, q = v_pair (x: cy, vy; y: cx, vx)
    double pp[5], qq[5], gpp[5] = {}, gqq[5] = {};
#pragma unroll
    for (int m = 0; m < 5; m++) {
      pp[m] = ld(DIR == 0 ? 4 : 5, m) * ld(1, m);
      qq[m] = ld(DIR == 0 ? 2 : 3, m);
    }
    lr_fb_adj(pp, qq, bf[fq], bf[fq + o], gql, gqr, gpp, gqq, gbm, gbp);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      if (DIR == 0) {
        acc(1, m, gpp[m] * ld(4, m));
        acc(4, m, gpp[m] * ld(1, m));
        acc(2, m, gqq[m]);
      } else {
        acc(1, m, gpp[m] * ld(5, m));
        acc(5, m, gpp[m] * ld(1, m));
        acc(3, m, gqq[m]);
      }
    }
  }
  // every lane's adds are in the row (one wavefront, LDS operations in order)
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double gc7[7];
#pragma unroll
  for (int k = 0; k < 7; k++) gc7[k] = row[k * (WAVE + 4) + lane + 2];
  const double gbpl = __shfl_up(gbp, 1, WAVE);
  const bool own = lane >= 3 && lane < 3 + SEGA && q >= 0 && q < n;
  if (!own) return;
  const long long c = base + o * q;
  const double gface = gbn + gbm + gbpl;
  // stencil rows -> adjoint arrays (ro vx vy vz cx cy bz pr)
  const int gi[7] = {0, DIR == 0 ? 1 : 2, DIR == 0 ? 2 : 3, DIR == 0 ? 3 : 1, DIR == 0 ? 5 : 6, DIR == 0 ? 6 : 4, 7};
  if (TR) {  // transposed scratch, written once per cell, added back by k_adj_untranspose_add
#pragma unroll
    for (int k = 0; k < 7; k++) p.gp[gi[k]][c] = gc7[k];
    p.gp[5][c] = gcn;
    p.gbf[1][c] = gface;
  } else {  // the x pass runs first and writes every cell: it sets the arrays
#pragma unroll
    for (int k = 0; k < 7; k++) p.gp[gi[k]][c] = gc7[k];
    p.gp[4][c] = gcn;
    p.gbf[0][c] = gface;
  }
}

// Both directions in one launch: workgroups [0, nbx) run the x pass, the
// rest the y pass (the two are independent: the x pass sets the adjoint
// arrays, the y pass writes transposed scratch).  On small domains one
// pass fills less than two rounds of the 1024 wave slots (one wave per
// SIMD), so running them side by side saves a partly empty round.  One
// wave per SIMD: the reverse reconstruction needs about 350 registers (held
// to 256 for two waves it spills and runs slower, measured).
template <int RMN>
__global__ __launch_bounds__(256, 1) void k_adj_sweep_xy(AdjPtrs px, AdjPtrs py, Geo g, double gam, int nsx,
                                                         int nsy, int nbx)
{
  const int b = (int)blockIdx.x;
  if (b < nbx) adj_sweep_body<RMN, 0, false>(px, g, gam, nsx, b);
  else adj_sweep_body<RMN, 1, true>(py, g, gam, nsy, b - nbx);
}

// ---------------------------------------------------------------- cell
// Adjoint of the stage input at cell s before the ghost fold (gI):
//   gin_prim_at: ro, mx, my, mz, bz, en from the primitive recovery
//     (prim2d_cell, the y sweep's staging) with the CUCT cell terms, plus the
//     RK pass-through b * g_out;
//   gin_face_at: bx, by from the face-B terms of the sweeps and the CUCT
//     corners, the adjoint of fcen (cx, cy of the kept cells, with their
//     ghost copies folded on), plus b * g_out.
// The cell-centred B adjoint totals (sweep terms + the pressure term of
// the primitive recovery) are formed where fcen's adjoint reads them.
// MLAU: the adjoint of the faces' dv1 = min(dvt(left cell), dvt(right
// cell)) onto the transverse velocity of cell (i, j), dvt(c) = min(fwd,
// bwd) of that velocity at c's resolved kept cell.  Gather form: for each
// kept centre Q within one cell of (i, j) along the difference axis whose
// selected difference reads (i, j), every face on the periodic images of Q
// that selected Q adds its stored adjoint.
__device__ __forceinline__ double dv1_adj_vy(const AdjPtrs& p, const Geo& g, int i, int j)
{
  const long long nx = g.nx;
  const double* vy = p.vy;
  double acc = 0.0;
#pragma unroll
  for (int d = -1; d <= 1; d++) {  // (i, j) = Q + d rows
    const int jc = j - d;
    if (!kept(i, g.nx) || !kept(jc, g.ny)) continue;
    const long long Q = nx * jc + i;
    const double fwd = vy[Q + nx] - vy[Q], bwd = vy[Q] - vy[Q - nx];
    const bool useb = fwd >= bwd;
    const double coef = d == 1 ? (useb ? 0.0 : 1.0) : (d == 0 ? (useb ? 1.0 : -1.0) : (useb ? -1.0 : 0.0));
    if (coef == 0.0) continue;
    const double dq = useb ? bwd : fwd;
    const Img ir = images(jc, g.ny), ic = images(i, g.nx);
    for (int b = 0; b < ir.n; b++)
      for (int a = 0; a < ic.n; a++) {
        const int r = ir.i[b], col = ic.i[a];
        int f = col + 1;  // this cell on the face's left
        if (f >= 3 && f <= g.nx - 3) {
          const double other = dv_per(vy, col + 1, r, g.nx, g.ny, 1, nx, nx);
          if (!(dq >= other)) acc += coef * p.gdv[0][nx * r + f];
        }
        f = col;  // this cell on the face's right
        if (f >= 3 && f <= g.nx - 3) {
          const double other = dv_per(vy, col - 1, r, g.nx, g.ny, 1, nx, nx);
          if (other >= dq) acc += coef * p.gdv[0][nx * r + f];
        }
      }
  }
  return acc;
}
__device__ __forceinline__ double dv1_adj_vx(const AdjPtrs& p, const Geo& g, int i, int j)
{
  const long long nx = g.nx;
  const double* vx = p.vx;
  double acc = 0.0;
#pragma unroll
  for (int d = -1; d <= 1; d++) {  // (i, j) = Q + d columns
    const int icq = i - d;
    if (!kept(icq, g.nx) || !kept(j, g.ny)) continue;
    const long long Q = nx * j + icq;
    const double fwd = vx[Q + 1] - vx[Q], bwd = vx[Q] - vx[Q - 1];
    const bool useb = fwd >= bwd;
    const double coef = d == 1 ? (useb ? 0.0 : 1.0) : (d == 0 ? (useb ? 1.0 : -1.0) : (useb ? -1.0 : 0.0));
    if (coef == 0.0) continue;
    const double dq = useb ? bwd : fwd;
    const Img ic = images(icq, g.nx), ir = images(j, g.ny);
    for (int a = 0; a < ic.n; a++)
      for (int b = 0; b < ir.n; b++) {
        const int c = ic.i[a], rr = ir.i[b];
        int f = rr + 1;  // this cell below the face
        if (f >= 3 && f <= g.ny - 3) {
          const double other = dv_per(vx, c, rr + 1, g.nx, g.ny, 1, nx, 1);
          if (!(dq >= other)) acc += coef * p.gdv[1][nx * f + c];
        }
        f = rr;  // this cell above the face
        if (f >= 3 && f <= g.ny - 3) {
          const double other = dv_per(vx, c, rr - 1, g.nx, g.ny, 1, nx, 1);
          if (other >= dq) acc += coef * p.gdv[1][nx * f + c];
        }
      }
  }
  return acc;
}

template <int CTW, bool DV>
__device__ __forceinline__ void gin_prim_at(const AdjPtrs& p, const Geo& g, double gam, long long s)
{
  const int i = (int)(s % g.nx), j = (int)(s / g.nx);
  const long long nx = g.nx;
  const double gm1 = gam - 1.0;
  const double b = p.b;
  {  // primitive recovery
    double cro = 0.0, cvx = 0.0, cvy = 0.0;
    if (CTW) {
      // this cell is entry `role` of the corners (i+1,j+1), (i,j+1), (i+1,j), (i,j)
#pragma unroll
      for (int role = 0; role < 4; role++) {
        const int ic = i + ((role == 0 || role == 2) ? 1 : 0);
        const int jc = j + (role <= 1 ? 1 : 0);
        const bool ok = corner_ok(ic, jc, g);
        const long long ss = ok ? nx * jc + ic : s;  // loads unconditional, terms selected
        const double a0 = p.gcc[0][ss], a1 = p.gcc[1][ss], a2 = p.gcc[2][ss];
        cro += ok ? a0 : 0.0;
        cvx += ok ? a1 : 0.0;
        cvy += ok ? a2 : 0.0;
      }
    }
    double gro = p.gp[0][s] + cro;
    double gvx = p.gp[1][s] + cvx, gvy = p.gp[2][s] + cvy, gvz = p.gp[3][s];
    if (DV) {  // MLAU shock sensor, transverse part
      gvx += dv1_adj_vx(p, g, i, j);
      gvy += dv1_adj_vy(p, g, i, j);
    }
    double gbz = p.gp[6][s];
    const double gpr = p.gp[7][s];
    // forward of the primitive recovery (prim2d_cell)
    const double ro = p.in[0][s];
    const double iro = 1.0 / ro;
    const double mx = p.in[1][s], my = p.in[2][s], mz = p.in[3][s];
    const double vx = mx * iro, vy = my * iro, vz = mz * iro;
    const double cz = p.in[6][s];
    const double v2 = vx * vx + vy * vy + vz * vz;
    // pr = (gam-1) (en - (ro v2 + b2) / 2)
    const double gen = gm1 * gpr;
    gro += -gm1 * 0.5 * v2 * gpr;
    gvx += -gm1 * ro * vx * gpr;
    gvy += -gm1 * ro * vy * gpr;
    gvz += -gm1 * ro * vz * gpr;
    gbz += -gm1 * cz * gpr;
    // v = m / ro
    const double giro = mx * gvx + my * gvy + mz * gvz;
    gro += -iro * iro * giro;
    const double go0 = p.gO[0][s], go1 = p.gO[1][s], go2 = p.gO[2][s], go3 = p.gO[3][s];
    const double go6 = p.gO[6][s], go7 = p.gO[7][s];
    p.gI[0][s] = b * go0 + gro;
    p.gI[1][s] = b * go1 + iro * gvx;
    p.gI[2][s] = b * go2 + iro * gvy;
    p.gI[3][s] = b * go3 + iro * gvz;
    p.gI[6][s] = b * go6 + gbz;
    p.gI[7][s] = b * go7 + gen;
  }
}

template <int CTW>
__device__ __forceinline__ void gin_face_at(const AdjPtrs& p, const Geo& g, double gam, long long s)
{
  const int i = (int)(s % g.nx), j = (int)(s / g.nx);
  const long long nx = g.nx;
  const double gm1 = gam - 1.0;
  const double b = p.b;
  // face B
  double gbx = b * p.gO[4][s] + p.gbf[0][s];
  double gby = b * p.gO[5][s] + p.gbf[1][s];
  if (CTW) {
    // bx(i,j) enters the corners (i,j) and (i,j+1); by(i,j) the corners
    // (i,j) and (i+1,j)
    const bool ok0 = corner_ok(i, j, g), ok1 = corner_ok(i, j + 1, g), ok2 = corner_ok(i + 1, j, g);
    const long long s1 = ok1 ? nx * (j + 1) + i : s, s2 = ok2 ? nx * j + i + 1 : s;
    const double b0 = p.gcc[3][s], c0 = p.gcc[4][s], b1 = p.gcc[3][s1], c2 = p.gcc[4][s2];
    gbx += ok0 ? b0 : 0.0;
    gby += ok0 ? c0 : 0.0;
    gbx += ok1 ? b1 : 0.0;
    gby += ok2 ? c2 : 0.0;
  }
  // cx(q) = fcen(bx(q-1..q+2)) at kept q; ghost copies of cx fold onto q.
  // Adjoint total of cx at q: sweep terms + pressure term -gm1 cx gpr.
  auto tcx = [&](long long q) { return p.gp[4][q] + -gm1 * p.cx[q] * p.gp[7][q]; };
  auto tcy = [&](long long q) { return p.gp[5][q] + -gm1 * p.cy[q] * p.gp[7][q]; };
  const double fw[4] = {-0.0625, 0.5625, 0.5625, -0.0625};  // offsets -1..2
  // (straight-line: out-of-range terms read a clamped kept cell and are
  // dropped by selects, so all the loads can be in flight together)
  const int jk = j < NG ? NG : (j > g.ny - NG - 1 ? g.ny - NG - 1 : j);
  const int ik = i < NG ? NG : (i > g.nx - NG - 1 ? g.nx - NG - 1 : i);
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const int iq = i - (kk - 1);
    const int iqk = iq < NG ? NG : (iq > g.nx - NG - 1 ? g.nx - NG - 1 : iq);
    const double v = fold_var(tcx, iqk, jk, g, p.bc, 4, true);
    gbx += (kept(j, g.ny) && kept(iq, g.nx)) ? fw[kk] * v : 0.0;
  }
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const int jq = j - (kk - 1);
    const int jqk = jq < NG ? NG : (jq > g.ny - NG - 1 ? g.ny - NG - 1 : jq);
    const double v = fold_var(tcy, ik, jqk, g, p.bc, 5, true);
    gby += (kept(i, g.nx) && kept(jq, g.ny)) ? fw[kk] * v : 0.0;
  }
  p.gI[4][s] = gbx;
  p.gI[5][s] = gby;
}

// the stage input adjoint at every cell, ghosts included (gI): workgroups
// [0, nbh) the six variables of the primitive recovery, the rest face B
// (two independent halves: shorter per-thread chains on small domains)
template <int CTW, bool DV>
__device__ __forceinline__ void gin_block(const AdjPtrs& p, const Geo& g, double gam, int nbh, int b)
{
  const long long t = (long long)(b < nbh ? b : b - nbh) * blockDim.x + threadIdx.x;
  if (t >= g.nd) return;
  if (b < nbh) gin_prim_at<CTW, DV>(p, g, gam, t);
  else gin_face_at<CTW>(p, g, gam, t);
}

template <int CTW, bool DV>
__global__ __launch_bounds__(256) void k_adj_gin(AdjPtrs p, Geo g, double gam, int nbh)
{
  gin_block<CTW, DV>(p, g, gam, nbh, (int)blockIdx.x);
}

// grid (cells, 8 variables): per kept cell, gI folded over the cell's
// ghost images (the adjoint of bound()) becomes the output adjoint of the
// previous stage, written in place of this stage's (each thread reads its
// own entry first; ghost entries stay zero).  The step-base adjoint takes
// a * g_out (set by the first stage processed, added by the others); the
// last one processed adds it to its result.
__device__ __forceinline__ void foldstage_at(const AdjPtrs& p, const Geo& g, int first, int last, long long t, int m)
{
  if (t >= g.nd) return;
  const int i = (int)(t % g.nx), j = (int)(t / g.nx);
  if (!source_cell(i, j, g, p.bc, m)) return;
  const double* gi = p.gI[m];
  double v = fold_var([gi](long long q) { return gi[q]; }, i, j, g, p.bc, m);
  const double go = p.gO[m][t];
  if (p.a != 0.0) {
    const double gu = first ? 0.0 : p.gU0[m][t];
    p.gU0[m][t] = first ? p.a * go : gu + p.a * go;
  }
  if (last) v += p.gU0[m][t];
  p.gO[m][t] = v;
}

__global__ __launch_bounds__(256) void k_adj_foldstage(AdjPtrs p, Geo g, int first, int last)
{
  foldstage_at(p, g, first, last, (long long)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y);
}

// ---------------------------------------------------------------- fold
struct FoldArgs {
  const double* src[8];
  double* dst[8];
  const double* extra[8];  // added at kept cells when non-null
  BcFlags bc;
};

// grid: (cells, 8 variables)
__device__ __forceinline__ void fold_cell(const FoldArgs& f, const Geo& g, long long t, int m)
{
  if (t >= g.nd) return;
  const int i = (int)(t % g.nx), j = (int)(t / g.nx);
  double v = 0.0;
  if (source_cell(i, j, g, f.bc, m)) {
    const double* src = f.src[m];
    v = fold_var([src](long long q) { return src[q]; }, i, j, g, f.bc, m);
    if (f.extra[m]) v += f.extra[m][t];
  }
  f.dst[m][t] = v;
}

__global__ __launch_bounds__(256) void k_adj_fold(FoldArgs f, Geo g)
{
  fold_cell(f, g, (long long)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y);
}

// ---------------------------------------------------------------- transposes
// dst[a][i * ny + j] = src[a][j * nx + i] for array a = blockIdx.z, through
// a 32 x 32 LDS tile (both global accesses run along a contiguous axis);
// the back form accumulates dst[a][j * nx + i] += src[a][i * ny + j] for
// a < nadd and stores it for the others.
constexpr int NTR = 15;
struct TrArgs {
  const double* src[NTR];
  double* dst[NTR];
  int nadd;
};

// tile (bx, by) of array bz: one 32 x 32 tile per 256-thread workgroup
__device__ __forceinline__ void transpose_tile(const TrArgs& a, int nx, int ny, int bx, int by, int bz)
{
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i0 = bx * 32, j0 = by * 32;
  const double* s = a.src[bz];
  double* d = a.dst[bz];
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < nx && j < ny) tile[r][tx] = s[(long long)j * nx + i];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + r;
    if (i < nx && j < ny) d[(long long)i * ny + j] = tile[tx][r];
  }
}

// One launch: workgroups [0, nbc) run the corner adjoint (one thread per
// corner), the rest transpose the arrays the y line pass reads (they do not
// depend on the corner outputs).
template <int CTW>
__device__ __forceinline__ void corner_tr_block(const AdjPtrs& p, const Geo& g, const TrArgs& a, int nbc, int b)
{
  if (b < nbc) {
    corner_at<CTW>(p, g, (long long)b * blockDim.x + threadIdx.x);
    return;
  }
  const int ntx = (g.nx + 31) / 32, nty = (g.ny + 31) / 32;
  const int k = b - nbc;
  const int bz = k / (ntx * nty), r = k - bz * (ntx * nty);
  transpose_tile(a, g.nx, g.ny, r % ntx, r / ntx, bz);
}

template <int CTW>
__global__ __launch_bounds__(256) void k_adj_corner_tr(AdjPtrs p, Geo g, TrArgs a, int nbc)
{
  corner_tr_block<CTW>(p, g, a, nbc, (int)blockIdx.x);
}

__device__ __forceinline__ void untranspose_tile(const TrArgs& a, int nx, int ny, int bx, int by, int bz)
{
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i0 = bx * 32, j0 = by * 32;
  const double* s = a.src[bz];
  double* d = a.dst[bz];
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + r;
    if (i < nx && j < ny) tile[r][tx] = s[(long long)i * ny + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < nx && j < ny) {
      if (bz < a.nadd) d[(long long)j * nx + i] += tile[tx][r];
      else d[(long long)j * nx + i] = tile[tx][r];
    }
  }
}

__global__ __launch_bounds__(256) void k_adj_untranspose_add(TrArgs a, int nx, int ny)
{
  untranspose_tile(a, nx, ny, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}

// ---------------------------------------------------------------- batches
// Batched stage adjoints (qas_batch_adj_backward): one launch per kernel for
// every instance of a batch.  Instance i owns the workgroups [off[i],
// off[i+1]) of a launch (per-kernel prefix sums built on the host), finds
// itself by binary search on its uniform workgroup index and runs the
// single-domain body with its own descriptor and local workgroup index.
struct AdjInst {
  AdjPtrs st[3];   // stage rk: the x pass and the element-wise kernels
  AdjPtrs tst[3];  // stage rk: the y pass on the transposed copies
  TrArgs tin[3];   // stage rk: the arrays the y pass reads, to transpose
  TrArgs tout;     // the y pass's outputs, added back
  FoldArgs fin;    // the staged final-state adjoint, folded into gO
  Geo g;
  int nb;          // cell workgroups (256 cells)
  int nbx, csx, csy;  // x line-pass workgroups; segments per x and y line
  int ntt;         // 32 x 32 tiles of one array
};
enum AdjBatchKernel { KA_FOLD = 0, KA_CORNER, KA_FACE, KA_SWEEP, KA_UNTR, KA_GIN, KA_FSTAGE, KA_COUNT };

#define QAS_ABI                                                                                         \
  const int b_ = (int)blockIdx.x, i_ = batch_find(off, n, b_), l = b_ - off[i_];                          \
  const AdjInst& I_ = I[i_];

// (the descriptors are used in place: a local copy indexed by a variable
// or array number would live in scratch memory)
__global__ __launch_bounds__(256) void kba_fold(const AdjInst* I, const int* off, int n)
{
  QAS_ABI
  const Geo g = I_.g;
  fold_cell(I_.fin, g, (long long)(l % I_.nb) * blockDim.x + threadIdx.x, l / I_.nb);
}

template <int CTW>
__global__ __launch_bounds__(256) void kba_corner_tr(const AdjInst* I, const int* off, int n, int rk)
{
  QAS_ABI
  const Geo g = I_.g;
  corner_tr_block<CTW>(I_.st[rk], g, I_.tin[rk], I_.nb, l);
}

__global__ __launch_bounds__(256) void kba_faceparts(const AdjInst* I, const int* off, int n, int rk)
{
  QAS_ABI
  const AdjPtrs p = I_.st[rk];
  const Geo g = I_.g;
  faceparts_at(p, g, (long long)l * blockDim.x + threadIdx.x);
}

template <int RMN>
__global__ __launch_bounds__(256, 1) void kba_sweep_xy(const AdjInst* I, const int* off, int n, int rk, double gam)
{
  QAS_ABI
  const Geo g = I_.g;
  if (l < I_.nbx) {
    const AdjPtrs p = I_.st[rk];
    adj_sweep_body<RMN, 0, false>(p, g, gam, I_.csx, l);
  } else {
    const AdjPtrs p = I_.tst[rk];
    adj_sweep_body<RMN, 1, true>(p, g, gam, I_.csy, l - I_.nbx);
  }
}

__global__ __launch_bounds__(256) void kba_untranspose(const AdjInst* I, const int* off, int n)
{
  QAS_ABI
  const int ntx = (I_.g.nx + 31) / 32;
  const int bz = l / I_.ntt, r = l - bz * I_.ntt;
  untranspose_tile(I_.tout, I_.g.nx, I_.g.ny, r % ntx, r / ntx, bz);
}

template <int CTW, bool DV>
__global__ __launch_bounds__(256) void kba_gin(const AdjInst* I, const int* off, int n, int rk, double gam)
{
  QAS_ABI
  const AdjPtrs p = I_.st[rk];
  const Geo g = I_.g;
  gin_block<CTW, DV>(p, g, gam, I_.nb, l);
}

__global__ __launch_bounds__(256) void kba_foldstage(const AdjInst* I, const int* off, int n, int rk, int first,
                                                     int last)
{
  QAS_ABI
  const Geo g = I_.g;
  foldstage_at(I_.st[rk], g, first, last, (long long)(l % I_.nb) * blockDim.x + threadIdx.x, l / I_.nb);
}
#undef QAS_ABI

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace qas

using namespace qas;

// ---------------------------------------------------------------- workspace
struct qas_adj_work {
  void* slab = nullptr;
  size_t stride = 0;
  double* S[3][8] = {};       // recomputed stage inputs S1, S2 and a discard output
  double* prim[3][6] = {};    // per stage: vx vy vz pr cx cy
  double* emf[3][8] = {};     // per stage: face parts (slots 0, 1, 4, 5 used)
  double* ez[3] = {};
  double* gO[8] = {};
  double* gU0[8] = {};
  double* gI[8] = {};
  double* gp[8] = {};
  double* gbf[2] = {};
  double* gemf[4] = {};
  double* gdv[2] = {};
  double *gez = nullptr, *ctv = nullptr;
  double* gcc[5] = {};
  double* tin[NTR] = {};      // transposed inputs of the y pass
  double* tout[9] = {};       // its transposed outputs: gp[0..7], gbf[1]
  double* ck = nullptr;       // checkpoints: nsteps x 8 x stride
  double* ks = nullptr;       // kept stages (qas_adj_keep_stages): nsteps x KEPT x stride
  int ks_cap = 0;
  bool kept = false;          // the recorded forward kept its stages
  double* d_dts = nullptr;    // the recorded step sizes (device), ck_cap + 1 entries
  double* dslot = nullptr;    // d_dts[ck_cap]: the step size a batched backward reads
  double* stage = nullptr;    // qas_adj_backward's device copies of g_final, g_init
  const double* cur_dt = nullptr;  // d_dts + the step being reversed
  int ck_cap = 0;
  int nrec = 0;               // steps recorded
  std::vector<double> dts;
};

namespace qas {
void adj_release(qas_ctx* c)
{
  if (!c || !c->adj) return;
  if (c->adj->slab) (void)hipFree(c->adj->slab);
  if (c->adj->ck) (void)hipFree(c->adj->ck);
  if (c->adj->ks) (void)hipFree(c->adj->ks);
  if (c->adj->d_dts) (void)hipFree(c->adj->d_dts);
  if (c->adj->stage) (void)hipFree(c->adj->stage);
  delete c->adj;
  c->adj = nullptr;
}
}  // namespace qas

static int afail(qas_ctx* c, int code, const std::string& m)
{
  if (c) c->err = m;
  return code;
}

#define QAS_ACHECK(c, expr)                                                                               \
  do {                                                                                                    \
    hipError_t _e = (expr);                                                                               \
    if (_e != hipSuccess) return afail((c), QAS_E_HIP, std::string(#expr) + ": " + hipGetErrorName(_e)); \
  } while (0)

static std::string adj_supported(const qas_ctx* c)
{
  if (c->g.ndim != 2) return "the adjoint covers the 2D path only";
  if (c->s.precision != 0) return "the adjoint covers the FP64 path only";
  if (c->s.rmn < 1 || c->s.rmn > 3) return "the adjoint covers the HLLD, LHLLD and MLAU solvers (rmn 1-3) only";
  if (c->s.odr != 4) return "the adjoint covers odr 4 only";
  if (c->s.r_k != 3) return "the adjoint covers r_k 3 only";
  if (c->d_phi || c->d_nu) return "the adjoint covers the ideal step without gravity or dissipation";
  // any boundary kind with HLLD; with LHLLD and MLAU periodic only (the
  // adjoint of their shock sensor gathers over periodic images)
  for (int a = 0; a < 2; a++)
    for (int m = 0; m < 8; m++)
      if (c->s.dn[a][m] != 0 && c->s.rmn != 1)
        return "the adjoint covers non-periodic boundaries with HLLD (rmn 1) only";
  return "";
}

static int adj_workspace(qas_ctx* c)
{
  if (c->adj) return QAS_OK;
  qas_adj_work* w = new qas_adj_work();
  c->adj = w;
  const size_t stride = (size_t)((c->geo.nd + 63) / 64 * 64);
  w->stride = stride;
  const size_t narr = 24 + 18 + 24 + 3 + 24 + 8 + 2 + 4 + 2 + 2 + 5 + NTR + 9;
  hipError_t e = hipMalloc(&w->slab, narr * stride * sizeof(double));
  if (e != hipSuccess) {
    adj_release(c);
    return afail(c, QAS_E_NOMEM, std::string("adjoint workspace: ") + hipGetErrorName(e));
  }
  double* p = (double*)w->slab;
  auto take = [&]() { double* r = p; p += stride; return r; };
  for (int s = 0; s < 3; s++)
    for (int m = 0; m < 8; m++) w->S[s][m] = take();
  for (int s = 0; s < 3; s++)
    for (int k = 0; k < 6; k++) w->prim[s][k] = take();
  for (int s = 0; s < 3; s++)
    for (int k = 0; k < 8; k++) w->emf[s][k] = take();
  for (int s = 0; s < 3; s++) w->ez[s] = take();
  for (int m = 0; m < 8; m++) w->gO[m] = take();
  for (int m = 0; m < 8; m++) w->gU0[m] = take();
  for (int m = 0; m < 8; m++) w->gI[m] = take();
  for (int m = 0; m < 8; m++) w->gp[m] = take();
  for (int k = 0; k < 2; k++) w->gbf[k] = take();
  for (int k = 0; k < 4; k++) w->gemf[k] = take();
  for (int k = 0; k < 2; k++) w->gdv[k] = take();
  w->gez = take();
  w->ctv = take();
  for (int k = 0; k < 5; k++) w->gcc[k] = take();
  for (int k = 0; k < NTR; k++) w->tin[k] = take();
  for (int k = 0; k < 9; k++) w->tout[k] = take();
  QAS_ACHECK(c, hipMemsetAsync(w->slab, 0, narr * stride * sizeof(double), c->stream));
  return QAS_OK;
}

// Where a stage's forward intermediates live: the recompute slots of the
// workspace, or a step's entries of the kept-stage store.
struct StageSlot {
  double* prim[6];  // vx vy vz pr cx cy
  double* emf[8];   // face parts (the 2D sweeps write slots 0, 1, 4, 5)
};
static StageSlot work_slot(const qas_adj_work* w, int s)
{
  StageSlot r;
  for (int k = 0; k < 6; k++) r.prim[k] = w->prim[s][k];
  for (int k = 0; k < 8; k++) r.emf[k] = w->emf[s][k];
  return r;
}
// kept-stage store, per step: the inputs of stages 2 and 3 (8 + 8 arrays),
// then per stage its 6 primitives and its 4 face parts
constexpr int KEPT = 16 + 3 * 6 + 3 * 4;
static double* kept_arr(const qas_adj_work* w, int n, int k)
{
  return w->ks + ((size_t)n * KEPT + (size_t)k) * w->stride;
}
static StageSlot kept_slot(const qas_adj_work* w, int n, int s)
{
  StageSlot r;
  for (int k = 0; k < 6; k++) r.prim[k] = kept_arr(w, n, 16 + 6 * s + k);
  for (int k = 0; k < 8; k++) r.emf[k] = nullptr;
  const int q[4] = {0, 1, 4, 5};
  for (int k = 0; k < 4; k++) r.emf[q[k]] = kept_arr(w, n, 34 + 4 * s + k);
  return r;
}

// forward stage rk with its primitives and face parts in `sl`; when
// `whole` is false the last stage's output is not needed (only its
// primitives and face parts are), so its corner EMF, CT and ghost fill are
// skipped; every other stage runs as ideal(dt) runs it, ghost fill included
static hipError_t fwd_stage(qas_ctx* c, const StageSlot& sl, double* ez, double* const in[8],
                            double* const u0[8], double* const out[8], int rk, bool whole)
{
  qas_adj_work* w = c->adj;
  StagePtrs p;
  std::memset(&p, 0, sizeof(p));
  for (int m = 0; m < 8; m++) {
    p.in[m] = in[m];
    p.out[m] = out[m];
    p.u0[m] = u0[m];
  }
  p.vx = sl.prim[0]; p.vy = sl.prim[1]; p.vz = sl.prim[2]; p.pr = sl.prim[3];
  p.cx = sl.prim[4]; p.cy = sl.prim[5];
  for (int k = 0; k < 6; k++) p.acc[k] = c->acc[k];
  for (int k = 0; k < 8; k++) p.emf[k] = sl.emf[k];
  p.ez = ez;
  p.dt = w->cur_dt;
  const double ra[3] = {0.0, 0.75, 1.0 / 3.0}, rb[3] = {1.0, 0.25, 2.0 / 3.0};
  p.a = ra[rk];
  p.b = rb[rk];
  const bool faces_only = rk == 2 && !whole;
  return launch_stage2d(c, p, faces_only);
}

// adjoint of stage rk (0..2) whose input is `in` and whose forward
// intermediates sit in slot rk; reads w->gO, leaves the folded input adjoint
// in w->gO (plus the step-base adjoint when `last`)
static hipError_t adj_stage(qas_ctx* c, int rk, const StageSlot& sl, double* const in[8], bool last)
{
  qas_adj_work* w = c->adj;
  const Geo& g = c->geo;
  hipStream_t st = c->stream;
  AdjPtrs p;
  std::memset(&p, 0, sizeof(p));
  for (int m = 0; m < 8; m++) {
    p.in[m] = in[m];
    p.gO[m] = w->gO[m];
    p.gU0[m] = w->gU0[m];
    p.gI[m] = w->gI[m];
    p.gp[m] = w->gp[m];
    p.emf[m] = sl.emf[m];
  }
  p.vx = sl.prim[0]; p.vy = sl.prim[1]; p.vz = sl.prim[2]; p.pr = sl.prim[3];
  p.cx = sl.prim[4]; p.cy = sl.prim[5];
  for (int k = 0; k < 2; k++) p.gbf[k] = w->gbf[k];
  for (int k = 0; k < 4; k++) p.gemf[k] = w->gemf[k];
  for (int k = 0; k < 2; k++) p.gdv[k] = w->gdv[k];
  p.gez = w->gez;
  p.ctv = w->ctv;
  for (int k = 0; k < 5; k++) p.gcc[k] = w->gcc[k];
  p.dt = w->cur_dt;
  const double ra[3] = {0.0, 0.75, 1.0 / 3.0}, rb[3] = {1.0, 0.25, 2.0 / 3.0};
  p.a = ra[rk];
  p.b = rb[rk];
  p.bc = c->bc;
  const unsigned nb = nblk(g.nd, 256);
  const bool ctw = c->s.ctw != 0;
  const int csx = (g.nx + SEGA - 1) / SEGA, csy = (g.ny + SEGA - 1) / SEGA;
  {
    // the y pass runs on transposed copies of what it reads (columns
    // contiguous), made in the corner launch; its outputs are added back
    // onto gp, gbf[1]
    const double* src[NTR] = {p.in[0], p.in[5], p.in[6], p.vx, p.vy, p.vz, p.cx, p.cy, p.pr,
                              p.gO[0], p.gO[1], p.gO[2], p.gO[3], p.gO[6], p.gO[7]};
    TrArgs ti;
    for (int k = 0; k < NTR; k++) {
      ti.src[k] = src[k];
      ti.dst[k] = w->tin[k];
    }
    ti.nadd = 0;
    const unsigned ntr = (unsigned)(((g.nx + 31) / 32) * ((g.ny + 31) / 32) * NTR);
    if (ctw) hipLaunchKernelGGL((k_adj_corner_tr<1>), nb + ntr, 256, 0, st, p, g, ti, (int)nb);
    else hipLaunchKernelGGL((k_adj_corner_tr<0>), nb + ntr, 256, 0, st, p, g, ti, (int)nb);
    hipLaunchKernelGGL(k_adj_faceparts, nb, 256, 0, st, p, g);
    AdjPtrs pt = p;
    pt.in[0] = w->tin[0];
    pt.in[5] = w->tin[1];
    pt.in[6] = w->tin[2];
    pt.vx = w->tin[3];
    pt.vy = w->tin[4];
    pt.vz = w->tin[5];
    pt.cx = w->tin[6];
    pt.cy = w->tin[7];
    pt.pr = w->tin[8];
    const int gom[6] = {0, 1, 2, 3, 6, 7};
    for (int v = 0; v < 6; v++) pt.gO[gom[v]] = w->tin[9 + v];
    // (pt.gemf[2], pt.gemf[3]: in place, see the line pass)
    for (int m = 0; m < 8; m++) pt.gp[m] = w->tout[m];
    pt.gbf[1] = w->tout[8];
    // x pass in place, y pass on the transposed copies, one launch
    const unsigned nbx = nblk((long long)csx * g.ny * WAVE, 256), nby = nblk((long long)csy * g.nx * WAVE, 256);
    if (c->s.rmn == 3)
      hipLaunchKernelGGL(k_adj_sweep_xy<3>, nbx + nby, 256, 0, st, p, pt, g, c->s.gam, csx, csy, (int)nbx);
    else if (c->s.rmn == 2)
      hipLaunchKernelGGL(k_adj_sweep_xy<2>, nbx + nby, 256, 0, st, p, pt, g, c->s.gam, csx, csy, (int)nbx);
    else
      hipLaunchKernelGGL(k_adj_sweep_xy<1>, nbx + nby, 256, 0, st, p, pt, g, c->s.gam, csx, csy, (int)nbx);
    TrArgs to;
    for (int k = 0; k < NTR; k++) {
      to.src[k] = nullptr;
      to.dst[k] = nullptr;
    }
    for (int m = 0; m < 8; m++) {
      to.src[m] = w->tout[m];
      to.dst[m] = w->gp[m];
    }
    to.src[8] = w->tout[8];
    to.dst[8] = w->gbf[1];
    to.nadd = 8;  // gp gets the y terms added, gbf[1] only has them
    const dim3 tb((unsigned)((g.nx + 31) / 32), (unsigned)((g.ny + 31) / 32), 9);
    hipLaunchKernelGGL(k_adj_untranspose_add, tb, 256, 0, st, to, g.nx, g.ny);
  }
  const bool dv = c->s.rmn == 2 || c->s.rmn == 3;  // the shock sensor feeds LHLLD and MLAU
  if (ctw && dv) hipLaunchKernelGGL((k_adj_gin<1, true>), 2 * nb, 256, 0, st, p, g, c->s.gam, (int)nb);
  else if (ctw) hipLaunchKernelGGL((k_adj_gin<1, false>), 2 * nb, 256, 0, st, p, g, c->s.gam, (int)nb);
  else if (dv) hipLaunchKernelGGL((k_adj_gin<0, true>), 2 * nb, 256, 0, st, p, g, c->s.gam, (int)nb);
  else hipLaunchKernelGGL((k_adj_gin<0, false>), 2 * nb, 256, 0, st, p, g, c->s.gam, (int)nb);
  // stage 3 is the first one processed, stage 1 the last
  hipLaunchKernelGGL(k_adj_foldstage, dim3(nb, 8), 256, 0, st, p, g, rk == 2 ? 1 : 0, last ? 1 : 0);
  return hipGetLastError();
}

namespace qas {
std::string adj_scope(const qas_ctx* c) { return adj_supported(c); }

// Set-up of a recorded forward of nsteps steps with the given step sizes:
// scope check, workspace, checkpoint room, step sizes on the device,
// gradient staging; no step recorded yet.
int adj_prepare(qas_ctx* c, int nsteps, const double* dts)
{
  if (!c || !c->slab) return afail(c, QAS_E_ARG, "qas_adj_forward: null context");
  if (!c->uploaded) return afail(c, QAS_E_STATE, "qas_adj_forward: no state uploaded");
  if (nsteps <= 0 || !dts) return afail(c, QAS_E_ARG, "qas_adj_forward: bad arguments");
  for (int n = 0; n < nsteps; n++)
    if (!(dts[n] > 0.0)) return afail(c, QAS_E_ARG, "qas_adj_forward: dt must be positive");
  const std::string why = adj_supported(c);
  if (!why.empty()) return afail(c, QAS_E_ARG, "qas_adj_forward: " + why);
  int r = adj_workspace(c);
  if (r) return r;
  qas_adj_work* w = c->adj;
  const size_t per = 8 * w->stride;
  if (nsteps > w->ck_cap) {
    if (w->ck) QAS_ACHECK(c, hipFree(w->ck));
    if (w->d_dts) QAS_ACHECK(c, hipFree(w->d_dts));
    w->ck = nullptr;
    w->d_dts = nullptr;
    w->ck_cap = 0;
    hipError_t e = hipMalloc(&w->ck, per * (size_t)nsteps * sizeof(double));
    if (e != hipSuccess) return afail(c, QAS_E_NOMEM, std::string("adjoint checkpoints: ") + hipGetErrorName(e));
    e = hipMalloc(&w->d_dts, (size_t)(nsteps + 1) * sizeof(double));
    if (e != hipSuccess) return afail(c, QAS_E_NOMEM, std::string("adjoint step sizes: ") + hipGetErrorName(e));
    w->ck_cap = nsteps;
  }
  w->dslot = w->d_dts + w->ck_cap;
  if (!w->stage) {  // device staging of the two gradients, kept with the workspace
    hipError_t e = hipMalloc(&w->stage, 2 * 8 * (size_t)c->geo.nd * sizeof(double));
    if (e != hipSuccess) return afail(c, QAS_E_NOMEM, std::string("adjoint staging: ") + hipGetErrorName(e));
  }
  w->dts.assign(dts, dts + nsteps);
  QAS_ACHECK(c, hipMemcpyAsync(w->d_dts, w->dts.data(), (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
  w->nrec = 0;
  w->kept = false;
  return QAS_OK;
}

void adj_bufs(const qas_ctx* c, AdjBufs* b)
{
  const qas_adj_work* w = c->adj;
  b->ck = w->ck;
  b->stride = w->stride;
  b->d_dts = w->d_dts;
  b->dslot = w->dslot;
  for (int s = 0; s < 3; s++) {
    for (int m = 0; m < 8; m++) b->S[s][m] = w->S[s][m];
    for (int k = 0; k < 6; k++) b->prim[s][k] = w->prim[s][k];
    for (int k = 0; k < 8; k++) b->emf[s][k] = w->emf[s][k];
    b->ez[s] = w->ez[s];
  }
  for (int m = 0; m < 8; m++) b->gO[m] = w->gO[m];
  b->stage = w->stage;
}

void adj_set_recorded(qas_ctx* c, int nsteps) { c->adj->nrec = nsteps; }
int adj_recorded(const qas_ctx* c) { return c->adj ? c->adj->nrec : 0; }

// ---- batched stage adjoints
struct AdjBatchDev {
  int n = 0;
  AdjInst* d_inst = nullptr;
  int* d_off = nullptr;  // KA_COUNT tables of n + 1 offsets
  int nblk[KA_COUNT] = {};
  int rmn = 1, ctw = 1;
  double gam = 0.0;
};

void adjb_destroy(AdjBatchDev* d)
{
  if (!d) return;
  if (d->d_inst) (void)hipFree(d->d_inst);
  if (d->d_off) (void)hipFree(d->d_off);
  delete d;
}

int adjb_create(AdjBatchDev** out, qas_ctx* const* ctxs, double* const* ub, int n, hipStream_t st, std::string* err)
{
  *out = nullptr;
  AdjBatchDev* d = new AdjBatchDev();
  d->n = n;
  d->rmn = ctxs[0]->s.rmn;
  d->ctw = ctxs[0]->s.ctw;
  d->gam = ctxs[0]->s.gam;
  std::vector<AdjInst> inst(n);
  std::vector<int> off((size_t)KA_COUNT * (n + 1), 0);
  const double ra[3] = {0.0, 0.75, 1.0 / 3.0}, rb[3] = {1.0, 0.25, 2.0 / 3.0};
  for (int i = 0; i < n; i++) {
    qas_ctx* c = ctxs[i];
    qas_adj_work* w = c->adj;
    AdjInst& A = inst[i];
    std::memset(&A, 0, sizeof(A));
    const Geo& g = c->geo;
    A.g = g;
    for (int rk = 0; rk < 3; rk++) {
      AdjPtrs& p = A.st[rk];
      for (int m = 0; m < 8; m++) {
        p.in[m] = rk == 2 ? w->S[1][m] : (rk == 1 ? w->S[0][m] : ub[8 * (size_t)i + m]);
        p.gO[m] = w->gO[m];
        p.gU0[m] = w->gU0[m];
        p.gI[m] = w->gI[m];
        p.gp[m] = w->gp[m];
        p.emf[m] = w->emf[rk][m];
      }
      p.vx = w->prim[rk][0]; p.vy = w->prim[rk][1]; p.vz = w->prim[rk][2]; p.pr = w->prim[rk][3];
      p.cx = w->prim[rk][4]; p.cy = w->prim[rk][5];
      for (int k = 0; k < 2; k++) p.gbf[k] = w->gbf[k];
      for (int k = 0; k < 4; k++) p.gemf[k] = w->gemf[k];
      for (int k = 0; k < 2; k++) p.gdv[k] = w->gdv[k];
      p.gez = w->gez;
      p.ctv = w->ctv;
      for (int k = 0; k < 5; k++) p.gcc[k] = w->gcc[k];
      p.dt = w->dslot;
      p.a = ra[rk];
      p.b = rb[rk];
      p.bc = c->bc;
      // the y pass's view on the transposed copies (as adj_stage)
      AdjPtrs pt = p;
      pt.in[0] = w->tin[0];
      pt.in[5] = w->tin[1];
      pt.in[6] = w->tin[2];
      pt.vx = w->tin[3];
      pt.vy = w->tin[4];
      pt.vz = w->tin[5];
      pt.cx = w->tin[6];
      pt.cy = w->tin[7];
      pt.pr = w->tin[8];
      const int gom[6] = {0, 1, 2, 3, 6, 7};
      for (int v = 0; v < 6; v++) pt.gO[gom[v]] = w->tin[9 + v];
      for (int m = 0; m < 8; m++) pt.gp[m] = w->tout[m];
      pt.gbf[1] = w->tout[8];
      A.tst[rk] = pt;
      const double* src[NTR] = {p.in[0], p.in[5], p.in[6], p.vx, p.vy, p.vz, p.cx, p.cy, p.pr,
                                p.gO[0], p.gO[1], p.gO[2], p.gO[3], p.gO[6], p.gO[7]};
      for (int k = 0; k < NTR; k++) {
        A.tin[rk].src[k] = src[k];
        A.tin[rk].dst[k] = w->tin[k];
      }
      A.tin[rk].nadd = 0;
    }
    for (int m = 0; m < 8; m++) {
      A.tout.src[m] = w->tout[m];
      A.tout.dst[m] = w->gp[m];
    }
    A.tout.src[8] = w->tout[8];
    A.tout.dst[8] = w->gbf[1];
    A.tout.nadd = 8;
    for (int m = 0; m < 8; m++) {
      A.fin.src[m] = w->stage + (size_t)m * g.nd;
      A.fin.dst[m] = w->gO[m];
      A.fin.extra[m] = nullptr;
    }
    A.fin.bc = c->bc;
    A.nb = (int)nblk(g.nd, 256);
    A.csx = (g.nx + SEGA - 1) / SEGA;
    A.csy = (g.ny + SEGA - 1) / SEGA;
    A.nbx = (int)nblk((long long)A.csx * g.ny * WAVE, 256);
    const int nby = (int)nblk((long long)A.csy * g.nx * WAVE, 256);
    A.ntt = ((g.nx + 31) / 32) * ((g.ny + 31) / 32);
    const int cnt[KA_COUNT] = {A.nb * 8, A.nb + A.ntt * NTR, A.nb, A.nbx + nby, A.ntt * 9, 2 * A.nb, A.nb * 8};
    for (int k = 0; k < KA_COUNT; k++) off[(size_t)k * (n + 1) + i + 1] = off[(size_t)k * (n + 1) + i] + cnt[k];
  }
  for (int k = 0; k < KA_COUNT; k++) d->nblk[k] = off[(size_t)k * (n + 1) + n];
  hipError_t e = hipMalloc(&d->d_inst, sizeof(AdjInst) * (size_t)n);
  if (e == hipSuccess) e = hipMalloc(&d->d_off, sizeof(int) * off.size());
  if (e == hipSuccess)
    e = hipMemcpyAsync(d->d_inst, inst.data(), sizeof(AdjInst) * (size_t)n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d->d_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    if (err) *err = std::string("batched adjoint tables: ") + hipGetErrorName(e);
    adjb_destroy(d);
    return e == hipErrorOutOfMemory ? QAS_E_NOMEM : QAS_E_HIP;
  }
  *out = d;
  return QAS_OK;
}

static const int* aboff(const AdjBatchDev* d, int k) { return d->d_off + (size_t)k * (d->n + 1); }

hipError_t adjb_fold_final(AdjBatchDev* d, hipStream_t st)
{
  hipLaunchKernelGGL(kba_fold, (unsigned)d->nblk[KA_FOLD], 256, 0, st, d->d_inst, aboff(d, KA_FOLD), d->n);
  return hipGetLastError();
}

hipError_t adjb_stage(AdjBatchDev* d, int rk, bool last, hipStream_t st)
{
  const AdjInst* I = d->d_inst;
  const int n = d->n;
  if (d->ctw) hipLaunchKernelGGL(kba_corner_tr<1>, (unsigned)d->nblk[KA_CORNER], 256, 0, st, I, aboff(d, KA_CORNER), n, rk);
  else hipLaunchKernelGGL(kba_corner_tr<0>, (unsigned)d->nblk[KA_CORNER], 256, 0, st, I, aboff(d, KA_CORNER), n, rk);
  hipLaunchKernelGGL(kba_faceparts, (unsigned)d->nblk[KA_FACE], 256, 0, st, I, aboff(d, KA_FACE), n, rk);
  if (d->rmn == 3)
    hipLaunchKernelGGL(kba_sweep_xy<3>, (unsigned)d->nblk[KA_SWEEP], 256, 0, st, I, aboff(d, KA_SWEEP), n, rk, d->gam);
  else if (d->rmn == 2)
    hipLaunchKernelGGL(kba_sweep_xy<2>, (unsigned)d->nblk[KA_SWEEP], 256, 0, st, I, aboff(d, KA_SWEEP), n, rk, d->gam);
  else
    hipLaunchKernelGGL(kba_sweep_xy<1>, (unsigned)d->nblk[KA_SWEEP], 256, 0, st, I, aboff(d, KA_SWEEP), n, rk, d->gam);
  hipLaunchKernelGGL(kba_untranspose, (unsigned)d->nblk[KA_UNTR], 256, 0, st, I, aboff(d, KA_UNTR), n);
  const bool dv = d->rmn == 2 || d->rmn == 3;
  const unsigned ng = (unsigned)d->nblk[KA_GIN];
  const int* og = aboff(d, KA_GIN);
  if (d->ctw && dv) hipLaunchKernelGGL((kba_gin<1, true>), ng, 256, 0, st, I, og, n, rk, d->gam);
  else if (d->ctw) hipLaunchKernelGGL((kba_gin<1, false>), ng, 256, 0, st, I, og, n, rk, d->gam);
  else if (dv) hipLaunchKernelGGL((kba_gin<0, true>), ng, 256, 0, st, I, og, n, rk, d->gam);
  else hipLaunchKernelGGL((kba_gin<0, false>), ng, 256, 0, st, I, og, n, rk, d->gam);
  hipLaunchKernelGGL(kba_foldstage, (unsigned)d->nblk[KA_FSTAGE], 256, 0, st, I, aboff(d, KA_FSTAGE), n, rk,
                     rk == 2 ? 1 : 0, last ? 1 : 0);
  return hipGetLastError();
}
}  // namespace qas

extern "C" int qas_adj_keep_stages(qas_ctx* c, int on)
{
  if (!c) return QAS_E_ARG;
  c->adj_keep = on != 0;
  return QAS_OK;
}

extern "C" int qas_adj_forward(qas_ctx* c, int nsteps, const double* dts)
{
  int r = adj_prepare(c, nsteps, dts);
  if (r) return r;
  qas_adj_work* w = c->adj;
  const size_t per = 8 * w->stride;
  if (c->adj_keep && nsteps > w->ks_cap) {
    if (w->ks) QAS_ACHECK(c, hipFree(w->ks));
    w->ks = nullptr;
    w->ks_cap = 0;
    hipError_t e = hipMalloc(&w->ks, (size_t)KEPT * w->stride * (size_t)nsteps * sizeof(double));
    if (e != hipSuccess) return afail(c, QAS_E_NOMEM, std::string("adjoint kept stages: ") + hipGetErrorName(e));
    w->ks_cap = nsteps;
  }
  if (!c->bounded) {
    QAS_ACHECK(c, launch_bound(c, c->U[c->cur], 8));
    c->bounded = true;
  }
  const size_t bytes = (size_t)c->geo.nd * sizeof(double);
  for (int n = 0; n < nsteps; n++) {
    double* U[8];
    for (int m = 0; m < 8; m++) {
      U[m] = w->ck + per * n + w->stride * m;
      QAS_ACHECK(c, hipMemcpyAsync(U[m], c->U[c->cur][m], bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    if (!c->adj_keep) {
      r = qas_step(c, dts[n]);
      if (r) return r;
      continue;
    }
    // ideal(dt) from the checkpoint, the stage intermediates written into
    // this step's entries of the store and the step's result into the
    // current state buffer (the kernels and buffers order of ctx_step)
    w->cur_dt = w->d_dts + n;
    double *S0[8], *S1[8];
    for (int m = 0; m < 8; m++) {
      S0[m] = kept_arr(w, n, m);
      S1[m] = kept_arr(w, n, 8 + m);
    }
    QAS_ACHECK(c, fwd_stage(c, kept_slot(w, n, 0), c->ez, U, U, S0, 0, true));
    QAS_ACHECK(c, fwd_stage(c, kept_slot(w, n, 1), c->ez, S0, U, S1, 1, true));
    QAS_ACHECK(c, fwd_stage(c, kept_slot(w, n, 2), c->ez, S1, U, c->U[c->cur], 2, true));
  }
  if (c->adj_keep) {
    // qas_step leaves the last dt in the context's device dt
    QAS_ACHECK(c, launch_set_dt(c, dts[nsteps - 1]));
    c->bounded = true;
    w->kept = true;
  }
  w->nrec = nsteps;
  QAS_ACHECK(c, hipStreamSynchronize(c->stream));
  return QAS_OK;
}

extern "C" int qas_adj_backward_device(qas_ctx* c, const double* g_final, double* g_init)
{
  if (!c || !c->slab || !g_final || !g_init) return afail(c, QAS_E_ARG, "qas_adj_backward: bad arguments");
  if (!c->adj || c->adj->nrec == 0) return afail(c, QAS_E_STATE, "qas_adj_backward: no steps recorded");
  qas_adj_work* w = c->adj;
  const Geo& g = c->geo;
  const size_t per = 8 * w->stride;
  const unsigned nb = nblk(g.nd, 256);
  // the final state came out of bound(): fold the given adjoint onto kept cells
  {
    FoldArgs f;
    for (int m = 0; m < 8; m++) {
      f.src[m] = g_final + (size_t)m * g.nd;
      f.dst[m] = w->gO[m];
      f.extra[m] = nullptr;
    }
    f.bc = c->bc;
    hipLaunchKernelGGL(k_adj_fold, dim3(nb, 8), 256, 0, c->stream, f, g);
    QAS_ACHECK(c, hipGetLastError());
  }
  for (int n = w->nrec - 1; n >= 0; n--) {
    w->cur_dt = w->d_dts + n;  // the step's dt, read by the recompute and the stage adjoints
    double* U[8];
    for (int m = 0; m < 8; m++) U[m] = w->ck + per * n + w->stride * m;
    if (w->kept) {  // the recorded forward kept this step's stages
      double *S0[8], *S1[8];
      for (int m = 0; m < 8; m++) {
        S0[m] = kept_arr(w, n, m);
        S1[m] = kept_arr(w, n, 8 + m);
      }
      QAS_ACHECK(c, adj_stage(c, 2, kept_slot(w, n, 2), S1, false));
      QAS_ACHECK(c, adj_stage(c, 1, kept_slot(w, n, 1), S0, false));
      QAS_ACHECK(c, adj_stage(c, 0, kept_slot(w, n, 0), U, true));
      continue;
    }
    // forward recompute: stage inputs S1, S2 and every stage's intermediates
    const StageSlot s0 = work_slot(w, 0), s1 = work_slot(w, 1), s2 = work_slot(w, 2);
    QAS_ACHECK(c, fwd_stage(c, s0, w->ez[0], U, U, w->S[0], 0, false));
    QAS_ACHECK(c, fwd_stage(c, s1, w->ez[1], w->S[0], U, w->S[1], 1, false));
    QAS_ACHECK(c, fwd_stage(c, s2, w->ez[2], w->S[1], U, w->S[2], 2, false));
    QAS_ACHECK(c, adj_stage(c, 2, s2, w->S[1], false));
    QAS_ACHECK(c, adj_stage(c, 1, s1, w->S[0], false));
    QAS_ACHECK(c, adj_stage(c, 0, s0, U, true));
  }
  for (int m = 0; m < 8; m++)
    QAS_ACHECK(c, hipMemcpyAsync(g_init + (size_t)m * g.nd, w->gO[m], (size_t)g.nd * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
  QAS_ACHECK(c, hipStreamSynchronize(c->stream));
  return QAS_OK;
}

extern "C" int qas_adj_backward(qas_ctx* c, const double* g_final, double* g_init)
{
  if (!c || !c->slab || !g_final || !g_init) return afail(c, QAS_E_ARG, "qas_adj_backward: bad arguments");
  if (!c->adj || c->adj->nrec == 0) return afail(c, QAS_E_STATE, "qas_adj_backward: no steps recorded");
  qas_adj_work* w = c->adj;
  const size_t bytes = 8 * (size_t)c->geo.nd * sizeof(double);
  if (!w->stage) {  // device staging of the two gradients, kept with the workspace
    hipError_t e = hipMalloc(&w->stage, 2 * bytes);
    if (e != hipSuccess) return afail(c, QAS_E_NOMEM, std::string("qas_adj_backward: staging: ") + hipGetErrorName(e));
  }
  double* dg = w->stage;
  double* di = w->stage + 8 * (size_t)c->geo.nd;
  // copies on the context's stream (a null-stream copy would also wait for
  // every other context's work)
  QAS_ACHECK(c, hipMemcpyAsync(dg, g_final, bytes, hipMemcpyHostToDevice, c->stream));
  int r = qas_adj_backward_device(c, dg, di);
  if (r) return r;
  QAS_ACHECK(c, hipMemcpyAsync(g_init, di, bytes, hipMemcpyDeviceToHost, c->stream));
  QAS_ACHECK(c, hipStreamSynchronize(c->stream));
  return QAS_OK;
}
