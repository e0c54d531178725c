// qas_batch.hpp -- device side of batched 2D domains (qas_batch_*): many
// independent domains of uneven sizes advanced by single launches.
//
// Each launch of a batched kernel covers every instance: instance i owns
// the blocks [off[i], off[i+1]) of the grid (a per-kernel-type prefix sum
// computed on the host), finds itself by a binary search over off[] on its
// (uniform) workgroup index, and runs the single-domain body with its own
// descriptor and local block index.
#pragma once
#include "qas_internal.hpp"

namespace qas {

// kernel types with their own block-offset table
enum BatchKernel { BK_SX = 0, BK_SY, BK_EMF, BK_CT, BK_BOUND, BK_CFL, BK_COUNT };

struct BatchInst {
  StagePtrs st[3];   // stage descriptors of one step (buffer rotation, rk_fac, dt)
  Geo g;
  BcFlags bc;
  double* cur[8];    // the state at the start of a step (and after it, R_K 2 and 3)
  double* cflpart;   // per-block CFL maxima
  double* dt;        // this instance's device dt
  double* hist;      // dt history (qas_batch_run with a history)
  long long* hcount;
  long long hcap;
  long long nghost;  // ghost cells per array
  double cfl_dr;     // cfl * min(dx, dy)
  int nsx, nsy;      // sweep segments along x and y
  int cflblocks;     // blocks of this instance's CFL reduction
  int pad;
};

// instance owning block b: off[0..n] ascending, off[0] = 0, off[n] = grid
__device__ __forceinline__ int batch_find(const int* off, int n, int b)
{
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

typedef void (*batch_fn)(const BatchInst*, const int*, int, int, double);
// the 2D sweeps: (stage descriptor, geometry, boundary flags, gam, segments)
// for one domain, (instance table, block offsets, instances, stage) for a batch
typedef void (*xsweep2d_fn)(StagePtrs, Geo, BcFlags, double, int, const BatchInst*, const int*, int, int);

// kernel tables (instantiated in qas_kernels2d.hip for FP64, in
// qas_mixed.hip for the mixed path and in qas_fma64.hip for contracted FP64)
xsweep2d_fn pick_ysweep2d(int rmn, int odr, bool batch);
xsweep2d_fn pick_ysweep2d_f32(int rmn, int odr, bool batch);
xsweep2d_fn pick_xsweep2d(int rmn, int odr, bool batch);
xsweep2d_fn pick_xsweep2d_f32(int rmn, int odr, bool batch);
// FP64 with FMA contraction (qas_fma64.hip)
xsweep2d_fn pick_ysweep2d_fma(int rmn, int odr, bool batch);
xsweep2d_fn pick_xsweep2d_fma(int rmn, int odr, bool batch);
// gravity instantiations (qas_gravity.hip)
xsweep2d_fn pick_grav_xsweep(int rmn, int odr);
xsweep2d_fn pick_grav_ysweep(int rmn, int odr);

}  // namespace qas
