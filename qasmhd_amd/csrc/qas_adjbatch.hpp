// qas_adjbatch.hpp -- interface between the batched domains (qas_batch.hip)
// and the discrete adjoint (qas_adjoint2d.hip) for the batched adjoint
// (qas_batch_adj_forward / qas_batch_adj_backward).
#pragma once
#include <string>

#include "qas_internal.hpp"

namespace qas {

// One context's adjoint buffers (allocated by adj_prepare).
struct AdjBufs {
  double* ck = nullptr;      // checkpoints: step n, variable m at ck + (8 n + m) * stride
  size_t stride = 0;
  double* d_dts = nullptr;   // the recorded step sizes (device)
  double* dslot = nullptr;   // the step size the recompute and the stage adjoints read
  double* S[3][8] = {};      // recompute: stage inputs S1, S2 and a discard output
  double* prim[3][6] = {};   // per stage: vx vy vz pr cx cy
  double* emf[3][8] = {};    // per stage: face parts
  double* ez[3] = {};
  double* gO[8] = {};        // adjoint of the current stage output (the result, after the last stage)
  double* stage = nullptr;   // device staging of g_final (first 8 nd) and g_init (next 8 nd)
};

// "" when the context is in the adjoint's scope, else the reason
std::string adj_scope(const qas_ctx* c);
// workspace, checkpoint room for nsteps steps, the step sizes on the device
// and the gradient staging; no step recorded yet
int adj_prepare(qas_ctx* c, int nsteps, const double* dts);
void adj_bufs(const qas_ctx* c, AdjBufs* b);
void adj_set_recorded(qas_ctx* c, int nsteps);
int adj_recorded(const qas_ctx* c);

// Batched stage adjoints: device tables for the contexts in the given
// order (the batch's block order); ub[8 i + m] is variable m of the buffer
// holding the input state of the step being reversed of instance i.  Every
// launch covers all instances.
struct AdjBatchDev;
int adjb_create(AdjBatchDev** out, qas_ctx* const* ctxs, double* const* ub, int n, hipStream_t st,
                std::string* err);
void adjb_destroy(AdjBatchDev* d);
// the final-state adjoint (staged in each context's first staging half)
// folded onto kept cells into gO
hipError_t adjb_fold_final(AdjBatchDev* d, hipStream_t st);
// the adjoint of stage rk (2, 1, 0 in that order) of every instance
hipError_t adjb_stage(AdjBatchDev* d, int rk, bool last, hipStream_t st);

}  // namespace qas
