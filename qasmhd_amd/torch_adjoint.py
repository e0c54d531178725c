"""torch.autograd.Function over the C ABI: n steps of ideal(dt) at frozen
step sizes, with the hand-written discrete adjoint as its backward.

    from qasmhd_amd import capi, torch_adjoint
    s = capi.Solver(shape, dx, dy, gam=..., ...)          # 2D, FP64, HLLD, odr 4, RK3, periodic
    u_final = torch_adjoint.ideal_steps(u0, s, dts)        # u0: (8, ny, nx) float64 on the GPU
    loss = ((u_final[0] - target) ** 2).sum(); loss.backward()   # u0.grad

The context runs on its own HIP stream; the wrapper synchronises torch's
stream before handing buffers over and the context's stream before
returning them.  The C++ host API (include/qasmhd.h) stays the primary
interface; this is a thin binding of qas_adj_forward / qas_adj_backward_device.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import capi

# torch ships its own libamdhip64 with the same soname as ROCm's; whichever
# is loaded first serves the process.  torch must come first, or it finds no
# GPU once the library has initialised the other copy.
if "torch" not in sys.modules and capi._lib is not None:
    raise ImportError("qasmhd_amd.torch_adjoint: import torch before loading libqasmhd_amd "
                      "(both bind libamdhip64.so.7; torch's copy must be loaded first)")
import torch  # noqa: E402


class IdealSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, solver, dts):
        if u0.dtype != torch.float64 or not u0.is_cuda:
            raise capi.QasError("ideal_steps: u0 must be a float64 GPU tensor")
        u0c = u0.detach().contiguous()
        if tuple(u0c.shape) != (8,) + solver.shape:
            raise capi.QasError("ideal_steps: u0 has shape %s, the solver %s" % (tuple(u0c.shape), solver.shape))
        torch.cuda.synchronize(u0c.device)
        solver.upload_device(u0c.data_ptr())
        solver.adj_forward(np.asarray(dts, dtype=np.float64))
        out = torch.empty_like(u0c)
        solver.download_device(out.data_ptr())
        solver.sync()
        ctx.solver = solver
        ctx.generation = solver.adj_generation
        return out

    @staticmethod
    def backward(ctx, g_out):
        s = ctx.solver
        if getattr(s, "adj_generation", None) != ctx.generation:
            # the solver keeps the checkpoints of its latest forward only:
            # a later ideal_steps (or adj_forward) on it replaced this node's
            raise capi.QasError("ideal_steps backward: the solver ran another forward since this node's; its "
                                "checkpoints are gone (use one solver per live ideal_steps node)")
        g = g_out.detach().contiguous()
        g_in = torch.empty_like(g)
        torch.cuda.synchronize(g.device)
        s._check(s.lib.qas_adj_backward_device(s.ctx, C.c_void_p(g.data_ptr()), C.c_void_p(g_in.data_ptr())),
                 "qas_adj_backward_device")
        return g_in, None, None


def ideal_steps(u0: torch.Tensor, solver, dts) -> torch.Tensor:
    """len(dts) steps of ideal(dt) from u0 (ghosts included; bound() is
    applied first), differentiable with respect to u0.  The solver holds
    the checkpoints of one forward: a second ideal_steps on the same solver
    before the first node's backward makes that backward raise."""
    return IdealSteps.apply(u0, solver, dts)
