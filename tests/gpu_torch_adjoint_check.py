"""Run by tests/test_gpu_adjoint.py in a fresh process: the
torch.autograd.Function over the C ABI (qasmhd_amd/torch_adjoint.py) gives
the same input gradient as the host adjoint entry points, and a node whose
checkpoints a later forward on the same solver replaced raises in backward."""
import os
import sys

import torch  # first: torch's HIP runtime must be the one the library binds to

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import numpy as np  # noqa: E402

from qasmhd_amd import capi, torch_adjoint  # noqa: E402
from test_gpu_adjoint import _setup  # noqa: E402


def main():
    p, s, dts = _setup(n=24, seed=9)
    u0 = torch.tensor(p.state, device="cuda:0", requires_grad=True)
    target = torch.tensor(p.state[0], device="cuda:0") * 1.01
    out = torch_adjoint.ideal_steps(u0, s, dts[:2])
    loss = ((out[0] - target) ** 2).sum()
    loss.backward()
    W = np.zeros_like(p.state)
    W[0] = 2.0 * (out[0].detach().cpu().numpy() - target.cpu().numpy())
    s.upload(p.state)
    s.adj_forward(dts[:2])
    g = s.adj_backward(W)
    # the forward through torch equals the host forward
    s.upload(p.state)
    s.bound()
    for dt in dts[:2]:
        s.step(dt)
    ref = s.download()
    # two live nodes on one solver: the second forward replaces the first's
    # checkpoints, so the first node's backward must raise, not return a
    # gradient of the other trajectory
    u1 = torch.tensor(p.state, device="cuda:0", requires_grad=True)
    o1 = torch_adjoint.ideal_steps(u1, s, dts[:2])
    o2 = torch_adjoint.ideal_steps(u1 * 1.0, s, dts[:1])
    stale_raised = False
    try:
        torch.autograd.grad(o1[0].sum(), u1, retain_graph=True)
    except capi.QasError:
        stale_raised = True
    g2 = torch.autograd.grad(o2[0].sum(), u1)[0]
    s.close()
    print("stale node raises:", stale_raised, " current node finite:", bool(torch.isfinite(g2).all()), flush=True)
    if not stale_raised or not bool(torch.isfinite(g2).all()):
        sys.exit(1)
    ok_f = np.array_equal(out.detach().cpu().numpy(), ref)
    ok_g = np.allclose(u0.grad.cpu().numpy(), g, rtol=1e-12, atol=1e-14 * np.abs(g).max())
    print("torch forward == host forward:", ok_f, " torch grad == host adjoint:", ok_g, flush=True)
    if not (ok_f and ok_g):
        sys.exit(1)


if __name__ == "__main__":
    main()
