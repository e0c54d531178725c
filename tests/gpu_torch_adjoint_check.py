"""Run by tests/test_gpu_adjoint.py in a fresh process: the
torch.autograd.Function over the C ABI (qasmhd_amd/torch_adjoint.py) gives
the same input gradient as the host adjoint entry points."""
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
    s.close()
    ok_f = np.array_equal(out.detach().cpu().numpy(), ref)
    ok_g = np.allclose(u0.grad.cpu().numpy(), g, rtol=1e-12, atol=1e-14 * np.abs(g).max())
    print("torch forward == host forward:", ok_f, " torch grad == host adjoint:", ok_g, flush=True)
    if not (ok_f and ok_g):
        sys.exit(1)


if __name__ == "__main__":
    main()
