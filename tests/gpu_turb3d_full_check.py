"""Run by tests/test_gpu_parity.py in a fresh process (torch must load its
HIP runtime before the library does): the 3D turbulence workload at its
full size and length -- 512^3, HLLD / WCNS-4 / 3D CT / RK3, 200 steps with
the setdt schedule -- checked through size-independent properties on the
device: the periodic totals of rho, m and E stay at round-off, and the
4th-order cell divergence of the face B stays at round-off (CT), with the
state finite throughout.

Forward: MHD3D::ideal (3D/MPI/OTvortex/mhd3d_solve.cpp:3-729), bit-exact
to the oracle on small grids (test_3d_vs_oracle)."""
import os
import sys

import torch  # first: torch's HIP runtime must be the one the library binds to

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from qasmhd_amd import capi, problems  # noqa: E402

NG = 4


def divb(st, d):
    I = slice(NG, -NG)

    def df(f, axis):
        m = f.shape[axis] - 2 * NG

        def sl(a):
            return f.narrow(axis, a, m)
        return (27.0 * (sl(NG + 1) - sl(NG)) - (sl(NG + 2) - sl(NG - 1))) / 24.0 / d
    bx, by, bz = st[4], st[5], st[6]
    return (df(bx, 2)[I, I, :] + df(by, 1)[I, :, I] + df(bz, 0)[:, I, I]).abs().max().item()


def main(n=512, steps=200, chunk=50):
    p = problems.turbulence3d(n=n, device="cuda:0")
    st0 = p.state
    d = p.dx
    I = slice(NG, -NG)
    tot0 = st0[:, I, I, I].sum(dim=(1, 2, 3))
    scale = st0[:, I, I, I].abs().sum(dim=(1, 2, 3)) + 1.0
    d0 = divb(st0, d)
    bmax = st0[4:7].abs().max().item()
    print("3d %d^3: initial div B %.3e (B max %.3f)" % (n, d0, bmax), flush=True)
    out = torch.empty_like(st0)
    with capi.solver_for(p) as s:
        torch.cuda.synchronize()
        s.upload_device(st0.data_ptr())
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            s.run(k, flg=1, history=False)
            s.sync()
            done += k
            print("  %d steps" % done, flush=True)
        s.download_device(out.data_ptr())
        s.sync()
    assert bool(torch.isfinite(out).all())
    tot = out[:, I, I, I].sum(dim=(1, 2, 3))
    rel = [((tot[v] - tot0[v]).abs() / scale[v]).item() for v in (0, 1, 2, 3, 7)]
    print("  total drift / sum|u| (ro mx my mz en):", " ".join("%.1e" % r for r in rel), flush=True)
    assert max(rel) < 1e-12, rel
    d1 = divb(out, d)
    print("  div B after %d steps %.3e" % (steps, d1), flush=True)
    assert d1 < 1e-11 * bmax / d + 10 * d0
    assert not torch.equal(out, st0)
    print("ok", flush=True)


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
