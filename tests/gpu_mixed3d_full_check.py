"""Run by tests/test_gpu_mixed.py in a fresh process (torch must load its
HIP runtime before the library does): the 3D turbulence workload at its
full size and length -- 512^3, HLLD / WCNS-4 / 3D CT / RK3, 200 steps at
one fixed dt -- in mixed precision (FP32 reconstruction, Riemann flux and
edge EMFs; FP64 state, pressure recovery, RK sum and CT update) against the
FP64 path: mean |rho_mixed - rho_fp64| and mean |p_mixed - p_fp64| over the
interior at most 1e-6 (rho ~ 1 and p ~ 1/gam: absolute and relative L1
agree to those factors)."""
import os
import sys

import torch  # first: torch's HIP runtime must be the one the library binds to

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from qasmhd_amd import capi, problems  # noqa: E402

NG = 4
TOL = 1e-6


def run(p, precision, steps, dt, chunk=50):
    out = torch.empty_like(p.state)
    with capi.solver_for(p, precision=precision) as s:
        torch.cuda.synchronize()
        s.upload_device(p.state.data_ptr())
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            s.run(k, flg=0, dt=dt, history=False)
            done += k
            print("  precision %d: %d steps" % (precision, done), flush=True)
        s.download_device(out.data_ptr())
        s.sync()
    return out


def pressure(st, gam):
    """Pressure with the 2nd-order cell-centred B (as setdt)."""
    ro = st[0]
    v2 = (st[1] ** 2 + st[2] ** 2 + st[3] ** 2) / ro
    bx = 0.5 * (st[4] + torch.roll(st[4], -1, dims=2))
    by = 0.5 * (st[5] + torch.roll(st[5], -1, dims=1))
    bz = 0.5 * (st[6] + torch.roll(st[6], -1, dims=0))
    return (gam - 1) * (st[7] - 0.5 * (v2 + bx ** 2 + by ** 2 + bz ** 2))


def main(n=512, steps=200):
    p = problems.turbulence3d(n=n, device="cuda:0")
    with capi.solver_for(p) as s:
        torch.cuda.synchronize()
        s.upload_device(p.state.data_ptr())
        dt = s.cfl_dt()
    print("3d %d^3, %d steps at dt %.6g: fp64 vs mixed" % (n, steps, dt), flush=True)
    a = run(p, 0, steps, dt)
    b = run(p, 1, steps, dt)
    assert bool(torch.isfinite(b).all())
    I = slice(NG, -NG)
    e_ro = (b[0, I, I, I] - a[0, I, I, I]).abs().mean().item()
    pa = pressure(a, p.gam)[I, I, I]
    pb = pressure(b, p.gam)[I, I, I]
    e_pr = (pb - pa).abs().mean().item()
    print("  L1(rho) %.3e  L1(p) %.3e" % (e_ro, e_pr), flush=True)
    assert e_ro <= TOL and e_pr <= TOL, (e_ro, e_pr)
    print("OK", flush=True)


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
