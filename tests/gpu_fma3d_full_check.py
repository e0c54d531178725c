"""Run by tests/test_gpu_fma64.py in a fresh process (torch must load its
HIP runtime before the library does): the 3D turbulence workload at its
full size and length -- 512^3, HLLD / WCNS-4 / 3D CT / RK3, 200 steps with
the setdt schedule -- with the contracted FP64 sweeps (precision 2) against
the bit-exact path (precision 0), set against the bit-exact path's own
response to a one-ulp change of the initial density: relative L1 distance
of every variable over the interior (a vector component against the mean
L1 norm of its vector's components) and of the final dt at most 10x that
response."""
import os
import sys

import torch  # first: torch's HIP runtime must be the one the library binds to

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from qasmhd_amd import capi, problems  # noqa: E402

NG = 4


def run(p, precision, steps, chunk):
    out = torch.empty_like(p.state)
    with capi.solver_for(p, precision=precision) as s:
        torch.cuda.synchronize()
        s.upload_device(p.state.data_ptr())
        done, dt = 0, None
        while done < steps:
            k = min(chunk, steps - done)
            _, dt = s.run(k, flg=1, history=False)
            done += k
            print("  precision %d: %d steps" % (precision, done), flush=True)
        s.download_device(out.data_ptr())
        s.sync()
    return out, dt


def dist(a, b):
    """Relative L1 distance per variable over the interior; a vector
    component against the mean L1 norm of its vector's components."""
    I = slice(NG, -NG)
    groups = ((0,), (1, 2, 3), (1, 2, 3), (1, 2, 3), (4, 5, 6), (4, 5, 6), (4, 5, 6), (7,))
    norm = [a[m, I, I, I].abs().sum().item() for m in range(8)]
    return [(b[m, I, I, I] - a[m, I, I, I]).abs().sum().item() / (sum(norm[k] for k in groups[m]) / len(groups[m]))
            for m in range(8)]


def main(n=512, steps=200, chunk=50):
    p = problems.turbulence3d(n=n, device="cuda:0")
    print("3d %d^3, %d steps: exact vs fma64 vs exact from a one-ulp change of rho" % (n, steps), flush=True)
    a, dta = run(p, 0, steps, chunk)
    b, dtb = run(p, 2, steps, chunk)
    assert bool(torch.isfinite(b).all())
    e = dist(a, b)
    del b
    p.state[0].mul_(1.0 + 2.0 ** -52)
    c, dtc = run(p, 0, steps, chunk)
    u = dist(a, c)
    print("  fma  rel L1 per variable: " + " ".join("%.2e" % x for x in e), flush=True)
    print("  1ulp rel L1 per variable: " + " ".join("%.2e" % x for x in u), flush=True)
    print("  final dt exact %.17g fma %.17g 1ulp %.17g" % (dta, dtb, dtc), flush=True)
    assert max(e) <= 10 * max(u), (e, u)
    assert abs(dtb / dta - 1) <= 10 * abs(dtc / dta - 1) + 1e-15
    print("OK", flush=True)


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
