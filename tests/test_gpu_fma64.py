"""FP64 with fused multiply-adds in the sweeps (qas_scheme.precision = 2,
qasmhd_amd/csrc/qas_fma64.hip) against the bit-exact path.

Bar at the small sizes (30 steps 2D, 8 steps 3D, against the oracle):
relative L1 distance of each of the 8 variables over the interior at most
1e-12, dt history (setdt schedule) within 1e-12 relative.  At the workloads'
full sizes and lengths the flow amplifies round-off beyond 1e-12 whatever
changes the rounding; there the bar is the bit-exact path's own response to
a one-ulp change of the initial state (see the tests), and div B at
round-off.  The distance of a vector component
(m_x, m_y, m_z; B_x, B_y, B_z) is taken against the mean L1 norm of the
three components of its vector, sum|u_c - u_ref,c| / (sum_k sum|u_ref,k| / 3):
a component that is zero in exact arithmetic (m_z, B_z of a 2D set-up) holds
only round-off, whose distance from other round-off is no measure.  The
bit-exact path (precision 0) equals the reference's own driver bit for bit
(tests/test_gpu_parity.py, tests/test_gpu_fullsize.py), and the oracle at
the small sizes; so these bound the distance to the reference itself.
Full sizes: the 2D decaying-turbulence workload (1024^2, HLLD, 1000 steps),
the low-Mach Kelvin-Helmholtz workload (2048^2, MLAU, 2000 steps) and the
3D turbulence workload (512^3, HLLD, 200 steps, in a fresh process)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


GROUPS = ((0,), (1, 2, 3), (1, 2, 3), (1, 2, 3), (4, 5, 6), (4, 5, 6), (4, 5, 6), (7,))


def l1rel_vars(a, b, interior):
    norm = [float(np.abs(b[m][interior]).sum()) for m in range(8)]
    return [float(np.abs(a[m][interior] - b[m][interior]).sum()) / (sum(norm[k] for k in GROUPS[m]) / len(GROUPS[m]))
            for m in range(8)]


def _run(p, steps, precision, flg=1):
    from qasmhd_amd import capi
    with capi.solver_for(p, precision=precision) as s:
        s.upload(p.state)
        hist, dt = s.run(steps, flg=flg)
        return s.download(), hist


def _divb2d(st, dx, dy):
    """4th-order cell divergence of the face B (the difference the CT update
    curls with, so it stays at round-off)."""
    bx, by = st[4], st[5]
    J, I = slice(4, -4), slice(4, -4)
    dbx = (27.0 * (bx[J, 5:-3] - bx[J, 4:-4]) - (bx[J, 6:-2] - bx[J, 3:-5])) / 24.0 / dx
    dby = (27.0 * (by[5:-3, I] - by[4:-4, I]) - (by[6:-2, I] - by[3:-5, I])) / 24.0 / dy
    return float(np.abs(dbx + dby).max())


@pytest.mark.parametrize("rmn,odr,bc", [(1, 4, "periodic"), (3, 4, "periodic"), (2, 4, "walls"), (0, 4, "walls"),
                                        (1, 3, "periodic"), (1, 2, "periodic")])
def test_fma64_2d_vs_oracle(oracle, rmn, odr, bc):
    from qasmhd_amd import problems
    _say("case: fma64 2D 48x40 vs oracle rmn", rmn, "odr", odr, bc)
    p = problems.turbulence2d(n=48, ny=40, seed=31, vrms=0.6)
    if bc == "walls":  # open along x, mirror along y (vy, by odd)
        dn = np.zeros((3, 8), dtype=int)
        dn[0] = [2] * 8
        dn[1] = [1, 1, -1, 1, 1, -1, 1, 1]
        p.dn = dn
    p.rmn, p.odr = rmn, odr
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=rmn, odr=odr, dn=p.dn)
    ref = p.state.copy()
    h_ref, _ = oracle.run(cfg, ref, 30, 1)
    out, hist = _run(p, 30, 2)
    I = (slice(4, -4), slice(4, -4))
    e = l1rel_vars(out, ref, I)
    _say("  rel L1 per variable: " + " ".join("%.1e" % x for x in e))
    assert max(e) <= TOL
    assert np.abs(hist / h_ref - 1).max() <= TOL
    assert not np.array_equal(out, ref)  # contraction is on: not the reference's rounding


def test_fma64_3d_vs_oracle(oracle):
    from qasmhd_amd import problems
    _say("case: fma64 3D 16^3 vs oracle")
    p = problems.turbulence3d(n=16, seed=7)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, p.dz, gam=p.gam, rmn=p.rmn, odr=p.odr)
    ref = p.state.copy()
    h_ref, _ = oracle.run(cfg, ref, 8, 1)
    out, hist = _run(p, 8, 2)
    I = (slice(4, -4), slice(4, -4), slice(4, -4))
    e = l1rel_vars(out, ref, I)
    _say("  rel L1 per variable: " + " ".join("%.1e" % x for x in e))
    assert max(e) <= TOL
    assert np.abs(hist / h_ref - 1).max() <= TOL


@pytest.mark.parametrize("name", ["turb2d_1024", "khi_2048_mlau"])
def test_fma64_full_size_vs_exact(name):
    """At the workloads' own sizes and lengths the flow amplifies round-off:
    the distance is set against the bit-exact path's own response to a
    one-ulp change of the initial density (rho (1 + 2^-52) in every cell).
    The contracted path must stay within 10x of that round-off response.
    A 1e-12 relative L1 bar is met only by the bit-exact path here: measured
    one-ulp responses are 6e-11 (1024^2 turbulence, 1000 steps) and 1.7e-6
    (Kelvin-Helmholtz at Mach 0.01, 2000 steps), and the contracted path
    lands at 2.4x and 1.0x of them."""
    from qasmhd_amd import problems
    if name == "turb2d_1024":
        p = problems.turbulence2d(n=1024, seed=2)
        steps = 1000
    else:
        p = problems.kelvin_helmholtz(n=2048)
        steps = 2000
    _say("case: fma64 vs exact,", name, steps, "steps")
    a, ha = _run(p, steps, 0)
    b, hb = _run(p, steps, 2)
    q = p.copy()
    q.state = p.state.copy()
    q.state[0] = q.state[0] * (1.0 + 2.0 ** -52)
    c, hc = _run(q, steps, 0)
    I = (slice(4, -4), slice(4, -4))
    e = l1rel_vars(b, a, I)
    u = l1rel_vars(c, a, I)
    _say("  fma  rel L1 per variable: " + " ".join("%.2e" % x for x in e))
    _say("  1ulp rel L1 per variable: " + " ".join("%.2e" % x for x in u))
    _say("  dt history max rel diff: fma %.2e, 1ulp %.2e" % (np.abs(hb / ha - 1).max(), np.abs(hc / ha - 1).max()))
    assert np.isfinite(b).all()
    assert max(e) <= 10 * max(u)
    assert np.abs(hb / ha - 1).max() <= 10 * np.abs(hc / ha - 1).max()
    db = _divb2d(b, p.dx, p.dy)
    scale = np.abs(b[4:6]).max() / min(p.dx, p.dy)
    _say("  max |div B| %.2e (B/dx %.2e)" % (db, scale))
    assert db <= 1e-12 * scale


def test_fma64_3d_full_size_vs_exact():
    """512^3 over 200 steps in a fresh process (tests/gpu_fma3d_full_check.py)."""
    _say("case: fma64 vs exact 3D 512^3, 200 steps (subprocess)")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_fma3d_full_check.py")
    r = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True, timeout=280)
    _say(r.stdout[-3000:])
    if r.returncode != 0:
        _say(r.stderr[-3000:])
    assert r.returncode == 0
    assert "OK" in r.stdout
