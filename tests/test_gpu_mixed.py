"""Mixed precision (FP32 reconstruction + Riemann flux; FP64 state, pressure
recovery, flux differencing, RK sum, CT) against the FP64 path.  The FP64
path is the parity-checked one; here the L1 error of the mixed path is
measured and bounded."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def l1rel(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def run(p, steps, precision, dt):
    from qasmhd_amd import capi
    with capi.solver_for(p, precision=precision) as s:
        s.upload(p.state)
        # fixed dt: both precisions take identical time steps
        s.run(steps, flg=0, dt=dt, history=False)
        return s.download()


def pressure(st, gam):
    ro = st[0]
    v2 = (st[1] ** 2 + st[2] ** 2 + st[3] ** 2) / ro
    return (gam - 1) * (st[7] - 0.5 * (v2 + st[4] ** 2 + st[5] ** 2 + st[6] ** 2))


@pytest.mark.parametrize("case", ["khi", "turb"])
def test_mixed_l1_error(case):
    from qasmhd_amd import capi, problems
    if case == "khi":
        p = problems.kelvin_helmholtz(n=128)
        steps, tol = 200, 1e-5
    else:
        p = problems.turbulence2d(n=128, seed=3)
        steps, tol = 100, 1e-4
    with capi.solver_for(p) as s:
        s.upload(p.state)
        dt = s.cfl_dt()
    a = run(p, steps, 0, dt)
    b = run(p, steps, 1, dt)
    assert np.isfinite(b).all()
    I = (slice(4, -4), slice(4, -4))
    e_ro = l1rel(b[0][I], a[0][I])
    e_pr = l1rel(pressure(b, p.gam)[I], pressure(a, p.gam)[I])
    _say("mixed vs fp64 %s: L1(rho) %.3e  L1(p) %.3e" % (case, e_ro, e_pr))
    assert e_ro < tol and e_pr < tol


def test_mixed_l1_error_vs_oracle(oracle):
    """The mixed path's L1 error measured against the oracle itself (the
    CPU restatement of the reference) rather than through the FP64 HIP
    path: KH 128^2, MLAU, 200 steps at a fixed dt; the FP64 HIP result is
    the oracle's bit for bit in the same run."""
    from qasmhd_amd import capi, problems
    p = problems.kelvin_helmholtz(n=128)
    with capi.solver_for(p) as s:
        s.upload(p.state)
        dt = s.cfl_dt()
    steps = 200
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, cfl=p.cfl, rmn=p.rmn, odr=p.odr, r_k=p.r_k,
                          ctw=p.ctw, dn=p.dn)
    ref = p.state.copy()
    oracle.run(cfg, ref, steps, 0, dt=dt)
    a = run(p, steps, 0, dt)
    b = run(p, steps, 1, dt)
    assert np.array_equal(a, ref)
    I = (slice(4, -4), slice(4, -4))
    e_ro = l1rel(b[0][I], ref[0][I])
    e_pr = l1rel(pressure(b, p.gam)[I], pressure(ref, p.gam)[I])
    _say("mixed vs oracle khi: L1(rho) %.3e  L1(p) %.3e" % (e_ro, e_pr))
    assert np.isfinite(b).all() and e_ro < 1e-5 and e_pr < 1e-5


def test_mixed_3d_runs(oracle):
    from qasmhd_amd import capi, problems
    p = problems.turbulence3d(n=16, seed=2, nmodes=8)
    with capi.solver_for(p) as s:
        s.upload(p.state)
        dt = s.cfl_dt()
    a = run(p, 4, 0, dt)
    b = run(p, 4, 1, dt)
    e = l1rel(b[0], a[0])
    _say("mixed vs fp64 3d: L1(rho) %.3e" % e)
    assert np.isfinite(b).all() and e < 1e-5


def test_mixed_full_size_khi():
    """The low-Mach Kelvin-Helmholtz workload at its full size and length
    (2048^2, M = 0.01, MLAU / WCNS-4 / CT / RK3, 2000 steps at one fixed dt):
    mean |rho_mixed - rho_fp64| and mean |p_mixed - p_fp64| over the
    interior at most 1e-6 (rho = 1 and p = 1/gam, so absolute and relative
    L1 coincide up to the factor gam for p)."""
    from qasmhd_amd import capi, problems
    p = problems.kelvin_helmholtz(n=2048)
    with capi.solver_for(p) as s:
        s.upload(p.state)
        dt = s.cfl_dt()
    _say("case: mixed vs fp64, KH 2048^2, 2000 steps, fp64 run")
    a = run(p, 2000, 0, dt)
    _say("  mixed run")
    b = run(p, 2000, 1, dt)
    assert np.isfinite(b).all()
    I = (slice(4, -4), slice(4, -4))
    e_ro = float(np.abs(b[0][I] - a[0][I]).mean())
    e_pr = float(np.abs(pressure(b, p.gam)[I] - pressure(a, p.gam)[I]).mean())
    _say("  L1(rho) %.3e  L1(p) %.3e" % (e_ro, e_pr))
    assert e_ro <= 1e-6 and e_pr <= 1e-6


def test_mixed_full_size_3d():
    """The 3D turbulence workload at its full size and length (512^3, 200
    steps at one fixed dt) in mixed precision against FP64: L1(rho) and
    L1(p) at most 1e-6 (tests/gpu_mixed3d_full_check.py, fresh process)."""
    import os
    import subprocess
    _say("case: mixed vs fp64 3D 512^3, 200 steps (subprocess)")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_mixed3d_full_check.py")
    r = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True, timeout=280)
    _say(r.stdout[-3000:])
    if r.returncode != 0:
        _say(r.stderr[-3000:])
    assert r.returncode == 0
    assert "OK" in r.stdout
