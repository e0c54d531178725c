"""Dissipative step on the GPU (qas_dsptv, qas_run with qas_set_dissipation;
qasmhd_amd/csrc/qas_dissipative.hip) against the oracle's restatement of
DMHD2D::dsptv and DMHD2D::exec_ (2D/SERIAL/MRX/dmhd2d_solve.cpp:3-151,
dmhd2d_class.cpp:15-33), which tests/test_oracle_dsptv.py pins bit for bit
to the reference's own code.  FP64 parity is array equality."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mrx(nx=24, ny=20, seed=0, rev=20.0, rem=10.0):
    from qasmhd_amd import problems
    p, nu, eta, dcmax = problems.reconnection2d(nx=nx, ny=ny, xlen=16.0, rev=rev, rem=rem)
    rng = np.random.default_rng(seed)
    ro = p.state[0]
    v = 0.05 * rng.standard_normal((3,) + ro.shape)
    p.state[1:4] = ro * v
    p.state[7] += 0.5 * ro * (v ** 2).sum(axis=0)
    return p, nu, eta, dcmax


@pytest.mark.parametrize("odr,r_k", [(4, 3), (2, 3), (2, 2), (4, 1)])
def test_dsptv_vs_oracle(oracle, odr, r_k):
    from qasmhd_amd import capi
    _say("case: dsptv vs oracle, MRX 24x20, odr", odr, "r_k", r_k)
    p, nu, eta, _ = _mrx()
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=p.rmn, odr=odr, r_k=r_k, dn=p.dn)
    ref = p.state.copy()
    for dt in (0.01, 0.013, 0.007):
        oracle.bound(cfg, ref)
        oracle.dsptv(cfg, ref, nu, eta, dt)
    with capi.solver_for(p, odr=odr, r_k=r_k) as s:
        s.upload(p.state)
        s.set_dissipation(nu, eta, 0.0)
        for dt in (0.01, 0.013, 0.007):
            s.dsptv(dt)
        out = s.download()
    assert np.array_equal(out, ref)


def test_dsptv_periodic_turbulence(oracle):
    from qasmhd_amd import capi, problems
    _say("case: dsptv vs oracle, periodic 32^2, uniform nu, eta")
    p = problems.turbulence2d(n=32, seed=6, vrms=0.5, brms=0.5)
    nu = np.full(p.state.shape[1:], 0.01)
    eta = np.linspace(0.0, 0.02, p.state[0].size).reshape(p.state.shape[1:])
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam)
    ref = p.state.copy()
    for _ in range(4):
        oracle.bound(cfg, ref)
        oracle.dsptv(cfg, ref, nu, eta, 0.002)
    with capi.solver_for(p) as s:
        s.upload(p.state)
        s.set_dissipation(nu, eta, 0.0)
        for _ in range(4):
            s.dsptv(0.002)
        out = s.download()
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("flg", [1, 0])
def test_dissipative_run_vs_oracle(oracle, flg):
    """qas_run with the dissipative step = DMHD2D::exec_ body: ideal(dt),
    sub-cycled dsptv, setdt every 2nd step."""
    from qasmhd_amd import capi
    _say("case: dissipative run vs oracle, MRX 32x24, flg", flg)
    p, nu, eta, dcmax = _mrx(nx=32, ny=24, seed=2, rev=0.5, rem=1.0)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=p.rmn, odr=p.odr, r_k=p.r_k, dn=p.dn)
    ref = p.state.copy()
    dt0 = oracle.cfl_dt(cfg, ref)
    hist_ref = oracle.run_dissipative(cfg, ref, nu, eta, dcmax, 6, flg=flg, dt=dt0)
    ncmax = 1 + int(2 * (dcmax * dt0) / (min(p.dx, p.dy) ** 2))
    assert ncmax >= 2
    with capi.solver_for(p) as s:
        s.upload(p.state)
        s.set_dissipation(nu, eta, dcmax)
        hist, _ = s.run(6, flg=flg, dt=dt0)
        out = s.download()
    assert np.array_equal(hist, hist_ref)
    assert np.array_equal(out, ref)


def test_dissipation_errors():
    from qasmhd_amd import capi, problems
    _say("case: dissipation error paths")
    p = problems.turbulence2d(n=16, seed=1)
    with capi.solver_for(p) as s:
        s.upload(p.state)
        with pytest.raises(capi.QasError):
            s.dsptv(0.01)  # no nu / eta set
        nu = np.zeros(p.state.shape[1:])
        s.set_dissipation(nu, nu, 0.0)
        with pytest.raises(capi.QasError):
            s.dsptv(-1.0)
        s.set_dissipation(None, None, 0.0)
        with pytest.raises(capi.QasError):
            s.dsptv(0.01)
    q = problems.turbulence3d(n=8, seed=1, nmodes=4)
    with capi.solver_for(q) as s:
        z = np.zeros(q.state.shape[1:])
        with pytest.raises(capi.QasError):
            s.set_dissipation(z, z, 0.0)
