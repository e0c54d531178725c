"""2D step with gravity on the GPU (qas_set_gravity; qasmhd_amd/csrc/
qas_gravity.hip and the gravity terms of the prim / sweep / bound / CFL
kernels) against the oracle's restatement of the reference's RTI driver
(2D/SERIAL/RTI/mhd2d_{solve,class}.cpp), which tests/test_oracle_gravity.py
pins bit for bit to the reference's own code.  Array equality, dt history
included."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("rmn,odr,r_k,flg", [(2, 4, 3, 1), (1, 4, 3, 1), (1, 2, 2, 1), (3, 4, 3, 0)])
def test_gravity_run_vs_oracle(oracle, rmn, odr, r_k, flg):
    from qasmhd_amd import capi, problems
    _say("case: gravity run vs oracle, RTI 32x48, rmn", rmn, "odr", odr, "r_k", r_k, "flg", flg)
    p, phi = problems.rayleigh_taylor2d(nx=32, ny=48, dv=0.05)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=rmn, odr=odr, r_k=r_k, dn=p.dn, phi=phi)
    ref = p.state.copy()
    nsteps = 8
    dt0 = None if flg else 0.002
    h_ref, _ = oracle.run(cfg, ref, nsteps, flg, dt=dt0)
    with capi.solver_for(p, rmn=rmn, odr=odr, r_k=r_k) as s:
        s.upload(p.state)
        s.set_gravity(phi)
        hist, _ = s.run(nsteps, flg=flg, dt=dt0)
        out = s.download()
    assert np.array_equal(hist, h_ref)
    assert np.array_equal(out, ref)


def test_gravity_steps_and_graph_replay(oracle):
    """qas_step / qas_bound / qas_cfl_dt individually and a long qas_run
    (graph replays of step pairs) against the oracle."""
    from qasmhd_amd import capi, problems
    _say("case: gravity step entry points + graph replay, RTI 24x40")
    p, phi = problems.rayleigh_taylor2d(nx=24, ny=40, dv=0.05)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=p.rmn, odr=p.odr, dn=p.dn, phi=phi)
    ref = p.state.copy()
    oracle.bound(cfg, ref)
    dt = oracle.cfl_dt(cfg, ref)
    oracle.step(cfg, ref, dt)
    oracle.bound(cfg, ref)
    h_ref, _ = oracle.run(cfg, ref, 20, 1, dt=dt)
    with capi.solver_for(p) as s:
        s.lib.qas_use_graphs(s.ctx, 1)
        s.upload(p.state)
        s.set_gravity(phi)
        s.bound()
        d = s.cfl_dt()
        s.step(d)
        s.bound()
        hist, _ = s.run(20, flg=1, dt=d)
        out = s.download()
    assert d == dt
    assert np.array_equal(hist, h_ref)
    assert np.array_equal(out, ref)


def test_gravity_off_again_and_errors():
    from qasmhd_amd import capi, problems
    _say("case: gravity switched off, error paths")
    p, phi = problems.rayleigh_taylor2d(nx=16, ny=24, dv=0.05)
    q = p.copy()
    with capi.solver_for(p) as s, capi.solver_for(q) as t:
        s.upload(p.state)
        t.upload(p.state)
        s.set_gravity(phi)
        s.set_gravity(None)
        s.run(4, flg=1)
        t.run(4, flg=1)
        assert np.array_equal(s.download(), t.download())
        s.set_gravity(phi)
        with pytest.raises(capi.QasError):
            capi.Batch([s, t])
    r = problems.turbulence3d(n=8, seed=1, nmodes=4)
    with capi.solver_for(r) as s:
        with pytest.raises(capi.QasError):
            s.set_gravity(np.zeros(r.state.shape[1:]))
