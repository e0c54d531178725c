"""The oracle's gravity terms (potential round trips in the primitive
recovery and in setdt, work done by gravity in the energy fluxes, momentum
sources, hydrostatic energy ghosts) against the reference's own 2D driver
with gravity (2D/SERIAL/RTI/mhd2d_{solve,class}.cpp, built from source by
oracle/build.py) on the same states: bit-exact float64, dt history included.
Skipped where /root/reference is absent."""
import numpy as np
import pytest

from oracle import build as obuild

pytestmark = pytest.mark.skipif(not obuild.reference_present(), reason="reference sources not present")


@pytest.mark.parametrize("rmn,odr,r_k,flg", [(2, 4, 3, 1), (1, 2, 2, 1), (2, 4, 3, 0)])
def test_gravity_run_matches_reference(oracle, rmn, odr, r_k, flg):
    from qasmhd_amd import problems
    p, phi = problems.rayleigh_taylor2d(nx=16, ny=24, dv=0.05)
    ny, nx = p.state.shape[1:]
    binary = obuild.build_ref(2, nx - 8, ny - 8, rmn, odr, r_k, 1, 0.4, gravity=True)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=rmn, odr=odr, r_k=r_k, dn=p.dn, phi=phi)
    nsteps = 6
    dt0 = 0.0 if flg else 0.003
    ref, hist_ref = obuild.run_ref(binary, p.state, p.dx, p.gam, nsteps, flg, dt=dt0, dnx=list(p.dn[0]),
                                   dny=list(p.dn[1]), dy=p.dy, phi=phi)
    st = p.state.copy()
    hist, _ = oracle.run(cfg, st, nsteps, flg, dt=None if flg else dt0)
    assert np.array_equal(hist, hist_ref)
    assert np.array_equal(st, ref)
    assert not np.array_equal(st, p.state)


def test_hydrostatic_state_stays_close(oracle):
    """Without the perturbation the set-up is in discrete hydrostatic
    balance to the accuracy of its midpoint-rule pressure: velocities stay
    small, and the energy ghosts follow the hydrostatic extrapolation."""
    from qasmhd_amd import problems
    p, phi = problems.rayleigh_taylor2d(nx=16, ny=32, dv=0.0)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam, rmn=2, odr=4, dn=p.dn, phi=phi)
    st = p.state.copy()
    oracle.run(cfg, st, 10, 1)
    v = np.abs(st[2] / st[0])[4:-4, 4:-4].max()
    assert v < 0.05
    fac = (2.0 - p.gam) / (p.gam - 1.0)
    j0 = 4
    assert np.allclose(st[7][:j0], st[7][j0] - fac * st[0][j0] * (phi[:j0] - phi[j0]), rtol=0, atol=1e-9)
