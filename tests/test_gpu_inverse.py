"""Velocity-mode inversion on batched 2D domains (qasmhd_amd/inverse.py):
the adjoint gradient of the loss with respect to the 8 mode amplitudes of
each instance against central finite differences (h = 1e-6, Richardson-
extrapolated) of the GPU forward over the frozen dt schedule, and the GPU
loss against the oracle's loss over the same schedule (bit-exact, so the
oracle's finite differences are the GPU's).

Forward: MHD2D::ideal (2D/SERIAL/OTvortex/mhd2d_solve.cpp:3-400) at the dt
values the batched run recorded.  Tolerance of the gradient: 1e-5 of the
instance's largest gradient entry."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NG = 4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def inv():
    from qasmhd_amd import capi, inverse
    capi.load()
    m = inverse.VyModeInversion([32, 48, 40], steps=6, seed=5)
    m.schedule()
    return m


def test_schedule_is_the_setdt_schedule(inv):
    """The batched run's dt history equals each instance's own qas_run."""
    _say("case: inverse schedule, sizes", inv.sizes)
    for i, d in enumerate(inv.dts):
        assert d.shape == (inv.steps,) and np.all(d > 0)
        with inv._solver(i) as s:
            s.upload(inv.state(i, inv.a0[i]))
            hist, _ = s.run(inv.steps, flg=1)
        assert np.array_equal(hist, d)


def test_gradient_vs_fd(inv):
    _say("case: inverse gradient vs FD, sizes", inv.sizes, "steps", inv.steps)
    a = inv.a0
    L, G = inv.loss_and_grad(a)
    assert L > 0 and np.all(np.isfinite(G))
    for i in range(inv.ninst):
        ms = range(inv.nmodes) if i == 0 else (0, inv.nmodes - 1)
        scale = np.abs(G[i]).max()
        for m in ms:
            fd = inv.fd_grad(i, a[i], m, h=1e-6)
            _say("  inst %d mode %d adjoint %.12e fd %.12e" % (i, m, G[i, m], fd))
            assert abs(G[i, m] - fd) <= 1e-5 * scale, (i, m, G[i, m], fd)


def test_threaded_gradient_equals_sequential(inv):
    # instances on their own contexts and streams, driven by host threads
    _say("case: inverse gradient, 3 threads vs 1")
    L1, G1 = inv.loss_and_grad(inv.a0, workers=1)
    L3, G3 = inv.loss_and_grad(inv.a0, workers=3)
    assert L1 == L3 and np.array_equal(G1, G3)


def test_loss_matches_oracle(inv, oracle):
    pyoracle = oracle
    _say("case: inverse loss vs oracle, instance 0")
    i = 0
    p = inv.base[i]
    cfg = pyoracle.make_cfg(p.state.shape[1:], p.dx, p.dy, gam=p.gam)

    def ofinal(a):
        st = inv.state(i, a)
        pyoracle.bound(cfg, st)
        for dt in inv.dts[i]:
            pyoracle.step(cfg, st, dt)
        return st
    tgt = ofinal(inv.a_target[i])
    assert np.array_equal(tgt[0], inv.targets()[i])
    a = inv.a0[i] + 0.01
    r = ofinal(a)[0] - tgt[0]
    lo = float(np.sum(r[NG:-NG, NG:-NG] ** 2))
    assert inv.instance_loss(i, a) == lo
