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


def test_batched_gradient_equals_per_instance(inv):
    """The batched adjoint (one launch per kernel for all instances) gives
    each instance's gradient bit for bit as its own context does."""
    _say("case: inverse gradient, batched adjoint vs per instance")
    L1, G1 = inv.loss_and_grad(inv.a0, workers=1)
    Lb, Gb = inv.loss_and_grad_batched(inv.a0, groups=1)    # one batch of all instances
    Lb2, Gb2 = inv.loss_and_grad_batched(inv.a0, groups=1)  # kept contexts, second call
    Lg, Gg = inv.loss_and_grad_batched(inv.a0, groups=2)    # two batches on two host threads
    inv.close()
    # batched recorded forward, then each instance's own backward on the
    # checkpoints the batch left in its context (twice: kept contexts)
    Lh, Gh = inv.loss_and_grad_batched(inv.a0, groups=2, backward="per_instance")
    Lh2, Gh2 = inv.loss_and_grad_batched(inv.a0, groups=2, backward="per_instance")
    inv.close()
    # per instance with the contexts kept for the next call
    Lk, Gk = inv.loss_and_grad(inv.a0, workers=2, keep=True)
    Lk2, Gk2 = inv.loss_and_grad(inv.a0, workers=2, keep=True)
    inv.close()
    assert L1 == Lb == Lb2 == Lg == Lh == Lh2 == Lk == Lk2
    assert np.array_equal(G1, Gb) and np.array_equal(G1, Gb2) and np.array_equal(G1, Gg)
    assert np.array_equal(G1, Gh) and np.array_equal(G1, Gh2)
    assert np.array_equal(G1, Gk) and np.array_equal(G1, Gk2)


def test_config_sizes_batched_gradient_vs_fd():
    """8 instances drawn from the workload's size set {128, ..., 512}^2 at 24
    steps: batched gradient equal to the per-instance one, and against
    central differences of the GPU forward for two amplitudes of the
    largest and the smallest instance."""
    from qasmhd_amd import capi, inverse
    capi.load()
    sizes = inverse.config_sizes(8, seed=11)
    _say("case: inverse, 8 instances of sizes", sizes, "24 steps")
    m = inverse.VyModeInversion(sizes, steps=24, seed=11)
    m.schedule()
    Lb, Gb = m.loss_and_grad_batched(m.a0)
    m.close()
    L1, G1 = m.loss_and_grad(m.a0, workers=4)
    assert Lb == L1 and np.array_equal(Gb, G1)
    big, small = int(np.argmax(sizes)), int(np.argmin(sizes))
    for i in (big, small):
        scale = np.abs(Gb[i]).max()
        for k in (0, m.nmodes - 1):
            fd = m.fd_grad(i, m.a0[i], k, h=1e-6)
            _say("  inst %d (n %d) mode %d adjoint %.12e fd %.12e" % (i, sizes[i], k, Gb[i, k], fd))
            assert abs(Gb[i, k] - fd) <= 1e-5 * scale
