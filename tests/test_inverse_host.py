"""Host logic of the velocity-mode inversion (qasmhd_amd/inverse.py) with
stand-in contexts, no GPU: which device contexts each gradient path keeps
and frees, and that the per-instance and batched paths assemble the same
loss and gradient from the same per-instance adjoints.  The stand-in
forward is linear, rho_final = rho + C my with every other variable kept,
and its adjoint is exact, so the gradient has a closed form."""
import numpy as np
import pytest

from qasmhd_amd import inverse

NG = 4
C = 0.5


def forward(u):
    uf = np.array(u)
    uf[0] = u[0] + C * u[2]
    return uf


class FakeSolver:
    live = []

    def __init__(self, shape):
        self.shape = shape
        self.open = True
        self.u = None
        FakeSolver.live.append(self)

    def upload(self, u):
        assert self.open
        self.u = np.array(u)

    def download(self):
        assert self.open
        return forward(self.u)

    def reset_counter(self):
        pass

    def adj_forward(self, dts):
        assert self.open and len(dts) > 0

    def adj_backward(self, g):
        assert self.open
        gi = np.array(g)
        gi[2] += C * g[0]
        return gi

    def close(self):
        self.open = False


class FakeBatch:
    live = []

    def __init__(self, solvers):
        self.solvers = solvers
        self.open = True
        FakeBatch.live.append(self)

    def adj_forward(self, dts):
        assert self.open and len(dts) == len(self.solvers)

    def adj_backward(self, gs):
        assert self.open
        return [s.adj_backward(g) for s, g in zip(self.solvers, gs)]

    def close(self):
        self.open = False


@pytest.fixture
def inv(monkeypatch):
    FakeSolver.live.clear()
    FakeBatch.live.clear()
    monkeypatch.setattr(inverse.capi, "Batch", FakeBatch)
    v = inverse.VyModeInversion(sizes=[16, 24, 20], steps=3, nmodes=4)
    v._solver = lambda i: FakeSolver(v.base[i].state.shape[1:])
    v.dts = [np.full(3, 0.01) for _ in v.sizes]
    v.rho_target = [forward(v.state(i, v.a_target[i]))[0] for i in range(v.ninst)]
    yield v
    v.close()


def _open():
    return [s for s in FakeSolver.live if s.open], [b for b in FakeBatch.live if b.open]


def _expected(v, a):
    L, G = 0.0, np.zeros_like(a)
    for i in range(v.ninst):
        u0 = v.state(i, a[i])
        r = (u0[0] + C * u0[2] - v.rho_target[i])[NG:-NG, NG:-NG]
        L += float(np.sum(r * r))
        # dL/da_m = sum 2 r C d(my)/da_m with my = my0 + ro psi_m . a
        # (energy carries no misfit)
        dmy = np.zeros_like(u0[0])
        dmy[NG:-NG, NG:-NG] = 2.0 * r * C
        G[i] = np.tensordot(v.psi[i], dmy * u0[0], axes=([1, 2], [0, 1]))
    return L, G


def test_paths_agree_and_hold_one_context_set(inv):
    a = inv.a0
    L0, G0 = _expected(inv, a)
    assert L0 > 0
    # the stand-in loss is quadratic in a: a central difference is exact
    # up to rounding
    for i, m in ((0, 1), (2, 3)):
        h = 1e-3
        ap, am = a.copy(), a.copy()
        ap[i, m] += h
        am[i, m] -= h
        fd = (_expected(inv, ap)[0] - _expected(inv, am)[0]) / (2 * h)
        assert abs(fd - G0[i, m]) < 1e-7 * (abs(G0[i, m]) + 1e-12)
    Lb, Gb = inv.loss_and_grad_batched(a, groups=2)
    s, b = _open()
    assert len(s) == 3 and len(b) == 2
    Lk, Gk = inv.loss_and_grad(a, keep=True, workers=1)
    s, b = _open()
    assert len(s) == 3 and not b and all(x in inv._ipool for x in s)
    Lk2, Gk2 = inv.loss_and_grad(a, keep=True, workers=2)
    assert _open()[0] == s
    Lp, Gp = inv.loss_and_grad_batched(a, groups=3, backward="per_instance")
    s, b = _open()
    assert len(s) == 3 and len(b) == 3 and inv._ipool is None
    Lf, Gf = inv.loss_and_grad(a, keep=False)
    assert _open() == ([], []) and inv._gpool is None
    for L, G in ((Lb, Gb), (Lk, Gk), (Lk2, Gk2), (Lp, Gp), (Lf, Gf)):
        assert L == L0
        assert np.array_equal(G, G0)


def test_close_frees_everything(inv):
    inv.loss_and_grad_batched(inv.a0, groups=2)
    inv.close()
    assert _open() == ([], [])
    inv.loss_and_grad(inv.a0, keep=True)
    inv.close()
    assert _open() == ([], [])


def test_backward_choice_is_checked(inv):
    with pytest.raises(ValueError):
        inv.loss_and_grad_batched(inv.a0, backward="both")
