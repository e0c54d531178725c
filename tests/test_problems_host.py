"""Initial conditions of the benchmark workloads (qasmhd_amd/problems.py),
checked on the CPU against what their definitions fix, not against the
formulas that build them:

- face B built as the 4th-order discrete curl of a potential (as
  2D/SERIAL/loop/mhd2d_init_.cpp:39-51 and 3D/MPI/loop/mhd3d_init_.cpp:36-75
  do) has zero 4th-order discrete divergence to round-off, and its interior
  mean is the imposed mean field (the differences telescope on a period);
- periodic workloads carry periodic ghosts in every variable;
- the stated rms of v, and the pressure recovered from the total energy;
- Kelvin-Helmholtz shear layer odd about y = 1/2;
- the Harris sheet in pressure balance; Rayleigh-Taylor in hydrostatic
  balance against an independent quadrature of dp/dy = -rho dphi/dy.
A wrong sign, a swapped dx / dy (grids with nx != ny), or a face offset by
one cell each move one of these checks by O(1)."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from qasmhd_amd import problems  # noqa: E402

NG = 4


def d4(fm1, f0, f1, f2):
    """Staggered 4th-order first difference times h: exact for cubics."""
    return (27.0 * (f1 - f0) - (f2 - fm1)) / 24.0


def test_d4_is_exact_on_cubics():
    h = 0.37
    x = np.array([-1.5, -0.5, 0.5, 1.5]) * h + 0.2
    f = 1.0 - 2.0 * x + 3.0 * x ** 2 + 5.0 * x ** 3
    assert abs(d4(*f) / h - (-2.0 + 6.0 * 0.2 + 15.0 * 0.2 ** 2)) < 1e-12


def _periodic_index(n):
    return (np.arange(n) - NG) % (n - 2 * NG) + NG


def _assert_periodic(state, axes, exact=range(8)):
    """Ghosts equal their periodic images: bit for bit in the variables the
    builder copies (`exact`), to round-off where it evaluates the analytic
    field in the ghost cells too."""
    for ax in axes:
        idx = _periodic_index(state.shape[ax])
        img = np.take(state, idx, axis=ax)
        for m in range(8):
            if m in exact:
                assert np.array_equal(state[m], img[m]), (ax, m)
            else:
                assert np.abs(state[m] - img[m]).max() < 1e-12 * (1 + np.abs(state[m]).max()), (ax, m)


def _pressure(st, gam, cx, cy, cz):
    ro = st[0]
    ke = 0.5 * (st[1] ** 2 + st[2] ** 2 + st[3] ** 2) / ro
    return (gam - 1.0) * (st[7] - ke - 0.5 * (cx * cx + cy * cy + cz * cz))


def test_turbulence2d_divergence_free_periodic_rms():
    p = problems.turbulence2d(n=64, ny=48, seed=2)
    st = p.state
    assert st.shape == (8, 56, 72)
    assert p.dx != p.dy
    bx, by = st[4], st[5]
    div = (d4(bx[1:-2, 0:-3], bx[1:-2, 1:-2], bx[1:-2, 2:-1], bx[1:-2, 3:]) / p.dx
           + d4(by[0:-3, 1:-2], by[1:-2, 1:-2], by[2:-1, 1:-2], by[3:, 1:-2]) / p.dy)
    scale = max(np.abs(bx).max(), np.abs(by).max()) / min(p.dx, p.dy)
    assert np.abs(div).max() < 1e-12 * scale
    # the same divergence with by's sign flipped is O(1): the check has teeth
    flip = (d4(bx[1:-2, 0:-3], bx[1:-2, 1:-2], bx[1:-2, 2:-1], bx[1:-2, 3:]) / p.dx
            - d4(by[0:-3, 1:-2], by[1:-2, 1:-2], by[2:-1, 1:-2], by[3:, 1:-2]) / p.dy)
    assert np.abs(flip).max() > 1e-2 * scale
    I = (slice(NG, -NG), slice(NG, -NG))
    # zero mean field: the face differences of a periodic potential telescope
    assert abs(bx[I].mean()) < 1e-13 and abs(by[I].mean()) < 1e-13
    _assert_periodic(st, (1, 2), exact=(0, 3, 4, 5, 6))
    ro = st[0]
    assert np.all(ro == (5.0 / 3.0) ** 2)
    v2 = (st[1] ** 2 + st[2] ** 2) / ro ** 2
    assert abs(math.sqrt(v2[I].mean()) - 1.0) < 1e-12
    assert np.all(st[3] == 0.0) and np.all(st[6] == 0.0)
    # cell-centred B from the face average (2nd order) against brms = 1 and the
    # stated pressure gam: both within the interpolation error at 16 cells
    # per shortest wavelength
    cx = 0.5 * (bx[:, :-1] + bx[:, 1:])[:-1]
    cy = 0.5 * (by[:-1] + by[1:])[:, :-1]
    J = (slice(NG, -NG), slice(NG, -NG + 1))
    assert abs(math.sqrt((cx ** 2 + cy ** 2)[J].mean()) - 1.0) < 0.03
    pr = _pressure(st[:, :-1, :-1], p.gam, cx, cy, 0.0)
    assert np.abs(pr[J] - p.gam).max() < 0.05 * p.gam


def test_turbulence2d_is_seeded():
    a = problems.turbulence2d(n=32, seed=5).state
    b = problems.turbulence2d(n=32, seed=5).state
    c = problems.turbulence2d(n=32, seed=6).state
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_turbulence3d_divergence_free_periodic_mean_field():
    p = problems.turbulence3d(n=24, seed=4, bz0=0.5)
    st = p.state
    assert st.shape == (8, 32, 32, 32)
    bx, by, bz = st[4], st[5], st[6]
    c = slice(1, -2)
    div = (d4(bx[c, c, 0:-3], bx[c, c, 1:-2], bx[c, c, 2:-1], bx[c, c, 3:])
           + d4(by[c, 0:-3, c], by[c, 1:-2, c], by[c, 2:-1, c], by[c, 3:, c])
           + d4(bz[0:-3, c, c], bz[1:-2, c, c], bz[2:-1, c, c], bz[3:, c, c])) / p.dx
    scale = max(np.abs(b).max() for b in (bx, by, bz)) / p.dx
    assert np.abs(div).max() < 1e-12 * scale
    I = (slice(NG, -NG),) * 3
    assert abs(bx[I].mean()) < 1e-13 and abs(by[I].mean()) < 1e-13
    assert abs(bz[I].mean() - 0.5) < 1e-13
    _assert_periodic(st, (1, 2, 3), exact=(0, 4, 5, 6))
    assert np.all(st[0] == 1.0)
    v2 = st[1] ** 2 + st[2] ** 2 + st[3] ** 2
    assert abs(math.sqrt(v2[I].mean()) - 1.0) < 1e-12


def test_kelvin_helmholtz_layer_and_pressure():
    p = problems.kelvin_helmholtz(n=64, mach=0.01, seed=3, b0=0.05)
    st = p.state
    N = 72
    assert st.shape == (8, N, N)
    _assert_periodic(st, (1, 2), exact=(0, 2, 3, 4, 5, 6))
    assert np.all(st[0] == 1.0)
    vx, vy = st[1], st[2]
    # odd about y = 1/2 (rows j and N-1-j sit at 1/2 -+ s; the mirrored
    # layers swap places in the sum, hence round-off)
    assert np.abs(vx + vx[::-1]).max() < 1e-12 * np.abs(vx).max()
    assert np.all(vx == vx[:, :1])
    assert np.abs(vx).max() <= 0.5 * 0.01 * (1 + 1e-9)
    y = (np.arange(N) - NG + 0.5) / 64
    outside = np.abs(y - 0.5) >= 0.05
    assert np.all(vy[outside] == 0.0)
    assert np.abs(vy).max() <= 1e-3 * 0.01
    assert np.count_nonzero(vy[NG:-NG, NG:-NG]) > 0
    assert np.all(st[4] == 0.05) and np.all(st[5] == 0.0) and np.all(st[6] == 0.0)
    pr = _pressure(st, p.gam, st[4], st[5], st[6])
    assert np.abs(pr - 1.0 / p.gam).max() < 1e-14
    assert p.rmn == 3


def test_brio_wu_states():
    p = problems.brio_wu(nx=64)
    st = p.state
    x = (np.arange(72) - NG + 0.5) / 64 - 0.5
    left = x <= 0
    assert np.count_nonzero(left) == 36
    assert np.all(st[0] == np.where(left, 1.0, 0.125))
    assert np.all(st[4] == 0.75) and np.all(st[5] == np.where(left, 1.0, -1.0))
    pr = _pressure(st, 2.0, st[4], st[5], st[6])
    assert np.allclose(pr, np.where(left, 1.0, 0.1), rtol=0, atol=1e-14)
    assert np.all(p.dn[0] == 2)


def test_harris_sheet_pressure_balance():
    beta, b0, bg = 0.2, 1.0, 0.3
    p, nu, eta, dcmax = problems.reconnection2d(nx=64, ny=16, beta=beta, b0=b0, b1=0.0, bg=bg)
    st = p.state
    # unperturbed: face bx equals the cell value, by = 0
    assert np.all(st[5] == 0.0)
    pr = _pressure(st, p.gam, st[4], st[5], st[6])
    total = pr + 0.5 * (st[4] ** 2 + st[6] ** 2)
    assert np.abs(total - 0.5 * (1 + beta) * (b0 * b0 + bg * bg)).max() < 1e-13
    # lobe plasma beta: pr / (B^2/2) far from the sheet
    lobe = pr[NG, NG] / (0.5 * (st[4][NG, NG] ** 2 + bg * bg))
    assert abs(lobe - beta) < 1e-3
    # y mirror flips my and by; x periodic
    assert list(p.dn[1]) == [1, 1, -1, 1, 1, -1, 1, 1] and np.all(p.dn[0] == 0)
    # equal Reynolds numbers: eta0 = nu0, so dcmax = 2 nu at the centre
    assert np.allclose(nu, eta) and abs(dcmax - 2.0 * nu.max()) < 1e-15
    # eta0 four times the default (rem 200 -> 50) now dominates: dcmax = eta0 = 2x
    _, _, eta2, dc2 = problems.reconnection2d(nx=64, ny=16, beta=beta, b0=b0, b1=0.0, bg=bg, rem=50.0)
    assert abs(dc2 - 2.0 * dcmax) < 1e-15 and abs(dc2 - eta2.max()) < 1e-15


def test_rayleigh_taylor_hydrostatic_balance():
    from scipy.integrate import quad
    nx, ny = 16, 96
    p, phi = problems.rayleigh_taylor2d(nx=nx, ny=ny)
    st = p.state
    gam, ro_u, ro_l, pr0, g0, lam = p.gam, 1.0, 0.1, 500.0, 1.0, 1.0
    dy = p.dy
    ly = 20.0 * ny / nx
    s0, lg = 0.35 * ly, 8 * dy
    y = (np.arange(ny + 2 * NG) - NG + 0.5) * dy - 0.5 * ly

    def rho(yy):
        # heavy above y = +s0 ... light between, tanh steps of width lam
        return 0.5 * (2.0 + math.tanh((yy - s0) / lam) - math.tanh((yy + s0) / lam)) * (ro_u - ro_l) + ro_l

    assert abs(st[0][0, 0] - rho(y[0])) < 1e-14 and abs(st[0][-1, 0] - rho(y[-1])) < 1e-14
    assert np.allclose(phi[:, 0], g0 * lg * np.log(np.cosh(y / lg)), rtol=0, atol=1e-12)
    ro = st[0]
    pr = _pressure(st, gam, st[4], st[5], st[6]) - (gam - 1.0) * ro * phi
    assert np.all(pr > 0)
    for j in (0, NG, ny // 2 - 3, ny // 2 + 3, ny + NG - 1, ny + 2 * NG - 1):
        exact = pr0 - quad(lambda s: rho(s) * g0 * math.tanh(s / lg), 0.0, y[j], limit=200)[0]
        assert abs(pr[j, NG] - exact) < 2e-3 * exact, (j, pr[j, NG], exact)
    assert np.all(p.dn[1] == 2) and np.all(p.dn[0] == 0)
    # one mode in x: vy changes sign between x = 0 and x = xlen/2
    vy = st[2] / ro
    j = int(np.argmax(np.abs(vy[:, NG])))
    assert vy[j, NG] * vy[j, NG + nx // 2] < 0
