"""Initial-condition map of the velocity-mode inversion (qasmhd_amd/
inverse.py), host only: the chain rule used to turn d(loss)/d(state) into
d(loss)/d(amplitudes) against central differences of the map itself, and the
mode set / size draws."""
import numpy as np


def test_chain_rule_of_the_ic_map():
    from qasmhd_amd import inverse
    inv = inverse.VyModeInversion([16, 24], steps=2, seed=3)
    rng = np.random.default_rng(7)
    for i in range(inv.ninst):
        a = inv.a0[i]
        u = inv.state(i, a)
        g = rng.standard_normal(u.shape)
        # the gradient expression of VyModeInversion.instance_grad
        w = g[2] * u[0] + g[7] * u[2]
        grad = np.tensordot(inv.psi[i], w, axes=([1, 2], [0, 1]))
        for m in range(inv.nmodes):
            h = 1e-4
            ap, am = a.copy(), a.copy()
            ap[m] += h
            am[m] -= h
            # the map is quadratic in a: the central difference is exact up to rounding
            fd = np.sum(g * (inv.state(i, ap) - inv.state(i, am))) / (2 * h)
            assert abs(fd - grad[m]) <= 1e-8 * np.abs(grad).max(), (i, m, fd, grad[m])


def test_map_changes_only_my_and_en():
    from qasmhd_amd import inverse
    inv = inverse.VyModeInversion([16], steps=2, seed=4)
    u0 = inv.state(0, np.zeros(inv.nmodes))
    assert np.array_equal(u0, inv.base[0].state)
    u = inv.state(0, inv.a_target[0])
    for v in (0, 1, 3, 4, 5, 6):
        assert np.array_equal(u[v], u0[v])
    # vy = my / rho moved by the mode sum
    d = u[2] / u[0] - u0[2] / u0[0]
    ref = np.tensordot(inv.a_target[0], inv.psi[0], axes=1)
    assert np.allclose(d, ref, rtol=0, atol=1e-14)
    # kinetic energy change only
    de = u[7] - u0[7]
    ke = 0.5 * (u[2] ** 2 - u0[2] ** 2) / u[0]
    assert np.allclose(de, ke, rtol=0, atol=1e-13)


def test_sizes_and_modes_are_seeded():
    from qasmhd_amd import inverse
    s1 = inverse.config_sizes(64, seed=5)
    assert s1 == inverse.config_sizes(64, seed=5)
    assert set(s1) <= {128, 192, 256, 384, 512} and len(s1) == 64
    a = inverse.VyModeInversion([16, 16], steps=1, seed=9)
    b = inverse.VyModeInversion([16, 16], steps=1, seed=9)
    assert np.array_equal(a.a0, b.a0) and np.array_equal(a.a_target, b.a_target)
    assert not np.array_equal(a.psi[0], a.psi[1])  # per-instance modes
