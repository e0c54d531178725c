"""Initial conditions for the benchmark workloads (host-side numpy).

States are float64 arrays of shape (8, ny, nx) / (8, nz, ny, nx) / (8, nx),
ghosts included (4 per side on every active axis), variable order
ro, mx, my, mz, bx, by, bz, en (common/mhd_class.hpp:34-40), x fastest
(ss = nx*(ny*k+j)+i).  Face-centred B is built from a vector potential with
the reference's 4th-order difference, as 2D/SERIAL/loop/mhd2d_init_.cpp:39-51
and 3D/MPI/loop/mhd3d_init_.cpp:36-75 do, so the discrete div B is zero to
round-off.  Coordinates are cell-centred: x_i = (i - 4 + 1/2) dx + xmin
(1D/SERIAL/shock/mhd1d_paras.cpp:24).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NG = 4  # ghost cells per side (mhd2d_class.hpp:10)


@dataclass
class Problem:
    """A state plus the scheme/grid settings it is meant to run with."""
    state: np.ndarray
    dx: float
    dy: float = 1.0
    dz: float = 1.0
    gam: float = 5.0 / 3.0
    cfl: float = 0.4
    rmn: int = 1
    odr: int = 4
    r_k: int = 3
    ctw: int = 1
    dn: np.ndarray = field(default_factory=lambda: np.zeros((3, 8), dtype=int))
    nsteps: int = 1
    name: str = ""

    @property
    def ndim(self):
        return self.state.ndim - 1

    @property
    def interior(self):
        return tuple(n - 2 * NG for n in self.state.shape[1:])

    def copy(self):
        p = Problem(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        p.state = self.state.copy()
        p.dn = np.array(self.dn)
        return p


def _df4(fm1, f0, f1, f2):
    """cal_df_4th (common_interp.h:21-24) on arrays."""
    return (27.0 * (f1 - f0) - (f2 - fm1)) / 24.0


def _fcen4(fm1, f0, f1, f2):
    """cal_fcen_4th (common_interp.h:10-13) on arrays."""
    return 0.0625 * (9.0 * (f0 + f1) - (fm1 + f2))


def energy(ro, vx, vy, vz, cx, cy, cz, pr, gam):
    """mhd_energy (common/mhd_func.c:18-28)."""
    v2 = vx * vx + vy * vy + vz * vz
    b2 = cx * cx + cy * cy + cz * cz
    return pr / (gam - 1) + 0.5 * (ro * v2 + b2)


def periodic_fill(a, axes):
    """Periodic ghost copy (bc2d/bc3d dn=0) on the given array axes."""
    for ax in axes:
        n = a.shape[ax]
        sl = [slice(None)] * a.ndim
        src = [slice(None)] * a.ndim
        sl[ax] = slice(0, NG)
        src[ax] = slice(n - 2 * NG, n - NG)
        a[tuple(sl)] = a[tuple(src)]
        sl[ax] = slice(n - NG, n)
        src[ax] = slice(NG, 2 * NG)
        a[tuple(sl)] = a[tuple(src)]


# ---------------------------------------------------------------- 1D
def brio_wu(nx: int = 1024, gam: float = 2.0) -> Problem:
    """Brio & Wu shock tube on [-0.5, 0.5] (1D/SERIAL/shock/mhd1d_init_.cpp:21-30,
    mhd1d_paras.cpp:6-25): open boundaries (dn=+2) for every variable."""
    n = nx + 2 * NG
    dx = 1.0 / nx
    x = (np.arange(n) - NG + 0.5) * dx - 0.5
    left = x <= 0
    ro = np.where(left, 1.0, 0.125)
    pr = np.where(left, 1.0, 0.1)
    by = np.where(left, 1.0, -1.0)
    bx = np.full(n, 0.75)
    z = np.zeros(n)
    st = np.zeros((8, n))
    st[0] = ro
    st[1] = ro * z
    st[2] = ro * z
    st[3] = ro * z
    st[4] = bx
    st[5] = by
    st[6] = z
    st[7] = energy(ro, z, z, z, bx, by, z, pr, gam)
    dn = np.zeros((3, 8), dtype=int)
    dn[0, :] = 2
    return Problem(state=st, dx=dx, gam=gam, cfl=0.4, rmn=1, odr=4, r_k=3, dn=dn, name="brio_wu_%d" % nx)


# ---------------------------------------------------------------- 2D
def _modes2d(rng, nmodes, kmax):
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(-kmax, kmax + 1)
          if 0 < kx * kx + ky * ky <= kmax * kmax]
    idx = rng.choice(len(ks), size=nmodes, replace=False)
    amp = rng.uniform(0.0, 1.0, size=nmodes)
    ph = rng.uniform(0.0, 2.0 * math.pi, size=nmodes)
    return [ks[i] for i in idx], amp, ph


def _potential2d(x, y, modes, amp, ph):
    k, a, p = modes, amp, ph
    out = np.zeros(np.broadcast(x, y).shape)
    for (kx, ky), aa, pp in zip(k, a, p):
        kk = math.sqrt(kx * kx + ky * ky)
        out += (aa / kk) * np.cos(kx * x + ky * y + pp)
    return out


def turbulence2d(n: int = 1024, seed: int = 2, nmodes: int = 16, kmax: int = 4, gam: float = 5.0 / 3.0,
                 vrms: float = 1.0, brms: float = 1.0, rho: float | None = None, pres: float | None = None,
                 ny: int | None = None) -> Problem:
    """Decaying 2D MHD turbulence on [0, 2pi]^2, periodic.

    rho = gam^2, p = gam (c_s = 1) by default.  Velocity = curl(psi z) and
    B = curl(Az z) with psi, Az sums of `nmodes` random Fourier modes
    0 < |k| <= kmax (numpy PCG64(seed)); rescaled to interior rms|v| = vrms and
    rms|B| = brms (cell-centred B from fcen).  Az lives at cell corners."""
    nx = n
    ny = n if ny is None else ny
    rng = np.random.Generator(np.random.PCG64(seed))
    L = 2.0 * math.pi
    dx, dy = L / nx, L / ny
    Nx, Ny = nx + 2 * NG, ny + 2 * NG
    xc = (np.arange(Nx) - NG + 0.5) * dx
    yc = (np.arange(Ny) - NG + 0.5) * dy
    X, Y = np.meshgrid(xc, yc)              # (Ny, Nx) cell centres
    XC, YC = X - 0.5 * dx, Y - 0.5 * dy      # corners (i-1/2, j-1/2)
    # velocity from a stream function (analytic derivatives)
    km, am, pm = _modes2d(rng, nmodes, kmax)
    vx = np.zeros_like(X)
    vy = np.zeros_like(X)
    for (kx, ky), aa, pp in zip(km, am, pm):
        kk = math.sqrt(kx * kx + ky * ky)
        s = -(aa / kk) * np.sin(kx * X + ky * Y + pp)
        vx += ky * s     # d psi / dy
        vy -= kx * s     # -d psi / dx
    # vector potential at corners -> face B (mhd2d_init_.cpp:48-51 pattern)
    ka, aa_, pa = _modes2d(rng, nmodes, kmax)
    az = _potential2d(XC, YC, ka, aa_, pa)
    bx = np.zeros_like(X)
    by = np.zeros_like(X)
    bx[1:-2, :] = _df4(az[0:-3, :], az[1:-2, :], az[2:-1, :], az[3:, :]) / dy
    by[:, 1:-2] = -_df4(az[:, 0:-3], az[:, 1:-2], az[:, 2:-1], az[:, 3:]) / dx
    periodic_fill(bx, (0, 1))
    periodic_fill(by, (0, 1))
    I = (slice(NG, Ny - NG), slice(NG, Nx - NG))
    cx = np.zeros_like(X)
    cy = np.zeros_like(X)
    cx[:, 1:-2] = _fcen4(bx[:, 0:-3], bx[:, 1:-2], bx[:, 2:-1], bx[:, 3:])
    cy[1:-2, :] = _fcen4(by[0:-3, :], by[1:-2, :], by[2:-1, :], by[3:, :])
    sv = vrms / math.sqrt(np.mean(vx[I] ** 2 + vy[I] ** 2))
    sb = brms / math.sqrt(np.mean(cx[I] ** 2 + cy[I] ** 2))
    vx *= sv
    vy *= sv
    bx *= sb
    by *= sb
    cx *= sb
    cy *= sb
    ro = np.full_like(X, gam * gam if rho is None else rho)
    pr = np.full_like(X, gam if pres is None else pres)
    z = np.zeros_like(X)
    periodic_fill(cx, (0, 1))
    periodic_fill(cy, (0, 1))
    st = np.zeros((8, Ny, Nx))
    st[0] = ro
    st[1] = ro * vx
    st[2] = ro * vy
    st[3] = ro * z
    st[4] = bx
    st[5] = by
    st[6] = z
    st[7] = energy(ro, vx, vy, z, cx, cy, z, pr, gam)
    return Problem(state=st, dx=dx, dy=dy, gam=gam, cfl=0.4, rmn=1, odr=4, r_k=3, ctw=1,
                   name="turb2d_%dx%d" % (nx, ny))


def kelvin_helmholtz(n: int = 2048, mach: float = 0.01, seed: int = 3, gam: float = 5.0 / 3.0,
                     width: float = 0.02, b0: float = 0.05) -> Problem:
    """Low-Mach Kelvin-Helmholtz on [0,1]^2, periodic, MLAU (RMN=3).

    rho = 1, p = 1/gam (c_s = 1), Bx = b0.  The shear layer
    vx = M/2 tanh((y-1/2)/w) is mirrored for periodicity as
    vx = M/2 [tanh(s/w) - tanh((s-1/2)/w) - tanh((s+1/2)/w)], s = y - 1/2,
    i.e. layers at y = 1/2 and y = 0 (= 1).  White-noise vy of amplitude
    1e-3 M (uniform in [-1, 1], PCG64(seed)) where |y - 1/2| < 0.05."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = n + 2 * NG
    d = 1.0 / n
    c = (np.arange(N) - NG + 0.5) * d
    X, Y = np.meshgrid(c, c)
    s = Y - 0.5
    vx = 0.5 * mach * (np.tanh(s / width) - np.tanh((s - 0.5) / width) - np.tanh((s + 0.5) / width))
    noise = rng.uniform(-1.0, 1.0, size=(n, n))
    vy = np.zeros_like(X)
    vy[NG:-NG, NG:-NG] = 1e-3 * mach * noise
    vy[np.abs(s) >= 0.05] = 0.0
    periodic_fill(vy, (0, 1))
    ro = np.ones_like(X)
    pr = np.full_like(X, 1.0 / gam)
    bx = np.full_like(X, b0)
    z = np.zeros_like(X)
    st = np.zeros((8, N, N))
    st[0] = ro
    st[1] = ro * vx
    st[2] = ro * vy
    st[3] = ro * z
    st[4] = bx
    st[5] = z
    st[6] = z
    st[7] = energy(ro, vx, vy, z, bx, z, z, pr, gam)
    return Problem(state=st, dx=d, dy=d, gam=gam, cfl=0.4, rmn=3, odr=4, r_k=3, ctw=1, name="khi_%d" % n)


# ---------------------------------------------------------------- 3D
def _modes3d(rng, nmodes, kmax):
    ks = [(a, b, c) for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1) for c in range(-kmax, kmax + 1)
          if 0 < a * a + b * b + c * c <= kmax * kmax]
    idx = rng.choice(len(ks), size=nmodes, replace=False)
    amp = rng.uniform(0.0, 1.0, size=(nmodes, 3))
    ph = rng.uniform(0.0, 2.0 * math.pi, size=(nmodes, 3))
    return [ks[i] for i in idx], amp, ph


class _Numpy:
    """Array operations of the 3D initial condition on the host."""
    sin = staticmethod(np.sin)
    cos = staticmethod(np.cos)

    @staticmethod
    def zeros(shape):
        return np.zeros(shape)

    @staticmethod
    def empty(shape):
        return np.empty(shape)

    @staticmethod
    def full(shape, v):
        return np.full(shape, v)

    @staticmethod
    def arange(n):
        return np.arange(n, dtype=np.float64)

    @staticmethod
    def mean(a):
        return float(np.mean(a))


class _Torch:
    """The same operations on a torch device (float64), for sizes where the
    host would take minutes (512^3: 8 arrays of 1.1 GB)."""

    def __init__(self, device):
        import torch
        self.t = torch
        self.dev = torch.device(device)
        self.sin = torch.sin
        self.cos = torch.cos

    def zeros(self, shape):
        return self.t.zeros(shape, dtype=self.t.float64, device=self.dev)

    def empty(self, shape):
        return self.t.empty(shape, dtype=self.t.float64, device=self.dev)

    def full(self, shape, v):
        return self.t.full(shape, v, dtype=self.t.float64, device=self.dev)

    def arange(self, n):
        return self.t.arange(n, dtype=self.t.float64, device=self.dev)

    def mean(self, a):
        return float(self.t.mean(a))


def _vec_modes(xp, shape, k, amp, ph, X, Y, Z):
    """Sum of modes a_m cos(k.x + p_m) for each of 3 components."""
    out = [xp.zeros(shape) for _ in range(3)]
    for (kx, ky, kz), a, p in zip(k, amp, ph):
        arg = kx * X + ky * Y + kz * Z
        kk = math.sqrt(kx * kx + ky * ky + kz * kz)
        for c in range(3):
            out[c] += (a[c] / kk) * xp.cos(arg + p[c])
    return out


def turbulence3d(n: int = 512, seed: int = 4, nmodes: int = 32, kmax: int = 4, gam: float = 5.0 / 3.0,
                 vrms: float = 1.0, brms: float = 0.5, bz0: float = 0.5, device: str | None = None) -> Problem:
    """3D MHD turbulence on [0, 2pi]^3, periodic, rho = 1, p = 1/gam.

    Velocity = curl(psi) and delta-B = curl(A) with psi, A sums of `nmodes`
    random modes 0 < |k| <= kmax (PCG64(seed)); A lives at cell edges
    (Ax at (i, j-1/2, k-1/2), Ay at (i-1/2, j, k-1/2), Az at (i-1/2, j-1/2, k))
    and face B is its 4th-order discrete curl, as in
    3D/MPI/loop/mhd3d_init_.cpp:36-75; plus a mean Bz = bz0.

    With `device` (a torch device string such as "cuda:0") the fields are
    evaluated there and `state` is a float64 torch tensor on that device;
    the mode draws are the same, the values agree with the host ones to
    rounding of sin/cos."""
    xp = _Numpy() if device is None else _Torch(device)
    rng = np.random.Generator(np.random.PCG64(seed))
    L = 2.0 * math.pi
    d = L / n
    N = n + 2 * NG
    shape = (N, N, N)
    c = (xp.arange(N) - NG + 0.5) * d
    X = c[None, None, :]
    Y = c[None, :, None]
    Z = c[:, None, None]
    h = 0.5 * d
    # velocity: analytic curl of psi at cell centres
    kv, av, pv = _modes3d(rng, nmodes, kmax)
    vx = xp.zeros(shape)
    vy = xp.zeros(shape)
    vz = xp.zeros(shape)
    for (kx, ky, kz), a, p in zip(kv, av, pv):
        arg = kx * X + ky * Y + kz * Z
        kk = math.sqrt(kx * kx + ky * ky + kz * kz)
        s = [-(a[q] / kk) * xp.sin(arg + p[q]) for q in range(3)]  # d/dx of cos = -k sin
        vx += ky * s[2] - kz * s[1]
        vy += kz * s[0] - kx * s[2]
        vz += kx * s[1] - ky * s[0]
        del s, arg
    ka, aa, pa = _modes3d(rng, nmodes, kmax)
    Ax = _vec_modes(xp, shape, ka, aa, pa, X, Y - h, Z - h)[0]
    Ay = _vec_modes(xp, shape, ka, aa, pa, X - h, Y, Z - h)[1]
    Az = _vec_modes(xp, shape, ka, aa, pa, X - h, Y - h, Z)[2]
    bx = xp.zeros(shape)
    by = xp.zeros(shape)
    bz = xp.zeros(shape)
    # bx(i-1/2,j,k) = dAz/dy - dAy/dz
    bx[1:-2, 1:-2, :] = (_df4(Az[1:-2, 0:-3, :], Az[1:-2, 1:-2, :], Az[1:-2, 2:-1, :], Az[1:-2, 3:, :]) / d
                         - _df4(Ay[0:-3, 1:-2, :], Ay[1:-2, 1:-2, :], Ay[2:-1, 1:-2, :], Ay[3:, 1:-2, :]) / d)
    # by(i,j-1/2,k) = dAx/dz - dAz/dx
    by[1:-2, :, 1:-2] = (_df4(Ax[0:-3, :, 1:-2], Ax[1:-2, :, 1:-2], Ax[2:-1, :, 1:-2], Ax[3:, :, 1:-2]) / d
                         - _df4(Az[1:-2, :, 0:-3], Az[1:-2, :, 1:-2], Az[1:-2, :, 2:-1], Az[1:-2, :, 3:]) / d)
    # bz(i,j,k-1/2) = dAy/dx - dAx/dy
    bz[:, 1:-2, 1:-2] = (_df4(Ay[:, 1:-2, 0:-3], Ay[:, 1:-2, 1:-2], Ay[:, 1:-2, 2:-1], Ay[:, 1:-2, 3:]) / d
                         - _df4(Ax[:, 0:-3, 1:-2], Ax[:, 1:-2, 1:-2], Ax[:, 2:-1, 1:-2], Ax[:, 3:, 1:-2]) / d)
    del Ax, Ay, Az
    for b in (bx, by, bz):
        periodic_fill(b, (0, 1, 2))
    I = (slice(NG, N - NG),) * 3
    # energy and the rms of B use cell-centred B (fcen of the face values)
    CX = xp.zeros(shape)
    CY = xp.zeros(shape)
    CZ = xp.zeros(shape)

    def centre():
        CX[..., 1:-2] = _fcen4(bx[..., 0:-3], bx[..., 1:-2], bx[..., 2:-1], bx[..., 3:])
        CY[:, 1:-2, :] = _fcen4(by[:, 0:-3, :], by[:, 1:-2, :], by[:, 2:-1, :], by[:, 3:, :])
        CZ[1:-2] = _fcen4(bz[0:-3], bz[1:-2], bz[2:-1], bz[3:])
        for q in (CX, CY, CZ):
            periodic_fill(q, (0, 1, 2))
    centre()
    sv = vrms / math.sqrt(xp.mean(vx[I] ** 2 + vy[I] ** 2 + vz[I] ** 2))
    sb = brms / math.sqrt(xp.mean(CX[I] ** 2 + CY[I] ** 2 + CZ[I] ** 2))
    for v in (vx, vy, vz):
        v *= sv
    for b in (bx, by, bz):
        b *= sb
    bz += bz0
    centre()  # of the scaled field, for the energy
    st = xp.empty((8,) + shape)
    st[0] = 1.0
    st[1] = vx
    st[2] = vy
    st[3] = vz
    st[4] = bx
    st[5] = by
    st[6] = bz
    del bx, by, bz
    ro = st[0]
    st[7] = energy(ro, vx, vy, vz, CX, CY, CZ, 1.0 / gam, gam)
    del CX, CY, CZ, vx, vy, vz
    return Problem(state=st, dx=d, dy=d, dz=d, gam=gam, cfl=0.4, rmn=1, odr=4, r_k=3, ctw=1,
                   name="turb3d_%d" % n)


# ---------------------------------------------------------------- 2D reconnection
def reconnection2d(nx: int = 1024, ny: int = 128, gam: float = 5.0 / 3.0, beta: float = 0.2, ro0: float = 1.0,
                   ro1: float = 0.2, b0: float = 1.0, b1: float = 0.05, bg: float = 0.0, lam: float = 1.0,
                   rev: float = 2e2, rem: float = 2e2, xlen: float = 64.0):
    """Harris current sheet with a localised resistivity and viscosity, as
    2D/SERIAL/MRX/dmhd2d_init_.cpp:7-103 and mhd2d_paras.cpp:6-40: x in
    [-xlen/2, xlen/2], y in the same aspect as ny/nx, ro = ro0 sech^2(y/lam)
    + ro1, Bx = b0 tanh(y/lam), pressure balance with plasma beta `beta` in
    the lobe, Zenitani's B perturbation of amplitude b1; nu = nu0 exp(-r/lam),
    eta = eta0 exp(-r/lam) with nu0 = vA lam / rev, eta0 = vA lam / rem.
    Boundaries: periodic in x; in y mirror with sign flips of my and by.
    Returns (Problem, nu, eta, dcmax) with dcmax = max(2 nu, eta) (setdc)."""
    xmin, xmax = -0.5 * xlen, 0.5 * xlen
    ymin = -0.5 * (xmax - xmin) * ny / nx
    ymax = -ymin
    dx = (xmax - xmin) / nx
    dy = (ymax - ymin) / ny
    Nx, Ny = nx + 2 * NG, ny + 2 * NG
    x = (np.arange(Nx) - NG + 0.5) * dx + xmin
    y = (np.arange(Ny) - NG + 0.5) * dy + ymin
    X, Y = np.meshgrid(x, y)
    w = lam + 1e-15
    ro = ro0 / np.cosh(Y / w) ** 2 + ro1
    bx = b0 * np.tanh(Y / w)
    by = np.zeros_like(X)
    bz = np.full_like(X, bg)
    cx = bx.copy()
    cy = by.copy()
    pr = 0.5 * ((1.0 + beta) * (b0 * b0 + bg * bg) - (cx * cx + cy * cy + bz * bz))
    xm = X - 0.5 * dx
    ym = Y - 0.5 * dy
    bx = bx - b1 * (Y / lam) * np.exp(-(xm * xm + Y * Y) / (4 * lam * lam))
    by = by + b1 * (X / lam) * np.exp(-(X * X + ym * ym) / (4 * lam * lam))
    g = np.exp(-(X * X + Y * Y) / (4 * lam * lam))
    cx = cx - b1 * (Y / lam) * g
    cy = cy + b1 * (X / lam) * g
    z = np.zeros_like(X)
    st = np.zeros((8, Ny, Nx))
    st[0] = ro
    st[1] = z
    st[2] = z
    st[3] = z
    st[4] = bx
    st[5] = by
    st[6] = bz
    st[7] = energy(ro, z, z, z, cx, cy, bz, pr, gam)
    dn = np.zeros((3, 8), dtype=int)
    dn[1] = [1, 1, -1, 1, 1, -1, 1, 1]
    al = math.sqrt((b0 * b0 + bg * bg) / ro0)
    nu0, eta0 = al * lam / rev, al * lam / rem
    r = np.sqrt(X * X + Y * Y)
    nu = nu0 * np.exp(-r / lam)
    eta = eta0 * np.exp(-r / lam)
    dcmax = float(np.max(np.maximum(2 * nu, eta)))
    p = Problem(state=st, dx=dx, dy=dy, gam=gam, cfl=0.4, rmn=2, odr=2, r_k=3, ctw=1, dn=dn,
                name="mrx_%dx%d" % (nx, ny))
    return p, nu, eta, dcmax


# ---------------------------------------------------------------- 2D Rayleigh-Taylor
def _sqrwave2(val_u, val_l, dx_u, dx_l):
    return 0.5 * (2.0 + np.tanh(dx_u) - np.tanh(dx_l)) * (val_u - val_l) + val_l


def rayleigh_taylor2d(nx: int = 128, ny: int = 1024, gam: float = 5.0 / 3.0, beta: float = 1e3, ro_u: float = 1.0,
                      ro_l: float = 0.1, pr0: float = 500.0, g0: float = 1.0, dv: float = 0.01, lam: float = 1.0,
                      xlen: float = 20.0):
    """Magnetised Rayleigh-Taylor set-up with gravity along y, as
    2D/SERIAL/RTI/mhd2d_init_.cpp:8-130 and mhd2d_paras.cpp:5-40: x in
    [0, xlen], y in the aspect ny/nx around 0; heavy fluid (ro_u) above light
    (ro_l) at y = +-0.35 Ly (tanh layers of width lam); potential
    phi = g0 lg log(cosh(y/lg)), lg = 8 dy; pressure from the hydrostatic
    balance integrated by the reference's midpoint rule (floor 1e-4); uniform
    Bz = sqrt(2 pr0/beta); single-mode vy perturbation of amplitude dv.
    Energy includes ro phi.  Boundaries: periodic in x, open (+2) in y.
    Returns (Problem, phi)."""
    xmin, xmax = 0.0, xlen
    ymin = -0.5 * (xmax - xmin) * ny / nx
    ymax = -ymin
    dx = (xmax - xmin) / nx
    dy = (ymax - ymin) / ny
    Nx, Ny = nx + 2 * NG, ny + 2 * NG
    x = (np.arange(Nx) - NG + 0.5) * dx + xmin
    y = (np.arange(Ny) - NG + 0.5) * dy + ymin
    ly = ymax - ymin
    s0 = 0.35 * ly
    lg = 8 * dy
    b0 = math.sqrt(2.0 * pr0 / beta)
    wlen = (xmax - xmin) / 1
    dvy = dv * np.cos(2 * math.pi * x / wlen)
    ro1 = _sqrwave2(ro_u, ro_l, (y - s0) / lam, (y + s0) / lam)
    phi1 = g0 * lg * np.log(np.cosh(y / lg))
    pr1 = np.empty(Ny)
    for j in range(Ny):  # cal_pressure, mhd2d_init_.cpp:105-120
        n = 1 + int(abs(y[j]) / dy + 0.5)
        dyc = (y[j] - 0.0) / n
        ans = pr0
        for k in range(n):
            yc = 0.0 + (k + 0.5) * dyc
            roc = _sqrwave2(ro_u, ro_l, (yc - s0) / lam, (yc + s0) / lam)
            ans += -roc * (g0 * math.tanh(yc / lg)) * dyc
        pr1[j] = max(ans, 1e-4)
    X, Y = np.meshgrid(x, y)
    ro = np.repeat(ro1[:, None], Nx, axis=1)
    vy = dvy[None, :] * (np.exp(-((Y - s0) ** 2) / (4 * lam * lam)) - np.exp(-((Y + s0) ** 2) / (4 * lam * lam)))
    pr = np.repeat(pr1[:, None], Nx, axis=1)
    phi = np.repeat(phi1[:, None], Nx, axis=1)
    z = np.zeros_like(X)
    bz = np.full_like(X, b0)
    st = np.zeros((8, Ny, Nx))
    st[0] = ro
    st[1] = ro * z
    st[2] = ro * vy
    st[3] = ro * z
    st[4] = z
    st[5] = z
    st[6] = bz
    st[7] = energy(ro, z, vy, z, z, z, bz, pr, gam) + ro * phi
    dn = np.zeros((3, 8), dtype=int)
    dn[1] = 2
    p = Problem(state=st, dx=dx, dy=dy, gam=gam, cfl=0.4, rmn=2, odr=4, r_k=3, ctw=1, dn=dn,
                name="rti_%dx%d" % (nx, ny))
    return p, phi
