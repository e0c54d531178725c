"""ctypes view of the C restatement (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  States are numpy float64 arrays of shape (8, nz, ny, nx) (or
(8, ny, nx) / (8, nx)), ghosts included, variable order ro, mx, my, mz, bx,
by, bz, en (common/mhd_class.hpp:34-40).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

_lib = None


class QorCfg(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xoff", C.c_int), ("yoff", C.c_int), ("zoff", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("gam", C.c_double), ("cfl", C.c_double),
                ("rmn", C.c_int), ("odr", C.c_int), ("r_k", C.c_int), ("ctw", C.c_int),
                ("dn", (C.c_int * 8) * 3), ("st", (C.c_int * 8) * 3),
                ("phi", C.POINTER(C.c_double))]


def lib():
    global _lib
    if _lib is None:
        path = _build.oracle_lib_path()
        if not os.path.exists(path):
            _build.build_oracle()
        l = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        l.qor_bound.argtypes = [C.POINTER(QorCfg), C.POINTER(dp)]
        l.qor_step.argtypes = [C.POINTER(QorCfg), C.POINTER(dp), C.c_double]
        l.qor_cfl_dt.argtypes = [C.POINTER(QorCfg), C.POINTER(dp)]
        l.qor_cfl_dt.restype = C.c_double
        l.qor_run.argtypes = [C.POINTER(QorCfg), C.POINTER(dp), C.c_int, C.c_int, dp, dp]
        l.qor_interp_batch.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        l.qor_fcen_batch.argtypes = [C.c_int, C.c_int, dp, dp]
        l.qor_df_batch.argtypes = [C.c_int, C.c_int, dp, dp]
        l.qor_lrstate_batch.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, dp, dp]
        l.qor_flux_batch.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_double, dp, dp]
        l.qor_cuct_batch.argtypes = [C.c_int, dp, dp, dp]
        l.qor_dsptv.argtypes = [C.POINTER(QorCfg), C.POINTER(dp), dp, dp, C.c_double]
        l.qor_run_dissipative.argtypes = [C.POINTER(QorCfg), C.POINTER(dp), dp, dp, C.c_double, C.c_int, C.c_int,
                                          dp, dp]
        _lib = l
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_cfg(shape, dx, dy=1.0, dz=1.0, gam=5.0 / 3.0, cfl=0.4, rmn=1, odr=4, r_k=3, ctw=1, dn=None, st=None,
             phi=None):
    """shape = per-variable array shape (nx,), (ny, nx) or (nz, ny, nx);
    phi: 2D gravitational potential per cell (the RTI driver's phi_g)."""
    c = QorCfg()
    if phi is not None:
        c._phi = np.ascontiguousarray(phi, dtype=np.float64)
        c.phi = _dp(c._phi)
    c.ndim = len(shape)
    dims = list(reversed(shape)) + [1] * (3 - len(shape))
    c.nx, c.ny, c.nz = dims
    c.xoff = 4
    c.yoff = 4 if c.ndim >= 2 else 0
    c.zoff = 4 if c.ndim >= 3 else 0
    c.dx, c.dy, c.dz = dx, dy, dz
    c.gam, c.cfl = gam, cfl
    c.rmn, c.odr, c.r_k, c.ctw = rmn, odr, r_k, ctw
    dn = np.zeros((3, 8), dtype=int) if dn is None else np.asarray(dn)
    if st is None:
        st = np.zeros((3, 8), dtype=int)
        st[0, 4] = 1
        st[1, 5] = 1
        st[2, 6] = 1
    st = np.asarray(st)
    for a in range(3):
        for m in range(8):
            c.dn[a][m] = int(dn[a][m])
            c.st[a][m] = int(st[a][m])
    return c


class _Val:
    def __init__(self, state):
        assert state.dtype == np.float64 and state.flags.c_contiguous
        self.state = state
        self.ptrs = (C.POINTER(C.c_double) * 8)(*[_dp(state[m]) for m in range(8)])


def bound(cfg, state):
    v = _Val(state)
    lib().qor_bound(C.byref(cfg), v.ptrs)


def step(cfg, state, dt):
    v = _Val(state)
    lib().qor_step(C.byref(cfg), v.ptrs, dt)


def cfl_dt(cfg, state):
    v = _Val(state)
    return lib().qor_cfl_dt(C.byref(cfg), v.ptrs)


def dsptv(cfg, state, nu, eta, dt):
    """DMHD2D::dsptv(dt) on state (2D), nu / eta per cell (ghosts included)."""
    v = _Val(state)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    lib().qor_dsptv(C.byref(cfg), v.ptrs, _dp(nu), _dp(eta), dt)


def run_dissipative(cfg, state, nu, eta, dcmax, nsteps, flg=1, dt=None):
    """DMHD2D::exec_ loop body nsteps times; returns the dt history."""
    v = _Val(state)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    if dt is None:
        dt = lib().qor_cfl_dt(C.byref(cfg), v.ptrs)
    d = C.c_double(dt)
    hist = np.zeros(max(nsteps, 1))
    lib().qor_run_dissipative(C.byref(cfg), v.ptrs, _dp(nu), _dp(eta), dcmax, nsteps, flg, C.byref(d), _dp(hist))
    return hist[:nsteps]


def run(cfg, state, nsteps, flg=1, dt=None):
    """exec_-style loop; returns the dt history.  dt None -> setdt(1) first."""
    v = _Val(state)
    if dt is None:
        dt = lib().qor_cfl_dt(C.byref(cfg), v.ptrs)
    d = C.c_double(dt)
    hist = np.zeros(max(nsteps, 1))
    lib().qor_run(C.byref(cfg), v.ptrs, nsteps, flg, C.byref(d), _dp(hist))
    return hist[:nsteps], d.value


def interp(kind, f5):
    f5 = np.ascontiguousarray(f5, dtype=np.float64)
    n = f5.shape[0]
    fl, fr = np.zeros(n), np.zeros(n)
    lib().qor_interp_batch(kind, n, _dp(f5), _dp(fl), _dp(fr))
    return fl, fr


def fcen(odr, f4):
    f4 = np.ascontiguousarray(f4, dtype=np.float64)
    out = np.zeros(f4.shape[0])
    lib().qor_fcen_batch(odr, f4.shape[0], _dp(f4), _dp(out))
    return out


def df(odr, f4):
    f4 = np.ascontiguousarray(f4, dtype=np.float64)
    out = np.zeros(f4.shape[0])
    lib().qor_df_batch(odr, f4.shape[0], _dp(f4), _dp(out))
    return out


def lrstate(odr, st, bn, gam):
    st = np.ascontiguousarray(st, dtype=np.float64)
    bn = np.ascontiguousarray(bn, dtype=np.float64)
    n = st.shape[0]
    vl, vr = np.zeros((n, 7)), np.zeros((n, 7))
    lib().qor_lrstate_batch(odr, n, _dp(st), _dp(bn), gam, _dp(vl), _dp(vr))
    return vl, vr


def flux(rmn, ql, qr, bn, gam, dv):
    ql, qr, bn, dv = (np.ascontiguousarray(a, dtype=np.float64) for a in (ql, qr, bn, dv))
    n = ql.shape[0]
    f = np.zeros((n, 7))
    lib().qor_flux_batch(rmn, n, _dp(ql), _dp(qr), _dp(bn), gam, _dp(dv), _dp(f))
    return f


def cuct(cells, faces):
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    faces = np.ascontiguousarray(faces, dtype=np.float64)
    out = np.zeros(cells.shape[0])
    lib().qor_cuct_batch(cells.shape[0], _dp(cells), _dp(faces), _dp(out))
    return out
