"""ctypes binding of libqasmhd_amd (include/qasmhd.h) and a small Python
driver over it.

The library must have been built (``python -m qasmhd_amd.build``); there is
no fallback: if the shared library is missing, or no HIP device is present,
every entry point raises.

States are float64 numpy arrays of shape (8, [nz,] [ny,] nx), ghosts
included, variable order ro, mx, my, mz, bx, by, bz, en.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libqasmhd_amd.so")

QAS_OK = 0


class QasGrid(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xoff", C.c_int), ("yoff", C.c_int), ("zoff", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class QasScheme(C.Structure):
    _fields_ = [("gam", C.c_double), ("cfl", C.c_double), ("rmn", C.c_int), ("odr", C.c_int),
                ("r_k", C.c_int), ("ctw", C.c_int), ("dn", (C.c_int * 8) * 3), ("st", (C.c_int * 8) * 3),
                ("precision", C.c_int)]


EXPORTS = ["qas_scheme_default", "qas_create", "qas_destroy", "qas_last_error", "qas_upload", "qas_download",
           "qas_upload_packed", "qas_download_packed", "qas_upload_device", "qas_download_device", "qas_bound",
           "qas_step", "qas_cfl_dt", "qas_run", "qas_reset_counter", "qas_use_graphs", "qas_sync", "qas_device_bytes",
           "qas_prof_enable", "qas_prof_query", "qas_kernel_model", "qas_point_flops", "qas_adjoint_selfcheck",
           "qas_batch_create", "qas_batch_run", "qas_batch_last_error", "qas_batch_destroy",
           "qas_batch_tune", "qas_batch_set_shape", "qas_batch_shape",
           "qas_adj_forward", "qas_adj_backward", "qas_adj_backward_device",
           "qas_test_interp", "qas_test_f4", "qas_test_lrstate", "qas_test_flux", "qas_test_cuct", "qas_test_rcp",
           "qas_set_dissipation", "qas_dsptv", "qas_set_gravity", "qas_run_static",
           "qas_batch_adj_forward", "qas_batch_adj_backward", "qas_adj_keep_stages", "qas_tune", "qas_set_shape",
           "qas_shape", "qas_batch_use_graphs"]

_lib = None


class QasError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load the HIP library (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise QasError("libqasmhd_amd.so is not built (%s); run `python -m qasmhd_amd.build`" % p)
    l = C.CDLL(p)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    l.qas_scheme_default.argtypes = [C.POINTER(QasScheme)]
    l.qas_scheme_default.restype = None
    l.qas_create.argtypes = [C.POINTER(vp), C.POINTER(QasGrid), C.POINTER(QasScheme)]
    l.qas_destroy.argtypes = [vp]
    l.qas_destroy.restype = None
    l.qas_last_error.argtypes = [vp]
    l.qas_last_error.restype = C.c_char_p
    for f in ("qas_upload", "qas_download"):
        getattr(l, f).argtypes = [vp, C.POINTER(dp)]
    for f in ("qas_upload_packed", "qas_download_packed"):
        getattr(l, f).argtypes = [vp, dp]
    for f in ("qas_upload_device", "qas_download_device"):
        getattr(l, f).argtypes = [vp, vp]
    l.qas_bound.argtypes = [vp]
    l.qas_step.argtypes = [vp, C.c_double]
    l.qas_cfl_dt.argtypes = [vp, dp]
    l.qas_run.argtypes = [vp, C.c_int, C.c_int, dp, dp]
    l.qas_reset_counter.argtypes = [vp]
    l.qas_sync.argtypes = [vp]
    l.qas_device_bytes.argtypes = [vp]
    l.qas_device_bytes.restype = C.c_longlong
    l.qas_adjoint_selfcheck.argtypes = [C.c_int, C.c_int, C.c_ulonglong, dp]
    l.qas_batch_create.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int]
    l.qas_batch_run.argtypes = [vp, C.c_int, C.c_int, dp, dp]
    l.qas_batch_tune.argtypes = [vp]
    l.qas_batch_set_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    l.qas_batch_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    l.qas_batch_last_error.argtypes = [vp]
    l.qas_batch_last_error.restype = C.c_char_p
    l.qas_batch_destroy.argtypes = [vp]
    l.qas_batch_destroy.restype = None
    l.qas_adj_forward.argtypes = [vp, C.c_int, dp]
    l.qas_adj_backward.argtypes = [vp, dp, dp]
    l.qas_adj_backward_device.argtypes = [vp, vp, vp]
    l.qas_adj_keep_stages.argtypes = [vp, C.c_int]
    l.qas_tune.argtypes = [vp]
    l.qas_batch_use_graphs.argtypes = [vp, C.c_int]
    l.qas_set_shape.argtypes = [vp, C.c_int, C.c_int]
    l.qas_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    l.qas_test_interp.argtypes = [C.c_int, C.c_int, dp, dp, dp]
    l.qas_test_f4.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    l.qas_test_lrstate.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, dp, dp]
    l.qas_test_flux.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_double, dp, dp]
    l.qas_test_cuct.argtypes = [C.c_int, dp, dp, dp]
    l.qas_test_rcp.argtypes = [C.c_int, dp, dp, dp]
    l.qas_set_dissipation.argtypes = [vp, dp, dp, C.c_double]
    l.qas_dsptv.argtypes = [vp, C.c_double]
    l.qas_set_gravity.argtypes = [vp, dp]
    l.qas_batch_adj_forward.argtypes = [vp, C.c_int, dp]
    l.qas_batch_adj_backward.argtypes = [vp, C.POINTER(dp), C.POINTER(dp)]
    l.qas_run_static.argtypes = [C.POINTER(QasGrid), C.POINTER(QasScheme), vp, C.c_int, C.c_longlong, C.c_int,
                                 C.c_double, vp, vp, C.c_char_p, C.c_int]
    _lib = l
    return l


ADJ_POINT = {"wcns4": 0, "lin4": 1, "lrstate": 2, "lrfb": 3, "cuct": 4, "hlld": 5, "mlau": 6, "lhlld": 7}


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _chk(r, what):
    if r != QAS_OK:
        raise QasError("%s failed (%d)" % (what, r))


class points:
    """Batched device evaluation of the per-point functions (qas_test_*)."""

    @staticmethod
    def interp(kind, f5):
        f5 = np.ascontiguousarray(f5, dtype=np.float64)
        n = f5.shape[0]
        fl, fr = np.zeros(n), np.zeros(n)
        _chk(load().qas_test_interp(kind, n, _p(f5), _p(fl), _p(fr)), "qas_test_interp")
        return fl, fr

    @staticmethod
    def f4(op, odr, f4):
        f4 = np.ascontiguousarray(f4, dtype=np.float64)
        out = np.zeros(f4.shape[0])
        _chk(load().qas_test_f4(op, odr, f4.shape[0], _p(f4), _p(out)), "qas_test_f4")
        return out

    @staticmethod
    def lrstate(odr, st, bn, gam):
        st = np.ascontiguousarray(st, dtype=np.float64)
        bn = np.ascontiguousarray(bn, dtype=np.float64)
        n = st.shape[0]
        vl, vr = np.zeros((n, 7)), np.zeros((n, 7))
        _chk(load().qas_test_lrstate(odr, n, _p(st), _p(bn), gam, _p(vl), _p(vr)), "qas_test_lrstate")
        return vl, vr

    @staticmethod
    def flux(rmn, ql, qr, bn, gam, dv):
        ql, qr, bn, dv = (np.ascontiguousarray(a, dtype=np.float64) for a in (ql, qr, bn, dv))
        f = np.zeros((ql.shape[0], 7))
        _chk(load().qas_test_flux(rmn, ql.shape[0], _p(ql), _p(qr), _p(bn), gam, _p(dv), _p(f)), "qas_test_flux")
        return f

    @staticmethod
    def cuct(cells, faces):
        cells = np.ascontiguousarray(cells, dtype=np.float64)
        faces = np.ascontiguousarray(faces, dtype=np.float64)
        out = np.zeros(cells.shape[0])
        _chk(load().qas_test_cuct(cells.shape[0], _p(cells), _p(faces), _p(out)), "qas_test_cuct")
        return out

    @staticmethod
    def rcp(d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        fast, div = np.zeros(d.shape[0]), np.zeros(d.shape[0])
        _chk(load().qas_test_rcp(d.shape[0], _p(d), _p(fast), _p(div)), "qas_test_rcp")
        return fast, div


def adjoint_selfcheck(name: str, ntrials: int = 1000, seed: int = 1):
    """Relative errors of one per-point adjoint against Richardson-extrapolated
    central differences of its forward function (host only, no device)."""
    l = C.CDLL(LIB_PATH)
    l.qas_adjoint_selfcheck.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_double)]
    err = np.zeros(ntrials)
    r = l.qas_adjoint_selfcheck(ADJ_POINT[name], ntrials, seed, err.ctypes.data_as(C.POINTER(C.c_double)))
    if r != QAS_OK:
        raise QasError("qas_adjoint_selfcheck(%s): %d" % (name, r))
    return err


def exported_symbols(path: str | None = None):
    """Names from EXPORTS that the library resolves (no device calls)."""
    l = C.CDLL(path or LIB_PATH)
    return [n for n in EXPORTS if hasattr(l, n)]


def make_scheme(gam=5.0 / 3.0, cfl=0.4, rmn=1, odr=4, r_k=3, ctw=1, dn=None, st=None, precision=0):
    s = QasScheme()
    load().qas_scheme_default(C.byref(s))
    s.gam, s.cfl, s.rmn, s.odr, s.r_k, s.ctw, s.precision = gam, cfl, rmn, odr, r_k, ctw, precision
    if dn is not None:
        dn = np.asarray(dn)
        for a in range(3):
            for m in range(8):
                s.dn[a][m] = int(dn[a][m])
    if st is not None:
        st = np.asarray(st)
        for a in range(3):
            for m in range(8):
                s.st[a][m] = int(st[a][m])
    return s


def make_grid(shape, dx, dy=1.0, dz=1.0):
    """shape = per-variable array shape (nx,), (ny, nx) or (nz, ny, nx)."""
    g = QasGrid()
    g.ndim = len(shape)
    dims = list(reversed(shape)) + [1] * (3 - len(shape))
    g.nx, g.ny, g.nz = dims
    g.xoff = 4
    g.yoff = 4 if g.ndim >= 2 else 0
    g.zoff = 4 if g.ndim >= 3 else 0
    g.dx, g.dy, g.dz = dx, dy, dz
    return g


# launch mode of the solvers and batches created from here on: None keeps
# the library's defaults (qas_run launches every kernel from the host,
# qas_batch_run replays captured step pairs); True replays pairs of steps as
# captured HIP graphs, False launches from the host (qas_use_graphs,
# qas_batch_use_graphs)
GRAPHS = None


class Solver:
    """One device-resident domain (owns a qas_ctx)."""

    def __init__(self, shape, dx, dy=1.0, dz=1.0, **scheme):
        self.lib = load()
        self.shape = tuple(shape)
        self.grid = make_grid(shape, dx, dy, dz)
        self.scheme = make_scheme(**scheme)
        self.ctx = C.c_void_p()
        r = self.lib.qas_create(C.byref(self.ctx), C.byref(self.grid), C.byref(self.scheme))
        if r != QAS_OK:
            msg = self.lib.qas_last_error(self.ctx).decode() if self.ctx else "qas_create failed"
            self.lib.qas_destroy(self.ctx)
            self.ctx = None
            raise QasError("qas_create: %s (%d)" % (msg, r))
        if GRAPHS is not None:
            self._check(self.lib.qas_use_graphs(self.ctx, 1 if GRAPHS else 0), "qas_use_graphs")

    def _check(self, r, what):
        if r != QAS_OK:
            raise QasError("%s: %s (%d)" % (what, self.lib.qas_last_error(self.ctx).decode(), r))

    def close(self):
        if self.ctx:
            self.lib.qas_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, state):
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert st.shape == (8,) + self.shape, (st.shape, self.shape)
        self._check(self.lib.qas_upload_packed(self.ctx, st.ctypes.data_as(C.POINTER(C.c_double))), "qas_upload")

    def download(self):
        out = np.empty((8,) + self.shape)
        self._check(self.lib.qas_download_packed(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "qas_download")
        return out

    def upload_device(self, ptr):
        """Copy a packed (8, ...) float64 device array (raw pointer) in."""
        self._check(self.lib.qas_upload_device(self.ctx, C.c_void_p(ptr)), "qas_upload_device")

    def download_device(self, ptr):
        """Copy the state into a packed (8, ...) float64 device array (raw
        pointer); ordered on the context's stream, call sync() before use."""
        self._check(self.lib.qas_download_device(self.ctx, C.c_void_p(ptr)), "qas_download_device")

    def bound(self):
        self._check(self.lib.qas_bound(self.ctx), "qas_bound")

    def step(self, dt):
        self._check(self.lib.qas_step(self.ctx, float(dt)), "qas_step")

    def cfl_dt(self):
        d = C.c_double(0.0)
        self._check(self.lib.qas_cfl_dt(self.ctx, C.byref(d)), "qas_cfl_dt")
        return d.value

    def run(self, nsteps, flg=1, dt=None, history=True):
        """exec_-style loop on the device; returns (dt_history, dt_now)."""
        d = C.c_double(dt if dt is not None else 0.0)
        hist = np.zeros(max(nsteps, 1))
        hp = hist.ctypes.data_as(C.POINTER(C.c_double)) if history else None
        self._check(self.lib.qas_run(self.ctx, int(nsteps), int(flg), C.byref(d), hp), "qas_run")
        return hist[:nsteps], d.value

    def reset_counter(self):
        self._check(self.lib.qas_reset_counter(self.ctx), "qas_reset_counter")

    def sync(self):
        self._check(self.lib.qas_sync(self.ctx), "qas_sync")

    def set_dissipation(self, nu, eta, dcmax):
        """Per-cell viscosity / resistivity (arrays of the domain shape, ghosts
        included) and max(2 nu, eta); None switches the dissipative step off."""
        if nu is None:
            self._check(self.lib.qas_set_dissipation(self.ctx, None, None, 0.0), "qas_set_dissipation")
            return
        self._nu = np.ascontiguousarray(nu, dtype=np.float64)
        self._eta = np.ascontiguousarray(eta, dtype=np.float64)
        assert self._nu.shape == self.shape and self._eta.shape == self.shape
        self._check(self.lib.qas_set_dissipation(self.ctx, _p(self._nu), _p(self._eta), float(dcmax)),
                    "qas_set_dissipation")

    def set_gravity(self, phi):
        """Gravitational potential per cell (domain shape, ghosts included);
        None switches gravity off."""
        if phi is None:
            self._check(self.lib.qas_set_gravity(self.ctx, None), "qas_set_gravity")
            return
        self._phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert self._phi.shape == self.shape
        self._check(self.lib.qas_set_gravity(self.ctx, _p(self._phi)), "qas_set_gravity")

    def dsptv(self, dt):
        """One visco-resistive step dsptv(dt)."""
        self._check(self.lib.qas_dsptv(self.ctx, float(dt)), "qas_dsptv")

    def adj_forward(self, dts):
        """Advance len(dts) steps at the given dt values, keeping checkpoints
        for adj_backward (qas_adj_forward).  The context holds the
        checkpoints of its latest adj_forward only; adj_generation counts
        the calls, so a caller that keeps a trajectory for later can tell
        that another forward has replaced it."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        self.adj_generation = getattr(self, "adj_generation", 0) + 1
        self._check(self.lib.qas_adj_forward(self.ctx, int(d.size), _p(d)), "qas_adj_forward")

    def tune(self):
        """Pick the 2D launch shape by timing (qas_tune); returns it."""
        self._check(self.lib.qas_tune(self.ctx), "qas_tune")
        return self.shape_now()

    def set_shape(self, xwpb, ewb):
        """Fix the 2D launch shape: x-sweep wavefronts per workgroup (1, 2, 4),
        element-wise workgroup size (64, 128, 256) (qas_set_shape)."""
        self._check(self.lib.qas_set_shape(self.ctx, int(xwpb), int(ewb)), "qas_set_shape")

    def shape_now(self):
        """The current 2D launch shape (xwpb, ewb) (qas_shape)."""
        a, b = C.c_int(), C.c_int()
        self._check(self.lib.qas_shape(self.ctx, C.byref(a), C.byref(b)), "qas_shape")
        return a.value, b.value

    def adj_keep_stages(self, on=True):
        """Later adj_forward calls keep every stage's intermediates (46 more
        arrays of the state's size per step) and adj_backward skips the
        forward recompute; same results (qas_adj_keep_stages)."""
        self._check(self.lib.qas_adj_keep_stages(self.ctx, 1 if on else 0), "qas_adj_keep_stages")

    def adj_backward(self, g_final):
        """Adjoint of the recorded steps: d(loss)/d(initial state) from
        d(loss)/d(final state), both (8, ...) arrays with ghosts."""
        g = np.ascontiguousarray(g_final, dtype=np.float64)
        assert g.shape == (8,) + self.shape
        out = np.empty_like(g)
        self._check(self.lib.qas_adj_backward(self.ctx, _p(g), _p(out)), "qas_adj_backward")
        return out

    def adj_backward_device(self, g_final, g_init):
        """adj_backward on device-resident buffers (DeviceArray of shape
        (8,) + shape each): no host copies (qas_adj_backward_device)."""
        for a in (g_final, g_init):
            assert a.shape == (8,) + self.shape
        self._check(self.lib.qas_adj_backward_device(self.ctx, g_final.ptr, g_init.ptr), "qas_adj_backward_device")

    @property
    def device_bytes(self):
        return int(self.lib.qas_device_bytes(self.ctx))


_hip = None


def hip_runtime():
    """The HIP runtime libqasmhd_amd is linked against (dlopen of the same
    soname returns the loaded copy)."""
    global _hip
    if _hip is None:
        load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip = h
    return _hip


class DeviceArray:
    """A float64 array in device memory (for the *_device entry points
    without torch)."""

    def __init__(self, shape):
        self.shape = tuple(int(n) for n in shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self._h = hip_runtime()
        p = C.c_void_p()
        if self._h.hipMalloc(C.byref(p), self.nbytes) != 0:
            raise QasError("hipMalloc of %d bytes failed" % self.nbytes)
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        if self._h.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) != 0:
            raise QasError("hipMemcpy to the device failed")

    def download(self):
        out = np.empty(self.shape, dtype=np.float64)
        if self._h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:
            raise QasError("hipMemcpy from the device failed")
        return out

    def close(self):
        if self.ptr:
            self._h.hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many 2D domains advanced together (qas_batch_*): one launch per stage
    kernel for all of them, two steps per graph replay.  `solvers` are
    Solver objects (same scheme), kept alive by the batch."""

    def __init__(self, solvers):
        self.lib = load()
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.ctx for s in self.solvers])
        self.b = C.c_void_p()
        r = self.lib.qas_batch_create(C.byref(self.b), arr, len(self.solvers))
        if r != QAS_OK:
            msg = self.lib.qas_batch_last_error(self.b).decode() if self.b else "qas_batch_create failed"
            self.lib.qas_batch_destroy(self.b)
            self.b = None
            raise QasError("qas_batch_create: %s (%d)" % (msg, r))
        if GRAPHS is not None:
            self.use_graphs(GRAPHS)

    def use_graphs(self, on=True):
        """Replay captured step pairs (True) or launch from the host (False)
        (qas_batch_use_graphs)."""
        r = self.lib.qas_batch_use_graphs(self.b, 1 if on else 0)
        if r != QAS_OK:
            raise QasError("qas_batch_use_graphs (%d)" % r)

    def run(self, nsteps, flg=1, dts=None, history=True):
        """qas_run on every domain; returns (dt_history[n, nsteps], dt_now[n])."""
        n = len(self.solvers)
        d = np.zeros(n) if dts is None else np.array(dts, dtype=np.float64)
        hist = np.zeros((n, max(nsteps, 1)))
        hp = hist.ctypes.data_as(C.POINTER(C.c_double)) if history else None
        r = self.lib.qas_batch_run(self.b, int(nsteps), int(flg), d.ctypes.data_as(C.POINTER(C.c_double)), hp)
        if r != QAS_OK:
            raise QasError("qas_batch_run: %s (%d)" % (self.lib.qas_batch_last_error(self.b).decode(), r))
        return hist[:, :nsteps], d

    def adj_forward(self, dts):
        """qas_adj_forward on every domain in one batched run: dts[i] are
        domain i's step sizes (all the same length); each domain keeps its
        checkpoints (qas_batch_adj_forward)."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        assert d.ndim == 2 and d.shape[0] == len(self.solvers)
        for s in self.solvers:
            s.adj_generation = getattr(s, "adj_generation", 0) + 1
        r = self.lib.qas_batch_adj_forward(self.b, int(d.shape[1]), _p(d))
        if r != QAS_OK:
            raise QasError("qas_batch_adj_forward: %s (%d)" % (self.lib.qas_batch_last_error(self.b).decode(), r))

    def adj_backward(self, g_final):
        """qas_adj_backward on every domain in one batched run: g_final[i]
        the (8, ny, nx) adjoint of domain i's final state; returns the list
        of initial-state adjoints (qas_batch_adj_backward)."""
        n = len(self.solvers)
        gf = [np.ascontiguousarray(g, dtype=np.float64) for g in g_final]
        out = [np.empty_like(g) for g in gf]
        for g, s in zip(gf, self.solvers):
            assert g.shape == (8,) + s.shape
        pf = (C.POINTER(C.c_double) * n)(*[_p(g) for g in gf])
        po = (C.POINTER(C.c_double) * n)(*[_p(g) for g in out])
        r = self.lib.qas_batch_adj_backward(self.b, pf, po)
        if r != QAS_OK:
            raise QasError("qas_batch_adj_backward: %s (%d)" % (self.lib.qas_batch_last_error(self.b).decode(), r))
        return out

    def tune(self):
        """Time the candidate launch shapes and keep the fastest (qas_batch_tune)."""
        r = self.lib.qas_batch_tune(self.b)
        if r != QAS_OK:
            raise QasError("qas_batch_tune: %s (%d)" % (self.lib.qas_batch_last_error(self.b).decode(), r))
        return self.shape()

    def set_shape(self, order, xwpb, ewb):
        r = self.lib.qas_batch_set_shape(self.b, int(order), int(xwpb), int(ewb))
        if r != QAS_OK:
            raise QasError("qas_batch_set_shape: %s (%d)" % (self.lib.qas_batch_last_error(self.b).decode(), r))

    def shape(self):
        """(order, x-sweep wavefronts per workgroup, element-wise workgroup size)."""
        o, x, e = C.c_int(), C.c_int(), C.c_int()
        self.lib.qas_batch_shape(self.b, C.byref(o), C.byref(x), C.byref(e))
        return o.value, x.value, e.value

    def close(self):
        if self.b:
            self.lib.qas_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def solver_for(problem, **over):
    """Solver configured from a problems.Problem (scheme fields overridable)."""
    kw = dict(gam=problem.gam, cfl=problem.cfl, rmn=problem.rmn, odr=problem.odr, r_k=problem.r_k,
              ctw=problem.ctw, dn=problem.dn)
    kw.update(over)
    return Solver(problem.state.shape[1:], problem.dx, problem.dy, problem.dz, **kw)
