"""The allocation-free run qas_run_static (device state, static scratch)
against qas_run on the same initial data: every FP64 value of the state and
the dt equal, on calls in a row with shrinking domains (the static scratch
keeps the previous call's contents), a continued step counter, R_K 1 (the
state ends in a scratch buffer and is copied home), 3D and the mixed path;
and the size limit rejected with a message before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _hip():
    """The HIP runtime the library is linked against (for the test's own
    device buffers; the entry point under test allocates nothing)."""
    from qasmhd_amd import capi
    capi.load()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _static_run(p, nsteps, dt=0.0, nstep0=0, flg=1, **over):
    """qas_run_static on a device copy of p.state; returns (state, dt)."""
    from qasmhd_amd import capi
    lib = capi.load()
    h = _hip()
    kw = dict(gam=p.gam, cfl=p.cfl, rmn=p.rmn, odr=p.odr, r_k=p.r_k, ctw=p.ctw, dn=p.dn)
    kw.update(over)
    g = capi.make_grid(p.state.shape[1:], p.dx, p.dy, p.dz)
    s = capi.make_scheme(**kw)
    host = np.ascontiguousarray(p.state, dtype=np.float64)
    d_state, d_dt = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(d_state), host.nbytes) == 0
    assert h.hipMalloc(C.byref(d_dt), 8) == 0
    try:
        assert h.hipMemcpy(d_state, host.ctypes.data, host.nbytes, H2D) == 0
        err = C.create_string_buffer(512)
        r = lib.qas_run_static(C.byref(g), C.byref(s), d_state, int(nsteps), int(nstep0), int(flg), float(dt),
                               d_dt, None, err, 512)
        assert r == 0, err.value.decode()
        assert h.hipDeviceSynchronize() == 0
        out = np.empty_like(host)
        dto = np.zeros(1)
        assert h.hipMemcpy(out.ctypes.data, d_state, host.nbytes, D2H) == 0
        assert h.hipMemcpy(dto.ctypes.data, d_dt, 8, D2H) == 0
    finally:
        h.hipFree(d_state)
        h.hipFree(d_dt)
    return out, float(dto[0])


def _ctx_run(p, runs, **over):
    """qas_run on a context: runs = list of step counts of successive calls."""
    from qasmhd_amd import capi
    out = []
    with capi.solver_for(p, **over) as s:
        s.upload(p.state)
        dt = None
        for n in runs:
            _, dt = s.run(n, flg=1, dt=dt)
            out.append((s.download(), dt))
    return out


def test_static_run_matches_qas_run_on_calls_in_a_row():
    from qasmhd_amd import problems
    # 1) 2D 48x40 HLLD / WCNS-4 / RK3, 6 steps, dt from setdt
    p = problems.turbulence2d(n=48, ny=40, seed=21, vrms=0.6)
    _say("case: static 2D 48x40 rk3 6 steps")
    st, dt = _static_run(p, 6)
    ref = _ctx_run(p, [6])[0]
    assert np.array_equal(st, ref[0]) and dt == ref[1]
    # 2) a smaller domain right after (scratch holds the larger call's data):
    #    32x24, R_K 1 over 5 steps -- the state ends in a scratch buffer
    q = problems.turbulence2d(n=32, ny=24, seed=22, vrms=0.5)
    _say("case: static 2D 32x24 rk1 5 steps")
    st, dt = _static_run(q, 5, r_k=1)
    ref = _ctx_run(q, [5], r_k=1)[0]
    assert np.array_equal(st, ref[0]) and dt == ref[1]
    # 3) smaller again, two calls continuing the step counter (setdt after
    #    even steps of the running count) against two qas_run calls
    q2 = problems.turbulence2d(n=24, ny=24, seed=23, vrms=0.7)
    _say("case: static 2D 24x24 4 + 3 steps")
    a, dta = _static_run(q2, 3)
    q2b = q2.copy()
    q2b.state = a
    b, dtb = _static_run(q2b, 3, dt=dta, nstep0=3)
    ref = _ctx_run(q2, [3, 3])
    assert np.array_equal(a, ref[0][0]) and dta == ref[0][1]
    assert np.array_equal(b, ref[1][0]) and dtb == ref[1][1]


@pytest.mark.parametrize("case", ["3d", "mixed"])
def test_static_run_3d_and_mixed(case):
    from qasmhd_amd import problems
    if case == "3d":
        p = problems.turbulence3d(n=16, seed=5)
        over = {}
    else:
        p = problems.turbulence2d(n=40, ny=32, seed=24, vrms=0.6)
        over = {"precision": 1}
    _say("case: static", case)
    st, dt = _static_run(p, 4, **over)
    ref = _ctx_run(p, [4], **over)[0]
    assert np.array_equal(st, ref[0]) and dt == ref[1]


def test_static_run_rejects_oversize_domain():
    from qasmhd_amd import capi
    lib = capi.load()
    n = 1024 + 8
    g = capi.make_grid((n, n), 1.0 / 1024, 1.0 / 1024)
    s = capi.make_scheme()
    err = C.create_string_buffer(512)
    # the size check comes before any device work: a dummy pointer is never touched
    r = lib.qas_run_static(C.byref(g), C.byref(s), C.c_void_p(0x1000), 1, 0, 1, 0.0, None, None, err, 512)
    assert r == -1
    assert b"static array holds" in err.value
