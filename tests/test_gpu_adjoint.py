"""Hand-written discrete adjoint of the 2D step (qasmhd_amd/csrc/
qas_adjoint2d.hip) against central finite differences of the GPU forward.

Forward: n steps of MHD2D::ideal (2D/SERIAL/OTvortex/mhd2d_solve.cpp:3-400)
at frozen dt, bit-identical to the reference (test_gpu_parity.py).  Loss
L(U0) = sum(W * U_n) for a random weight W; the directional derivative
<dL/dU0, d> from the adjoint is compared with a Richardson-extrapolated
central difference of the forward (h = 1e-7 of the variable's scale, small
enough that no min/max selection of the HLLD wave speeds changes branch
inside the interval), per variable, at Mach 0.3 (all cells on
the subsonic projection branch, as the inverse problem runs), with smooth
nonzero vz, bz.  Tolerance 1e-5 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NG = 4


def _with_vz_bz(p, a=0.1):
    """Adds smooth vz, bz (cell-centred) to a 2D state.  With vz = bz = 0
    exactly, the reference's exact-zero flags (common/mhd_eigen.c:575-587)
    and WCNS's eps = 1e-40 make the step non-differentiable in those
    components (the directional derivative at 0 is not linear in the
    direction), so a finite difference cannot be compared there."""
    st = p.state
    ny, nx = st.shape[1:]
    x = (np.arange(nx) - NG + 0.5) * p.dx
    y = (np.arange(ny) - NG + 0.5) * p.dy
    X, Y = np.meshgrid(x, y)
    ro = st[0]
    vz = a * np.sin(X + 2.0 * Y + 0.3)
    bz = a * np.cos(2.0 * X - Y + 0.7)
    st[7] += 0.5 * ro * vz * vz + 0.5 * bz * bz - 0.5 * st[3] ** 2 / ro - 0.5 * st[6] ** 2
    st[3] = ro * vz
    st[6] = bz
    return p


def _setup(n=32, seed=5, ctw=1, ny=None, rmn=1, dn=None):
    from qasmhd_amd import capi, problems
    p = _with_vz_bz(problems.turbulence2d(n=n, seed=seed, vrms=0.3, brms=0.3, ny=ny))
    if dn is not None:
        p.dn = np.array([dn[0], dn[1], [0] * 8])
    s = capi.solver_for(p, ctw=ctw, rmn=rmn)
    s.upload(p.state)
    hist, _ = s.run(3, flg=1)          # a realistic dt schedule
    return p, s, hist


def _forward(s, u0, dts):
    s.upload(u0)
    s.bound()
    for dt in dts:
        s.step(dt)
    return s.download()


def _dir_deriv(s, u0, dts, W, d, h):
    def D(hh):
        up = _forward(s, u0 + hh * d, dts)
        um = _forward(s, u0 - hh * d, dts)
        return float(np.sum(W * (up - um))) / (2.0 * hh)
    return (4.0 * D(h / 2) - D(h)) / 3.0


# (nx, ny): the y line pass runs on transposed copies and reads the EMF-part
# adjoints in place, so a non-square domain checks both index maps.  rmn 3
# (MLAU) and 2 (LHLLD) add the shock sensor: dv0 along the line, dv1 from
# the transverse velocity differences (gathered onto cells of the
# neighbouring lines).
# boundary flags (dnxs, dnys of the reference drivers): 2 open, +-1 mirror
# (-1: the normal velocity and normal B odd), 0 periodic; bx is staggered
# along x and by along y (the default st flags)
WALLS = {"open_x_mirror_y": ([2] * 8, [1, 1, -1, 1, 1, -1, 1, 1]),
         "mirror_x_periodic_y": ([1, -1, 1, 1, -1, 1, 1, 1], [0] * 8)}


@pytest.mark.parametrize("walls,ctw,nx,ny", [("open_x_mirror_y", 1, 32, 32), ("open_x_mirror_y", 0, 40, 24),
                                             ("mirror_x_periodic_y", 1, 24, 40)])
def test_adjoint_vs_fd_walls(walls, ctw, nx, ny):
    """Non-periodic boundaries (HLLD): the adjoint of bound() gathers every
    ghost onto its source with the ghost's factor (open boundaries copy the
    edge cell into all four ghosts, mirrors flip the odd components)."""
    print("case: adjoint vs FD, %s, %dx%d, 3 steps, ctw %d" % (walls, nx, ny, ctw), flush=True)
    p, s, dts = _setup(n=nx, ny=ny, ctw=ctw, rmn=1, dn=WALLS[walls])
    rng = np.random.default_rng(13)
    u0 = p.state.copy()
    W = rng.standard_normal(u0.shape)
    s.upload(u0)
    s.adj_forward(dts)
    g = s.adj_backward(W)
    errs = []
    for m in range(8):
        # directions on the interior (the forward overwrites ghosts)
        d = np.zeros_like(u0)
        d[m][NG:-NG, NG:-NG] = rng.standard_normal((u0.shape[1] - 2 * NG, u0.shape[2] - 2 * NG))
        scale = np.abs(u0[m][NG:-NG, NG:-NG]).max() + 1e-3
        ad = float(np.sum(g * d))
        fd = _dir_deriv(s, u0, dts, W, d, 1e-7 * scale)
        err = abs(ad - fd) / max(abs(ad), abs(fd), 1e-300)
        print("  var %d: adjoint %.12e  fd %.12e  rel %.2e" % (m, ad, fd, err), flush=True)
        errs.append(err)
    s.close()
    assert max(errs) < 1e-5, errs


def test_adjoint_walls_need_hlld():
    # LHLLD and MLAU keep periodic boundaries only (their shock sensor's adjoint)
    from qasmhd_amd import capi
    p, s, dts = _setup(n=24, rmn=3, dn=WALLS["open_x_mirror_y"])
    s.upload(p.state)
    with pytest.raises(capi.QasError, match="HLLD"):
        s.adj_forward(dts)
    s.close()


@pytest.mark.parametrize("rmn,ctw,nx,ny", [(1, 1, 32, 32), (1, 0, 32, 32), (1, 1, 40, 24), (1, 1, 24, 40),
                                           (3, 1, 32, 32), (3, 0, 32, 32), (3, 1, 40, 24), (3, 1, 24, 40),
                                           (2, 1, 32, 32), (2, 1, 40, 24)])
def test_adjoint_vs_fd(rmn, ctw, nx, ny):
    print("case: adjoint vs FD, rmn %d, %dx%d, 3 steps, ctw %d" % (rmn, nx, ny, ctw), flush=True)
    p, s, dts = _setup(n=nx, ny=ny, ctw=ctw, rmn=rmn)
    rng = np.random.default_rng(11)
    u0 = p.state.copy()
    W = rng.standard_normal(u0.shape)
    s.upload(u0)
    s.adj_forward(dts)
    g = s.adj_backward(W)
    # ghost entries of the input adjoint are zero (bound() overwrites them)
    inner = np.zeros(u0.shape[1:], bool)
    inner[NG:-NG, NG:-NG] = True
    assert np.all(g[:, ~inner] == 0.0)
    errs = []
    for m in range(8):
        d = np.zeros_like(u0)
        d[m][inner] = rng.standard_normal(inner.sum())
        scale = np.abs(u0[m][inner]).max() + 1e-3
        ad = float(np.sum(g * d))
        fd = _dir_deriv(s, u0, dts, W, d, 1e-7 * scale)
        err = abs(ad - fd) / max(abs(ad), abs(fd), 1e-300)
        print("  var %d: adjoint %.12e  fd %.12e  rel %.2e" % (m, ad, fd, err), flush=True)
        errs.append(err)
    s.close()
    assert max(errs) < 1e-5, errs


@pytest.mark.parametrize("rmn,ctw,nx,ny", [(1, 1, 40, 24), (3, 1, 32, 32), (1, 0, 24, 40)])
def test_kept_stages_equal_recompute(rmn, ctw, nx, ny):
    """qas_adj_keep_stages: the recorded forward keeps every stage's
    intermediates and the backward reads them instead of recomputing the
    three stages of each step.  The state after the recorded forward equals
    qas_step's, and the gradients equal those of the checkpoint-and-
    recompute path bit for bit, for a repeated backward and for a second,
    shorter recorded forward on the same store."""
    print("case: kept stages vs recompute, rmn %d ctw %d, %dx%d" % (rmn, ctw, nx, ny), flush=True)
    p, s, dts = _setup(n=nx, ny=ny, ctw=ctw, rmn=rmn)
    dts = np.concatenate([dts, dts[::-1]])
    W = np.random.default_rng(12).standard_normal(p.state.shape)
    s.upload(p.state)
    s.adj_forward(dts)
    fin_a = s.download()
    g_a = s.adj_backward(W)
    s.upload(p.state)
    s.adj_forward(dts[:3])
    g_short_a = s.adj_backward(W)
    fin_step = _forward(s, p.state, dts)
    s.adj_keep_stages(True)
    s.upload(p.state)
    s.adj_forward(dts)
    fin_b = s.download()
    g_b = s.adj_backward(W)
    g_b2 = s.adj_backward(W)
    s.upload(p.state)
    s.adj_forward(dts[:3])
    g_short_b = s.adj_backward(W)
    s.close()
    assert np.array_equal(fin_a, fin_step)
    assert np.array_equal(fin_b, fin_step)
    assert np.array_equal(g_a, g_b)
    assert np.array_equal(g_b, g_b2)
    assert np.array_equal(g_short_a, g_short_b)
    assert not np.array_equal(g_a, g_short_a)


def test_adjoint_linear_in_seed():
    # the adjoint is linear in the final-state weight
    print("case: adjoint linearity", flush=True)
    p, s, dts = _setup(n=24, seed=7)
    rng = np.random.default_rng(3)
    W1 = rng.standard_normal(p.state.shape)
    W2 = rng.standard_normal(p.state.shape)
    s.upload(p.state)
    s.adj_forward(dts[:2])
    g1 = s.adj_backward(W1)
    g2 = s.adj_backward(W2)
    g12 = s.adj_backward(2.0 * W1 - 3.0 * W2)
    s.close()
    assert np.allclose(g12, 2.0 * g1 - 3.0 * g2, rtol=1e-12, atol=1e-12 * np.abs(g12).max())


def test_adjoint_device_entry_matches_host():
    # qas_adj_backward_device on device-resident buffers gives the host entry's result
    from qasmhd_amd import capi
    print("case: device-resident backward vs host entry, 40x24", flush=True)
    p, s, dts = _setup(n=40, ny=24, seed=9)
    W = np.random.default_rng(4).standard_normal(p.state.shape)
    s.upload(p.state)
    s.adj_forward(dts)
    g = s.adj_backward(W)
    with capi.DeviceArray(W.shape) as dg, capi.DeviceArray(W.shape) as di:
        dg.upload(W)
        s.adj_backward_device(dg, di)
        gd = di.download()
    s.close()
    assert np.array_equal(g, gd)


def test_adjoint_rejects_unsupported():
    from qasmhd_amd import capi, problems
    p = problems.turbulence2d(n=16, seed=1)
    s = capi.solver_for(p, rmn=0)
    s.upload(p.state)
    with pytest.raises(capi.QasError):
        s.adj_forward([1e-3])
    s.close()


def test_torch_autograd_function():
    """The torch.autograd.Function wrapper, in a fresh process: torch ships
    its own libamdhip64 (same soname as ROCm's), so it must be imported
    before the library is loaded; this test process has loaded it already."""
    import os
    import subprocess
    import sys
    print("case: torch autograd wrapper (subprocess)", flush=True)
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_torch_adjoint_check.py")
    r = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0


def _ref_binary(rmn, n):
    """The reference's own 2D driver (2D/SERIAL/OTvortex/mhd2d_{solve,class}.cpp,
    built from source by oracle/build.py where the sources lie); None where
    it was not built (the reference sources are absent on the GPU host, so
    the binary travels prebuilt)."""
    import os
    from oracle import build as obuild
    path = os.path.join(obuild.REFOUT, obuild.ref_binary_name(2, n, n, rmn, 4, 3, 1, 0.4))
    return path if os.access(path, os.X_OK) else None


@pytest.mark.parametrize("rmn,walls", [(1, None), (3, None), (1, "open_x_mirror_y")])
def test_adjoint_vs_reference_fd(rmn, walls):
    """The adjoint gradient against central finite differences of the
    reference's own CPU code: its 2D driver run at a fixed dt (exec_ with
    flg = 0: bound, then MHD2D::ideal(dt), every step) on U0 +- h d, with
    periodic boundaries or the reference's bound() of open x / mirror y."""
    from oracle import build as obuild
    n = 32
    binary = _ref_binary(rmn, n)
    if binary is None:
        pytest.skip("reference 2D driver %dx%d rmn %d not built" % (n, n, rmn))
    print("case: adjoint vs FD of the reference driver, rmn %d, %dx%d, 3 steps, walls %s"
          % (rmn, n, n, walls), flush=True)
    dn = WALLS[walls] if walls else ([0] * 8, [0] * 8)
    p, s, hist = _setup(n=n, rmn=rmn, dn=None if walls is None else dn)
    dts = [float(hist[0])] * 3
    rng = np.random.default_rng(17)
    u0 = p.state.copy()
    W = rng.standard_normal(u0.shape)
    s.upload(u0)
    s.adj_forward(dts)
    g = s.adj_backward(W)
    s.close()
    inner = np.zeros(u0.shape[1:], bool)
    inner[NG:-NG, NG:-NG] = True

    def ref_final(u):
        out, _ = obuild.run_ref(binary, u, p.dx, p.gam, len(dts), 0, dt=dts[0], dnx=dn[0], dny=dn[1],
                                dy=p.dy, threads=1)
        return out

    def D(d, hh):
        return float(np.sum(W * (ref_final(u0 + hh * d) - ref_final(u0 - hh * d)))) / (2.0 * hh)

    errs = []
    for m in range(8):
        d = np.zeros_like(u0)
        d[m][inner] = rng.standard_normal(inner.sum())
        h = 1e-7 * (np.abs(u0[m][inner]).max() + 1e-3)
        fd = (4.0 * D(d, h / 2) - D(d, h)) / 3.0
        ad = float(np.sum(g * d))
        err = abs(ad - fd) / max(abs(ad), abs(fd), 1e-300)
        print("  var %d: adjoint %.12e  reference fd %.12e  rel %.2e" % (m, ad, fd, err), flush=True)
        errs.append(err)
    assert max(errs) < 1e-5, errs
