"""Bit-exact parity at the workloads' own sizes.

* 2D decaying turbulence at 1024^2 (HLLD, WCNS-4, CUCT, RK3, CFL 0.4; the
  benchmark workload) and the low-Mach Kelvin-Helmholtz layer at 2048^2
  (MLAU) against the reference's own 2D driver (MHD2D::ideal/bound/setdt,
  2D/SERIAL/OTvortex/mhd2d_solve.cpp:3-400, mhd2d_class.cpp:11-33) built
  from its sources in oracle/_ref and run on the host cores: every FP64
  value of the final state and the whole dt history are equal.
* 3D at 72^3 against the oracle: large enough that the z-marching EMF and
  CT kernels run several 32-plane chunks (window re-seeded at k0 > 3) and
  two 64-wide x tiles, with CUCT weights on and off and open/mirror walls.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ref_binary(p):
    from oracle import build as obuild
    nx, ny = p.interior[1], p.interior[0]
    path = os.path.join(obuild.REFOUT, obuild.ref_binary_name(2, nx, ny, p.rmn, p.odr, p.r_k, p.ctw, p.cfl))
    if not os.access(path, os.X_OK):
        pytest.skip("reference 2D driver for %dx%d rmn %d not built (oracle/_ref; __graft_entry__.build())"
                    % (nx, ny, p.rmn))
    return path


def _gpu_run(p, nsteps):
    from qasmhd_amd import capi
    with capi.solver_for(p) as s:
        s.upload(p.state)
        hist, _ = s.run(nsteps, flg=1)
        return s.download(), hist


@pytest.mark.parametrize("name,nsteps", [("turb2d_1024", 20), ("khi_2048_mlau", 10), ("turb2d_1024_walls", 6)])
def test_2d_full_size_bit_exact_vs_reference(name, nsteps):
    """(The walls case: open x, mirror y with the odd components flipped,
    bx / by staggered: the ghost fill inside the CT launch and the
    two-corner EMF at full size against the reference's bound().)"""
    from oracle import build as obuild
    from qasmhd_amd import problems
    if name == "turb2d_1024":
        p = problems.turbulence2d(n=1024, seed=2)
    elif name == "turb2d_1024_walls":
        p = problems.turbulence2d(n=1024, seed=2)
        p.dn = np.array([[2] * 8, [1, 1, -1, 1, 1, -1, 1, 1], [0] * 8])
    else:
        p = problems.kelvin_helmholtz(n=2048)
        assert p.rmn == 3
    binary = _ref_binary(p)
    _say("full size vs reference:", name, p.interior, "steps", nsteps)
    out, hist = _gpu_run(p, nsteps)
    _say("  gpu done; reference driver on %d host threads" % THREADS)
    ref, ref_hist = obuild.run_ref(binary, p.state, p.dx, p.gam, nsteps, 1, dnx=list(p.dn[0]), dny=list(p.dn[1]),
                                   dy=p.dy, threads=THREADS)
    assert np.isfinite(ref).all()
    assert ref_hist.shape == hist.shape and np.array_equal(hist, ref_hist)
    neq = int(np.count_nonzero(out != ref))
    assert neq == 0, "%d of %d values differ (max |diff| %g)" % (neq, out.size, np.abs(out - ref).max())


@pytest.mark.parametrize("ctw,bc", [(1, "periodic"), (0, "mixed")])
def test_3d_chunks_and_tiles_vs_oracle(oracle, ctw, bc):
    from qasmhd_amd import capi, problems
    n = 72
    p = problems.turbulence3d(n=n, seed=8, nmodes=12)
    dn = np.zeros((3, 8), dtype=int)
    if bc == "mixed":
        dn[0] = [2] * 8
        dn[2] = [1, 1, 1, -1, 1, 1, -1, 1]
    # z-marching kernels: 32-plane chunks over nz - 5 edges, 64-wide x tiles
    assert (n + 8) - 5 > 2 * 32 and (n + 8) - 5 > 64
    _say("3d chunks vs oracle", n, ctw, bc)
    cfg = oracle.make_cfg(p.state.shape[1:], p.dx, p.dy, p.dz, gam=p.gam, ctw=ctw, dn=dn)
    ref = p.state.copy()
    h_ref, dt_ref = oracle.run(cfg, ref, 4, 1)
    with capi.solver_for(p, ctw=ctw, dn=dn) as s:
        s.upload(p.state)
        hist, dt_now = s.run(4, flg=1)
        out = s.download()
    assert np.isfinite(ref).all()
    assert np.array_equal(hist, h_ref) and dt_now == dt_ref
    neq = int(np.count_nonzero(out != ref))
    assert neq == 0, "%d values differ" % neq
