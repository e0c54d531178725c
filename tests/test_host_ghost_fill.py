"""The CT launch with the fused ghost fill (qas_stage2d.hpp ctb2d_at), run
on the host through the kernels' own __host__ __device__ functions
(qas_test_ctb2d_host), against CT followed by the plain bound()
(qas_test_bound_host, the k_bound body; common/common_func.c:34-138).

The fused launch splits the work: CT threads update the kept bx / by faces
and write the ghost images of each face they update; ghost threads fill the
cell-centred ghosts and the bx / by ghosts whose source no CT thread
updates.  Every ghost must be written once, with bound()'s source and
factors, and no thread may read a cell another thread of the launch
writes.  Running the two parts in both orders catches such a read: the
stale output values differ from the updated ones, so one order would
disagree with the reference.  With zero EMF each CT face becomes
a u0 + b in; in the last RK stage u0 is the output buffer itself (alias).
No GPU is used."""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

NG = 4
# (x flags, y flags) of the 8 variables: 0 periodic, +-1 mirror (odd for
# -1), +-2 open / zero-fixed; stagger flags: bx along x, by along y
BCS = {
    "periodic": ([0] * 8, [0] * 8),
    "open_x_mirror_y": ([2] * 8, [1, 1, -1, 1, 1, -1, 1, 1]),
    "mirror_x_periodic_y": ([1, -1, 1, 1, -1, 1, 1, 1], [0] * 8),
    "zero_x_open_y": ([-2] * 8, [2] * 8),
    "mirror_both": ([1, -1, 1, 1, -1, 1, 1, -1], [1, 1, -1, 1, 1, -1, 1, -1]),
}
ST_DEFAULT = ([0, 0, 0, 0, 1, 0, 0, 0], [0, 0, 0, 0, 0, 1, 0, 0])
ST_NONE = ([0] * 8, [0] * 8)


@pytest.fixture(scope="module")
def lib():
    from qasmhd_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libqasmhd_amd.so not built here")
    l = C.CDLL(capi.LIB_PATH)
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    l.qas_test_bound_host.argtypes = [C.c_int, C.c_int, ip, ip, dp]
    l.qas_test_ctb2d_host.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_double, C.c_int, C.c_int]
    return l


def _ip(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ct_faces(nx, ny):
    """Masks of the faces CT updates: bx for i in [4, nx-4], j in [4, ny-5];
    by for i in [4, nx-5], j in [4, ny-4]."""
    mx = np.zeros((ny, nx), bool)
    my = np.zeros((ny, nx), bool)
    mx[4:ny - 4, 4:nx - 3] = True
    my[4:ny - 3, 4:nx - 4] = True
    return mx, my


@pytest.mark.parametrize("bcname,stname,nx,ny", [(b, s, nx, ny) for b in BCS for s in ("default", "none")
                                                 for nx, ny in ((24, 20), (25, 31), (16, 16), (16, 19))])
def test_fused_ct_ghost_fill_matches_ct_then_bound(lib, bcname, stname, nx, ny):
    dn = np.array(BCS[bcname], dtype=np.int32).ravel()
    st = np.array(ST_DEFAULT if stname == "default" else ST_NONE, dtype=np.int32).ravel()
    rng = np.random.default_rng(nx * 31 + ny)
    shape = (8, ny, nx)
    inp = rng.standard_normal(shape)
    u0 = rng.standard_normal(shape)
    out0 = rng.standard_normal(shape)   # the stage output before the launch (x sweep results, stale faces)
    a, b = 0.25, 0.75
    mx, my = _ct_faces(nx, ny)
    for alias, ghosts_first in itertools.product((0, 1), (0, 1)):
        base = out0 if alias else u0
        ref = out0.copy()
        ref[4][mx] = a * base[4][mx] + b * inp[4][mx]
        ref[5][my] = a * base[5][my] + b * inp[5][my]
        assert lib.qas_test_bound_host(nx, ny, _ip(dn), _ip(st), _dp(ref)) == 0
        got = out0.copy()
        assert lib.qas_test_ctb2d_host(nx, ny, _ip(dn), _ip(st), _dp(inp), _dp(u0), _dp(got), a, b, alias,
                                       ghosts_first) == 0
        bad = np.argwhere(got != ref)
        assert bad.size == 0, (alias, ghosts_first, bad[:8].tolist())


def test_bound_host_periodic_images(lib):
    """The reference fill itself on a closed form: periodic in both axes,
    every ghost equals its interior image."""
    nx, ny = 20, 18
    dn = np.zeros(16, np.int32)
    st = np.array(ST_DEFAULT, np.int32).ravel()
    f = np.random.default_rng(3).standard_normal((8, ny, nx))
    assert lib.qas_test_bound_host(nx, ny, _ip(dn), _ip(st), _dp(f)) == 0
    ii = (np.arange(nx) - NG) % (nx - 2 * NG) + NG
    jj = (np.arange(ny) - NG) % (ny - 2 * NG) + NG
    assert np.array_equal(f, f[:, jj][:, :, ii])


def test_host_runs_reject_bad_flags(lib):
    dn = np.zeros(16, np.int32)
    st = np.zeros(16, np.int32)
    f = np.zeros((8, 20, 20))
    bad = dn.copy()
    bad[3] = 5
    assert lib.qas_test_bound_host(20, 20, _ip(bad), _ip(st), _dp(f)) != 0
    assert lib.qas_test_bound_host(12, 20, _ip(dn), _ip(st), _dp(f)) != 0
