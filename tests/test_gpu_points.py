"""Device per-point functions against vectors written by the reference's own
functions (tests/golden/unit_vectors.npz, made by tests/golden/make_golden.py
through oracle/ref/unit.cpp).  One GPU thread per vector, through the same
device functions the stage kernels inline (qas_device.hpp); bit-exact:
array equality of float64 values (NaN where the reference gives NaN).

References: interpol / l_interp common/common_interp.c:7-150; fcen / df1
common/common_interp.h:10-24; mhd_lrstate common/mhd_func.c:108-161 with
mhd_m_reconst / mhd_a_reconst common/mhd_eigen.c:513-665; riemann[RMN]
common/mhd_flux.c:7-698; mhd_cuct_weight common/mhd_func.c:56-76."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

U = np.load(os.path.join(GOLDEN, "unit_vectors.npz"))


def _eq_nan(a, b):
    both = np.isnan(a) & np.isnan(b)
    return np.array_equal(np.where(both, 0, a), np.where(both, 0, b))


@pytest.fixture(scope="module")
def pts():
    from qasmhd_amd import capi
    return capi.points


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 10, 11, 12, 13])
def test_interp(pts, kind):
    print("case: interp kind", kind, flush=True)
    fl, fr = pts.interp(kind, U["interp_in"])
    ref = U["interp_%d" % kind]
    assert np.array_equal(fl, ref[:, 0]) and np.array_equal(fr, ref[:, 1])


@pytest.mark.parametrize("odr", [1, 2, 3, 4])
def test_fcen_df(pts, odr):
    print("case: fcen/df odr", odr, flush=True)
    assert np.array_equal(pts.f4(0, odr, U["f4_in"]), U["fcen_%d" % odr])
    assert np.array_equal(pts.f4(1, odr, U["f4_in"]), U["df_%d" % odr])


@pytest.mark.parametrize("odr", [1, 2, 3, 4])
def test_lrstate(pts, odr):
    print("case: lrstate odr", odr, flush=True)
    vl, vr = pts.lrstate(odr, U["lrstate_st"], U["lrstate_bn"], float(U["lrstate_gam"]))
    assert np.array_equal(vl, U["lrstate_%d_vl" % odr])
    assert np.array_equal(vr, U["lrstate_%d_vr" % odr])


@pytest.mark.parametrize("rmn", [0, 1, 2, 3])
@pytest.mark.parametrize("g", ["g53", "g2", "g14"])
def test_flux(pts, rmn, g):
    print("case: flux rmn", rmn, g, flush=True)
    f = pts.flux(rmn, U["flux_L"], U["flux_R"], U["flux_bn"], float(U["flux_gam_" + g]), U["flux_dv"])
    assert _eq_nan(f, U["flux_%d_%s" % (rmn, g)])


def test_cuct(pts):
    print("case: cuct", flush=True)
    assert np.array_equal(pts.cuct(U["cuct_c"], U["cuct_b"]), U["cuct"])
