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


def test_rcp_inrange_is_the_division(pts):
    """The reciprocal the WCNS-4 weights take (qas_device.hpp rcp_inrange:
    the estimate, Newton steps and correction of the FP64 division without
    its operand scaling and fix-up) equals 1.0 / d bit for bit over the
    whole range where the kernels use it, 2^-767 <= |d| <= 2^1022:
    log-uniform random values of both signs, the range ends, powers of two
    and significands of all ones."""
    print("case: rcp_inrange vs 1.0 / d", flush=True)
    rng = np.random.default_rng(21)
    e = rng.uniform(-767.0, 1022.0, 2_000_000)
    d = np.exp2(e) * rng.choice([-1.0, 1.0], e.size)
    ends = np.array([2.0 ** -767, 2.0 ** 1022, np.nextafter(2.0 ** 1022, 0), np.nextafter(2.0 ** -767, 1),
                     1.0, 3.0, 24.0, 1e-40, 1e40, 2.0 ** 700, 0.65 * 2.0 ** -700])
    k = np.arange(-767, 1022)
    allones = (2.0 - 2.0 ** -52) * np.exp2(k.astype(float))
    d = np.concatenate([d, ends, -ends, np.exp2(k.astype(float)), allones, -allones])
    fast, div = pts.rcp(d)
    assert np.array_equal(fast.view(np.int64), div.view(np.int64))
    assert np.array_equal(div, 1.0 / d)      # the device division is IEEE (numpy: correctly rounded)


def test_wcns4_across_magnitudes(pts, oracle):
    """WCNS-4 on stencils scaled from 1e-300 to 1e300: the smoothness
    denominators cross the in-range test (2^700) of the device's fast
    reciprocal both ways, and constant stencils hit d = eps.  Bit-exact
    against the oracle's C restatement (plain divisions), NaN where it has
    NaN."""
    print("case: wcns4 over magnitudes", flush=True)
    rng = np.random.default_rng(8)
    base = rng.standard_normal((201 * 64, 5))
    scale = np.repeat(10.0 ** np.arange(-300, 301, 3), 64)[:, None]
    f5 = base * scale
    const = np.repeat(rng.standard_normal((64, 1)), 5, axis=1)
    smooth = np.cumsum(rng.standard_normal((64, 5)) * 1e-3, axis=1) + 1.0
    f5 = np.concatenate([f5, const, smooth, smooth * 1e120, smooth * 1e-120])
    fl, fr = pts.interp(3, f5)
    ol, or_ = oracle.interp(3, f5)
    assert _eq_nan(fl, ol) and _eq_nan(fr, or_)
    ok = ~np.isnan(ol)
    assert np.array_equal(fl[ok].view(np.int64), ol[ok].view(np.int64))
    assert np.array_equal(fr[ok].view(np.int64), or_[ok].view(np.int64))


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
