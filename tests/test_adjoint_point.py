"""Per-point hand-written adjoints (qas_adjoint.hpp) against finite
differences of the forward device functions, evaluated on the host.

The forward functions are the ones the parity tests pin bit for bit against
the reference (WCNS-4 common/common_interp.c:102-150, characteristic
reconstruction common/mhd_eigen.c:513-665 via common/mhd_func.c:108-161,
HLLD common/mhd_flux.c:435-561, MLAU common/mhd_flux.c:7-192, B-flux face states common/mhd_func.c:163-188,
CUCT weight common/mhd_func.c:56-76).  Each adjoint is checked with a
dot-product test <g, J d> = <J^T g, d> where J d comes from a
Richardson-extrapolated central difference.  Trials where a branch (HLLD
region, positivity fallback, a stiff WCNS weight) changes inside the FD
interval give isolated large errors, so the bar is on quantiles.
"""
import os

import numpy as np
import pytest

from qasmhd_amd import capi

pytestmark = pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="library not built")


@pytest.mark.parametrize("name", ["wcns4", "lin4", "lrfb", "cuct", "hlld", "mlau", "lhlld"])
def test_point_adjoint_smooth(name):
    e = capi.adjoint_selfcheck(name, 2000, seed=11)
    assert np.median(e) < 1e-9, np.median(e)
    assert np.quantile(e, 0.999) < 1e-5, np.quantile(e, 0.999)


def test_point_adjoint_lrstate():
    e = capi.adjoint_selfcheck("lrstate", 2000, seed=11)
    assert np.median(e) < 1e-9, np.median(e)
    assert np.quantile(e, 0.98) < 1e-5, np.quantile(e, 0.98)

