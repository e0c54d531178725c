"""The C restatement (oracle/qas_oracle.c) against golden vectors written by
the reference's own code (tests/golden/make_golden.py).  Bit-exact: every
comparison is array equality of float64 values."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

U = np.load(os.path.join(GOLDEN, "unit_vectors.npz"))
S1 = np.load(os.path.join(GOLDEN, "steps_1d.npz"))
S2 = np.load(os.path.join(GOLDEN, "steps_2d.npz"))


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 10, 11, 12, 13])
def test_interp(oracle, kind):
    fl, fr = oracle.interp(kind, U["interp_in"])
    ref = U["interp_%d" % kind]
    assert np.array_equal(fl, ref[:, 0]) and np.array_equal(fr, ref[:, 1])


@pytest.mark.parametrize("odr", [1, 2, 3, 4])
def test_fcen_df(oracle, odr):
    assert np.array_equal(oracle.fcen(odr, U["f4_in"]), U["fcen_%d" % odr])
    assert np.array_equal(oracle.df(odr, U["f4_in"]), U["df_%d" % odr])


@pytest.mark.parametrize("odr", [1, 2, 3, 4])
def test_lrstate(oracle, odr):
    vl, vr = oracle.lrstate(odr, U["lrstate_st"], U["lrstate_bn"], float(U["lrstate_gam"]))
    assert np.array_equal(vl, U["lrstate_%d_vl" % odr])
    assert np.array_equal(vr, U["lrstate_%d_vr" % odr])


@pytest.mark.parametrize("rmn", [0, 1, 2, 3])
@pytest.mark.parametrize("g", ["g53", "g2", "g14"])
def test_flux(oracle, rmn, g):
    f = oracle.flux(rmn, U["flux_L"], U["flux_R"], U["flux_bn"], float(U["flux_gam_" + g]), U["flux_dv"])
    ref = U["flux_%d_%s" % (rmn, g)]
    both_nan = np.isnan(f) & np.isnan(ref)
    assert np.array_equal(np.where(both_nan, 0, f), np.where(both_nan, 0, ref))


def test_cuct(oracle):
    assert np.array_equal(oracle.cuct(U["cuct_c"], U["cuct_b"]), U["cuct"])


@pytest.mark.parametrize("rmn", [0, 1, 2, 3])
def test_brio_wu_steps(oracle, rmn):
    st = S1["in"].copy()
    cfg = oracle.make_cfg(st.shape[1:], float(S1["dx"]), gam=float(S1["gam"]), rmn=rmn, dn=S1["dn"])
    dt0 = oracle.cfl_dt(cfg, st.copy())
    assert dt0 == float(S1["dt0"])
    oracle.run(cfg, st, int(S1["steps_r%d" % rmn]), 0, dt=dt0)
    assert np.array_equal(st, S1["out_r%d" % rmn])


@pytest.mark.parametrize("name", [str(n) for n in S2["names"]])
def test_2d_steps(oracle, name):
    c = S2["cfg_" + name]
    rmn, odr, r_k, ctw = (int(v) for v in c[:4])
    dn = np.zeros((3, 8), dtype=int)
    dn[0], dn[1] = c[4:12], c[12:20]
    st = S2["in"].copy()
    cfg = oracle.make_cfg(st.shape[1:], float(S2["dx"]), float(S2["dy"]), gam=float(S2["gam"]), rmn=rmn, odr=odr,
                          r_k=r_k, ctw=ctw, dn=dn)
    hist, _ = oracle.run(cfg, st, int(S2["nsteps"]), 1)
    assert np.array_equal(hist, S2["dt_" + name])
    assert np.array_equal(st, S2["out_" + name])


def test_2d_long(oracle):
    st = S2["long_in"].copy()
    n = st.shape[-1] - 8
    d = 2 * np.pi / n
    cfg = oracle.make_cfg(st.shape[1:], d, d)
    hist, _ = oracle.run(cfg, st, 40, 1)
    assert np.array_equal(hist, S2["long_dt"])
    assert np.array_equal(st, S2["long_out"])
