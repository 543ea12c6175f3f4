"""Pins of the oracle's restatement of GFS_rrtmgp_pre's input preparation:
cmp_tlev against the reference's own Fortran (oracle/_ref), in both vertical
orders; the bounds, the water-vapour floor and the mixing-ratio conversions
by their invariants."""
import ctypes as C

import numpy as np
import pytest

AMDW = 28.9644 / 18.0154
AMDO3 = 28.9644 / 47.9982


def prep_inputs(kd, ncol=9, nlay=40, top=False, seed=3):
    """Synthetic caller fields with values outside the table range, a vapour
    below the floor, a saturated layer (q = 1) and negative ozone."""
    from ccpprad import synthetic as syn
    pr = syn.make_profiles(ncol, nlay, seed=seed, clouds=False, kd=kd)
    rng = np.random.default_rng(seed)
    prsi, prsl, tgrs = pr.p_lev.copy(), pr.p_lay.copy(), pr.t_lay.copy()
    q = rng.uniform(0.0, 0.02, (nlay, ncol))
    q[0, 0], q[1, 0] = 1e-9, 1.0
    o3 = rng.uniform(-1e-8, 1e-5, (nlay, ncol))
    tmin, tmax = float(kd.temp_ref[0]), float(kd.temp_ref[-1])
    pmin, pmax = float(np.min(kd.press_ref)), float(np.max(kd.press_ref))
    tgrs[2, 1], tgrs[3, 1], tgrs[4, 1], tgrs[5, 1] = tmin - 5.0, tmax + 5.0, tmin, tmax
    prsl[5, 2], prsl[nlay - 1, 2] = pmax * 1.01, pmin * 0.5
    tsfc = pr.tsfc.copy()
    tsfc[3] = tmax + 7.0
    if top:
        prsi, prsl, tgrs, q, o3 = prsi[::-1], prsl[::-1], tgrs[::-1], q[::-1], o3[::-1]
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return dict(prsi=c(prsi), prsl=c(prsl), tgrs=c(tgrs), q=c(q), o3=c(o3), tsfc=c(tsfc)), (pmin, pmax, tmin, tmax)


@pytest.mark.parametrize("top", [False, True])
def test_tlev_matches_reference_cmp_tlev(kd_lw, top):
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    x, (pmin, pmax, tmin, tmax) = prep_inputs(kd_lw, top=top)
    out, t1 = O.prepare(kd_lw, x["prsi"], x["prsl"], x["tgrs"], x["q"], x["o3"], x["tsfc"])
    assert t1 == top
    nl, nc = x["prsl"].shape
    ref = np.zeros((nl + 1, nc))
    P = C.POINTER(C.c_double)
    R.ref_cmp_tlev(nc, nl, pmin, out["p_lay"].ctypes.data_as(P), out["t_lay"].ctypes.data_as(P),
                   out["p_lev"].ctypes.data_as(P), x["tsfc"].ctypes.data_as(P), ref.ctypes.data_as(P))
    ref = np.where(ref <= tmin, tmin + np.finfo(float).eps, ref)
    ref = np.where(ref >= tmax, tmax - np.finfo(float).eps, ref)
    assert np.array_equal(out["t_lev"], ref)


@pytest.mark.parametrize("top", [False, True])
def test_bounds_floor_and_mixing_ratios(kd_lw, top):
    from oracle import oracle as O
    x, (pmin, pmax, tmin, tmax) = prep_inputs(kd_lw, top=top)
    out, _ = O.prepare(kd_lw, x["prsi"], x["prsl"], x["tgrs"], x["q"], x["o3"], x["tsfc"])
    for k, lo, hi in (("t_lay", tmin, tmax), ("p_lay", pmin, pmax), ("t_lev", tmin, tmax)):
        assert (out[k] >= lo).all() and (out[k] <= hi).all(), k
    inside = (x["tgrs"] > tmin) & (x["tgrs"] < tmax)
    assert np.array_equal(out["t_lay"][inside], x["tgrs"][inside])
    inside = (x["prsl"] > pmin) & (x["prsl"] < pmax)
    assert np.array_equal(out["p_lay"][inside], x["prsl"][inside])
    assert np.array_equal(out["p_lev"], x["prsi"])
    assert out["q_lay"].min() == 1e-6
    # vmr = q / (1 - q) amd / amw inverts to q = vmr / (amd / amw + vmr)
    ok = out["q_lay"] != 1.0
    qb = out["vmr_h2o"][ok] / (AMDW + out["vmr_h2o"][ok])
    assert np.allclose(qb, out["q_lay"][ok], rtol=1e-13, atol=0)
    assert (out["vmr_h2o"][~ok] == 0).all() and (~ok).sum() == 1
    assert np.allclose(out["vmr_o3"] / AMDO3, np.maximum(x["o3"], 1e-10), rtol=1e-15)
