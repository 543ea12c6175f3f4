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


def post_inputs(nc=37, nl=20, nday=1, seed=4):
    """Random level fluxes, components and diagnostics for GFS_rrtmgp_post;
    numpy arrays shaped as the Fortran arrays reversed (column fastest)."""
    rng = np.random.default_rng(seed)
    f = lambda *s: rng.uniform(0.0, 400.0, s[::-1]) if s else None
    arr = {k: f(nc, nl + 1) for k in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky",
                                      "fluxlwDOWN_clrsky", "fluxswUP_allsky", "fluxswDOWN_allsky",
                                      "fluxswUP_clrsky", "fluxswDOWN_clrsky")}
    arr["scmpsw"] = rng.uniform(0.0, 300.0, (nc, 6))
    for k in ("sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir", "sfc_alb_uvvis_dif"):
        arr[k] = rng.uniform(0.05, 0.8, nc)
    arr["tsfa"] = rng.uniform(250.0, 300.0, nc)
    arr["coszen"] = rng.uniform(-0.5, 1.0, nc)
    arr["coszdg"] = rng.uniform(0.0, 1.0, nc)
    arr["htrlw"] = rng.normal(0.0, 1e-5, (nl, nc))
    arr["p_lev"] = np.linspace(1e5, 100.0, nl + 1)[:, None] * np.ones((1, nc))
    arr["cldsa"] = rng.uniform(0.0, 1.0, (5, nc))
    mb = rng.integers(1, nl // 2, (3, nc))
    arr["mbota"] = mb.astype(np.int32)
    arr["mtopa"] = (mb + rng.integers(0, nl // 2, (3, nc))).astype(np.int32)
    arr["cldtaulw"] = rng.uniform(0.0, 2.0, (nl, nc))
    arr["cldtausw"] = rng.uniform(0.0, 5.0, (nl, nc))
    arr["aerodp"] = rng.uniform(0.0, 0.5, (6, nc))
    z = lambda *s: np.zeros(s[::-1])
    out = dict(topflw=np.zeros((nc, 2)), sfcflw=np.zeros((nc, 4)), topfsw=np.zeros((nc, 3)),
               sfcfsw=np.zeros((nc, 4)), htrlwu=z(nc, nl), fluxr=rng.uniform(0.0, 10.0, (45, nc)))
    for k in ("sfcdlw", "sfculw", "tsflw", "nirbmdi", "nirdfdi", "visbmdi", "visdfdi", "nirbmui", "nirdfui",
              "visbmui", "visdfui", "sfcnsw", "sfcdsw"):
        out[k] = np.zeros(nc)
    ints = dict(nCol=nc, nLev=nl, nDay=nday, iSFC=1, iTOA=nl + 1, doLWrad=1, doSWrad=1, save_diag=1,
                naerodp=6, nfxr=45)
    reals = dict(fhlwr=3600.0, fhswr=3600.0, raddt=1800.0)
    return ints, reals, arr, out


def test_post_bookkeeping_hand_cases():
    """TOA / surface picks, surface components and the time-weighted fluxr
    sums against their definitions; a sunless call zeroes the SW fields."""
    from oracle import oracle as O
    ints, reals, arr, out = post_inputs()
    f0 = out["fluxr"].copy()
    a = dict(arr)
    a.update(out)
    O.post(ints, reals, a)
    nl, it, isf = ints["nLev"], ints["iTOA"] - 1, ints["iSFC"] - 1
    assert np.array_equal(out["topflw"][:, 0], arr["fluxlwUP_allsky"][it])
    assert np.array_equal(out["sfcflw"][:, 2], arr["fluxlwDOWN_allsky"][isf])
    assert np.array_equal(out["topfsw"][:, 1], arr["fluxswDOWN_allsky"][it])
    assert np.array_equal(out["sfcfsw"][:, 3], arr["fluxswDOWN_clrsky"][isf])
    assert np.array_equal(out["sfcnsw"], out["sfcfsw"][:, 1] - out["sfcfsw"][:, 0])
    assert np.array_equal(out["nirbmui"], arr["scmpsw"][:, 2] * arr["sfc_alb_nir_dir"])
    assert np.array_equal(out["htrlwu"], arr["htrlw"])
    fr = out["fluxr"]
    # the sums may be contracted into FMAs (one rounding instead of two)
    assert np.allclose(fr[0], f0[0] + 3600.0 * arr["fluxlwUP_allsky"][it], rtol=1e-15, atol=0)
    day = arr["coszen"] > 0
    w = 3600.0 * arr["coszdg"] / arr["coszen"]
    assert np.allclose(fr[1][day], (f0[1] + out["topfsw"][:, 0] * w)[day], rtol=1e-15, atol=0)
    assert np.array_equal(fr[1][~day], f0[1][~day])
    assert np.array_equal(fr[33:39], arr["aerodp"])
    # cloud emissivity of the high cloud (j = 1 -> fluxr(45)): sum of tau over base..top
    c = 5
    b_, t_ = arr["mbota"][0, c], arr["mtopa"][0, c]
    tau = arr["cldtaulw"][b_ - 1:t_, c].sum()
    assert fr[44, c] == pytest.approx(f0[44, c] + 1800.0 * arr["cldsa"][0, c] * (1 - np.exp(-tau)), rel=1e-14)
    ints["nDay"] = 0
    O.post(ints, reals, a)
    assert not out["sfcfsw"].any() and not out["topfsw"].any() and not out["nirbmui"].any()
