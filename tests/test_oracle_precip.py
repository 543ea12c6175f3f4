"""Pins of the oracle's precipitation optics by hand cases of the
reference's own formulas: the LW rain and snow absorption depths
(rrtmgp_lw_main.F90:404-428, constants rrtmgp_lw_cloud_optics.F90:64-65)
and the SW delta-scaled rain and snow optics (rrtmgp_sw_main.F90:436-468,
band tables rrtmgp_sw_cloud_optics.F90:369-381, a0r/a0s/a1s :58-61), with
the reference's default-REAL literals taken at single precision.

Each case runs the oracle's whole scheme on a column whose only cloud is
one fully covered layer without condensate (so the McICA mask covers every
g-point there and the cloud's own optics vanish) and compares its all-sky
fluxes with the oracle's solver fed the hand-computed precipitation optics:
a wrong coefficient, threshold, band index, unit or delta-scaling step
changes the fluxes."""
import copy

import numpy as np
import pytest

f32 = lambda x: float(np.float32(x))

# rrtmgp_sw_cloud_optics.F90:369-381
B0R = [0.496, 0.466, 0.437, 0.416, 0.391, 0.374, 0.352, 0.183, 0.048, 0.012, 0.0, 0.0, 0.0, 0.0]
B0S = [0.460] * 8 + [0.0] * 6
B1S = [0.0] * 8 + [1.62e-5, 1.62e-5] + [0.0] * 4
C0R = [0.980, 0.975, 0.965, 0.960, 0.955, 0.952, 0.950, 0.944, 0.894, 0.884, 0.883, 0.883, 0.883, 0.883]
C0S = [0.970] * 10 + [0.700] * 4

# (rwp g/m2, swp g/m2, snow radius um): rain only; rain and snow; snow below
# the 10 um threshold (no snow depth)
CASES = [(80.0, 0.0, 20.0), (80.0, 40.0, 30.0), (0.0, 40.0, 8.0)]
NLAY, CLOUDY = 12, 6


def _columns(kd):
    from ccpprad import synthetic as syn, capi
    nc = len(CASES)
    pr = syn.make_profiles(nc, NLAY, seed=61, clouds=False, kd=kd)
    arr = capi.profile_fields(pr)
    arr["cld_frac"][CLOUDY, :] = 1.0
    for c, (rwp, swp, res) in enumerate(CASES):
        arr["cld_rwp"][CLOUDY, c] = rwp
        arr["cld_swp"][CLOUDY, c] = swp
        arr["cld_resnow"][CLOUDY, c] = res
    return pr, arr


def lw_precip_tau(rwp, swp, res):
    """rrtmgp_lw_main.F90:415-421: absrain rwp + abssnow0 1.05756 swp / re
    when swp > 0 and re > 10 um."""
    t = f32(0.33e-3) * rwp
    if swp > 0.0 and res > 10.0:
        t += 1.5 * f32(1.05756) * swp / res
    return t


def sw_precip_optics(rwp, swp, res, b):
    """rrtmgp_sw_main.F90:440-466 for band b (0-based): (tau, ssa, g)."""
    tr = rwp * f32(3.07e-3)
    ts = swp * f32(1.09087) * (0.0 + 1.5 / (f32(1.0315) * res)) if (swp > 0.0 and res > 10.0) else 0.0
    sr = tr * (1.0 - f32(B0R[b]))
    ar = sr * f32(C0R[b])
    sn = ts * (1.0 - (f32(B0S[b]) + f32(B1S[b]) * f32(1.0315) * res))
    an = sn * f32(C0S[b])
    tp = max(1e-12, tr + ts)
    sp = max(1e-12, sr + sn)
    ap = max(1e-12, ar + an)
    asyw = ap / max(1e-12, sp)
    ssaw = min(1.0 - f32(0.000001), sp / tp)
    za2 = ssaw * asyw * asyw
    return (1.0 - za2) * tp, (ssaw - za2) / (1.0 - za2), asyw / (1.0 + asyw)


def test_lw_precipitation_depths(kd_lw, lut_lw):
    from ccpprad import capi
    from oracle import oracle as O
    pr, arr = _columns(kd_lw)
    nc, ng, nb = len(CASES), kd_lw.ngpt, kd_lw.nband
    ints = capi.lw_default_ints(nc, NLAY, blksz=nc, nangle=3)
    out = dict(arr)
    out.update(O.lw_outputs(nc, NLAY))
    O.lw_main_run(kd_lw, lut_lw, ints, out)
    tau, pf = O.gas_optics(kd_lw, pr)
    src = O.lw_sources(kd_lw, pr, pf)
    for c, case in enumerate(CASES):
        tau[c, CLOUDY, :] += lw_precip_tau(*case)
    emis = np.repeat(np.asarray(pr.semis)[:, None], nb, axis=1)
    up, dn, _ = O.lw_solve(kd_lw, nc, NLAY, tau, src, emis, nang=3)
    assert np.allclose(out["fluxlwUP_allsky"], up, rtol=1e-12, atol=0)
    assert np.allclose(out["fluxlwDOWN_allsky"], dn, rtol=1e-12, atol=0)
    # the hand depths: 0.0264 (rain), 0.0264 + 3.17268 (snow), 0 (snow below 10 um)
    assert lw_precip_tau(*CASES[0]) == pytest.approx(80 * 0.33e-3, rel=1e-7)
    assert lw_precip_tau(*CASES[1]) == pytest.approx(80 * 0.33e-3 + 1.5 * 1.05756 * 40 / 30, rel=1e-7)
    assert lw_precip_tau(*CASES[2]) == 0.0
    # and the precipitation is visible in the fluxes
    base = dict(arr)
    base["cld_rwp"] = np.zeros_like(arr["cld_rwp"])
    base["cld_swp"] = np.zeros_like(arr["cld_swp"])
    base.update(O.lw_outputs(nc, NLAY))
    O.lw_main_run(kd_lw, lut_lw, ints, base)
    assert np.abs(base["fluxlwDOWN_allsky"][:, :2] - out["fluxlwDOWN_allsky"][:, :2]).max() > 0.5


def test_sw_precipitation_optics(kd_sw, lut_sw):
    """No gas absorption and no Rayleigh (zero mixing ratios, zero Rayleigh
    table) and equal surface albedos: the all-sky column is the one
    precipitating layer, whose per-band optics come from the hand formula."""
    from ccpprad import capi
    from oracle import oracle as O
    kd = copy.deepcopy(kd_sw)
    kd.krayl_lower = np.zeros_like(kd.krayl_lower)
    kd.krayl_upper = np.zeros_like(kd.krayl_upper)
    pr, arr = _columns(kd)
    nc, ng, nb = len(CASES), kd.ngpt, kd.nband
    for k in list(arr):
        if k.startswith("vmr_"):
            arr[k] = np.zeros_like(arr[k])
    for k in ("sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir", "sfc_alb_uvvis_dif"):
        arr[k] = np.full(nc, 0.2)
    arr["coszen"] = np.array([0.6, 0.35, 0.9])
    arr["idx"] = np.arange(1, nc + 1, dtype=np.int32)
    ints = capi.sw_default_ints(nc, nc, NLAY, blksz=nc)
    out = dict(arr)
    out.update(O.sw_outputs(nc, NLAY))
    O.sw_main_run(kd, lut_sw, ints, out, solcon=1361.0)
    t = np.zeros((nc, NLAY, ng))
    s = np.zeros_like(t)
    g = np.zeros_like(t)
    lims = np.asarray(kd.band_lims_gpt).reshape(nb, 2)
    for c, case in enumerate(CASES):
        for b in range(nb):
            tb, sb, gb = sw_precip_optics(*case, b)
            g0, g1 = lims[b, 0] - 1, lims[b, 1]
            t[c, CLOUDY, g0:g1], s[c, CLOUDY, g0:g1], g[c, CLOUDY, g0:g1] = tb, sb, gb
    sol = np.asarray(kd.solar_src, dtype=np.float64)
    toa = np.repeat((sol * 1361.0 / sol.sum())[None, :], nc, axis=0)
    bu, bd, _ = O.sw_solve(kd, nc, NLAY, arr["coszen"], t, s, g, toa, np.full((nc, nb), 0.2),
                           np.full((nc, nb), 0.2))
    up, dn = bu.sum(axis=2).T, bd.sum(axis=2).T
    assert np.allclose(out["fluxswUP_allsky"], up, rtol=1e-10, atol=1e-9)
    assert np.allclose(out["fluxswDOWN_allsky"], dn, rtol=1e-10, atol=1e-9)
    # rain only, band 1: delta-scaled depth below the unscaled 0.2456, 0 < ssa < 1, g < 1/2
    tb, sb, gb = sw_precip_optics(*CASES[0], 0)
    assert 0.0 < sb < 1.0 and 0.0 < gb < 0.5 and tb < 80 * 3.07e-3
    # the precipitation is visible in the fluxes (the snow below 10 um is not)
    clear = dict(arr)
    clear["cld_rwp"] = np.zeros_like(arr["cld_rwp"])
    clear["cld_swp"] = np.zeros_like(arr["cld_swp"])
    clear.update(O.sw_outputs(nc, NLAY))
    O.sw_main_run(kd, lut_sw, ints, clear, solcon=1361.0)
    d = np.abs(out["fluxswDOWN_allsky"] - clear["fluxswDOWN_allsky"]).max(axis=0)
    assert d[0] > 0.5 and d[1] > 0.5 and d[2] < 1e-6
