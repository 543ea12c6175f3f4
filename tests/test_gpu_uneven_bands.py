"""k-distributions whose bands hold different numbers of g-points (the real
files' g128 / g112 sets are split unevenly over their 16 / 14 bands; the
synthetic tables use 8 per band): LW and SW all sky against the FP64 oracle.
Every per-band quantity (the g-point to band map, band Planck functions and
Planck-fraction normalisation, cloud optics and surface albedos by band,
the band sums of scmpsw) then meets band limits that are not multiples of 8.
Tolerances: fluxes 0.05 W/m2, heating rates 0.01 K/day."""
import copy

import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

LW_SIZES = [4, 12, 8, 6, 10, 16, 4, 8, 8, 12, 4, 8, 6, 10, 6, 6]      # 128
SW_SIZES = [6, 10, 8, 12, 4, 8, 10, 6, 8, 8, 12, 4, 10, 6]            # 112


def uneven(kd, sizes):
    """The table with its g-points regrouped into bands of the given sizes:
    each g-point takes the flavors of its new band's first g-point in the
    original table, and the Planck fractions are renormalised per band."""
    kd = copy.deepcopy(kd)
    assert sum(sizes) == kd.ngpt and len(sizes) == kd.nband
    lims = np.cumsum([0] + sizes)
    kd.band_lims_gpt = np.array([[lims[b] + 1, lims[b + 1]] for b in range(kd.nband)], dtype=np.int32)
    gf = kd.gpoint_flavor.copy()
    for b in range(kd.nband):
        gf[lims[b]:lims[b + 1]] = kd.gpoint_flavor[8 * b]
    kd.gpoint_flavor = gf
    if kd.is_lw:
        pf = kd.planck_frac.copy()
        for b in range(kd.nband):
            blk = pf[..., lims[b]:lims[b + 1]]
            pf[..., lims[b]:lims[b + 1]] = blk / blk.sum(axis=-1, keepdims=True)
        kd.planck_frac = np.ascontiguousarray(pf)
    return kd


def _check(ref, out, names, hr, cols=None):
    sel = (lambda a: a) if cols is None else (lambda a: a[:, cols])
    for n in names:
        mx, _ = report(n, sel(out[n]), sel(ref[n]), "W/m2")
        assert mx < 0.05, n
    mx, _ = report(hr, out[hr] * 86400, ref[hr] * 86400, "K/day")
    assert mx < 0.01


def test_lw_uneven_bands(kd_lw, lut_lw):
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    kd = uneven(kd_lw, LW_SIZES)
    ncol, nlay = 96, 127
    pr = syn.make_profiles(ncol, nlay, seed=71, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    ints = capi.lw_default_ints(ncol, nlay, blksz=32, nangle=3, clrsky=True, jacobian=True)
    ref = dict(arr)
    ref.update(O.lw_outputs(ncol, nlay))
    O.lw_main_run(kd, lut_lw, ints, ref)
    s = lib.LWScheme(kd, lut_lw, nlay, blksz=32)
    dev = device.to_device(arr)
    dev.update(device.lw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
                      "fluxlwUP_jac"], "htrlw")


def test_sw_uneven_bands(kd_sw, lut_sw):
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    kd = uneven(kd_sw, SW_SIZES)
    ncol, nlay = 96, 127
    pr = syn.make_profiles(ncol, nlay, seed=72, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
    ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=32, clrsky=True)
    ref = dict(arr)
    ref.update(O.sw_outputs(ncol, nlay))
    O.sw_main_run(kd, lut_sw, ints, ref)
    s = lib.SWScheme(kd, lut_sw, nlay, blksz=32)
    dev = device.to_device(arr)
    dev.update(device.sw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    _check(ref, out, ["fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky"], "htrsw")
    mx, _ = report("scmpsw", out["scmpsw"], ref["scmpsw"], "W/m2")
    assert mx < 0.05
