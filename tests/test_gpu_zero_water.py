"""Clouds that hold no condensate leave the fluxes of the clear sky, bit for bit.

With zero liquid, ice, rain and snow paths the cloud optics are zero depths
(tau = 0, co-albedo 1, g = 0), and the reference's increments
(rrtmgp_sw_main.F90:603-604, rrtmgp_lw_main.F90:579-580) add nothing.  On
the GPU the all-sky solve then runs its cloudy-segment code (McICA masks,
cloud records, the general two-stream form with asymmetry 0) while the
clear-sky solve runs the clear-segment code (no cloud loads, the
asymmetry-free two-stream form), so equal bits check that the two forms
agree exactly and that a zero cloud depth is added exactly.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NCOL, NLAY, BLKSZ = 256, 127, 64


def _zero_water_fields(kd, seed):
    from ccpprad import synthetic as syn, capi
    pr = syn.make_profiles(NCOL, NLAY, seed=seed, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    for k in ("cld_lwp", "cld_iwp", "cld_rwp", "cld_swp"):
        arr[k] = np.zeros_like(arr[k])
    # every scheme block holds cloudy columns, so the all-sky solve takes its
    # cloudy-block path
    cf = arr["cld_frac"].reshape(NLAY, NCOL // BLKSZ, BLKSZ)
    assert (cf > 0).any(axis=(0, 2)).all()
    return arr


def test_sw_zero_water_clouds_give_clear_sky_bits(kd_sw, lut_sw):
    from ccpprad import capi, lib, device
    arr = _zero_water_fields(kd_sw, 71)
    arr["idx"] = np.arange(1, NCOL + 1, dtype=np.int32)
    ints = capi.sw_default_ints(NCOL, NCOL, NLAY, blksz=BLKSZ, clrsky=True)
    s = lib.SWScheme(kd_sw, lut_sw, NLAY, blksz=BLKSZ)
    dev = device.to_device(arr)
    dev.update(device.sw_outputs(NCOL, NLAY))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    assert np.isfinite(out["fluxswUP_allsky"]).all()
    for a, c in (("fluxswUP_allsky", "fluxswUP_clrsky"), ("fluxswDOWN_allsky", "fluxswDOWN_clrsky"),
                 ("htrsw", "htrswc")):
        assert np.array_equal(out[a], out[c]), a
    assert np.abs(out["cldtausw"]).max() == 0.0


def test_lw_zero_water_clouds_give_clear_sky_bits(kd_lw, lut_lw):
    from ccpprad import capi, lib, device
    arr = _zero_water_fields(kd_lw, 72)
    ints = capi.lw_default_ints(NCOL, NLAY, blksz=BLKSZ, nangle=3, clrsky=True)
    s = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=BLKSZ)
    dev = device.to_device(arr)
    dev.update(device.lw_outputs(NCOL, NLAY))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    assert np.isfinite(out["fluxlwUP_allsky"]).all()
    for a, c in (("fluxlwUP_allsky", "fluxlwUP_clrsky"), ("fluxlwDOWN_allsky", "fluxlwDOWN_clrsky"),
                 ("htrlw", "htrlwc")):
        assert np.array_equal(out[a], out[c]), a
