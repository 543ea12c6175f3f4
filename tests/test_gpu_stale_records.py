"""The cloud-optics kernel writes no record for a layer without cloud (the
McICA mask has no bit there, so no solver reads it), and the random stream
stops after the last cloudy layer.  A record a call does not write keeps the
previous call's contents: here a first call on the same context fills every
layer with thick overcast cloud, and the second call, with clouds only
between 850 and 300 hPa, must still match the FP64 oracle (LW without and
with cloud scattering, SW; flux 0.05 W/m2, heating rate 0.01 K/day)."""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FLUX_TOL = 0.05
HR_TOL = 0.01


def _overcast(arr):
    a = dict(arr)
    a["cld_frac"] = np.ones_like(arr["cld_frac"])
    a["cld_lwp"] = np.full_like(arr["cld_lwp"], 500.0)
    a["cld_reliq"] = np.full_like(arr["cld_reliq"], 12.0)
    a["cld_iwp"] = np.full_like(arr["cld_iwp"], 200.0)
    a["cld_reice"] = np.full_like(arr["cld_reice"], 40.0)
    return a


@pytest.mark.parametrize("scheme", ["lw", "lw_scat", "sw"])
def test_records_left_by_an_earlier_call_are_not_read(kd_lw, lut_lw, kd_sw, lut_sw, scheme):
    from ccpprad import capi, device, lib, synthetic as syn
    from oracle import oracle as O
    ncol, nlay = 128, 127
    lw = scheme != "sw"
    kd, lut = (kd_lw, lut_lw) if lw else (kd_sw, lut_sw)
    pr = syn.make_profiles(ncol, nlay, seed=23, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    if lw:
        ints = capi.lw_default_ints(ncol, nlay, blksz=ncol, nangle=3, scat=scheme == "lw_scat")
        s, outs, ref_out, run_ref = lib.LWScheme(kd, lut, nlay, blksz=ncol), device.lw_outputs, O.lw_outputs, \
            O.lw_main_run
        names, hr = ("fluxlwUP_allsky", "fluxlwDOWN_allsky"), "htrlw"
    else:
        arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
        ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=ncol)
        s, outs, ref_out, run_ref = lib.SWScheme(kd, lut, nlay, blksz=ncol), device.sw_outputs, O.sw_outputs, \
            O.sw_main_run
        names, hr = ("fluxswUP_allsky", "fluxswDOWN_allsky"), "htrsw"
    print(f"stale records: {scheme}: overcast call, then the test profiles", flush=True)
    for a in (_overcast(arr), arr):
        dev = device.to_device(a)
        dev.update(outs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
    out = device.to_host(dev)
    ref = dict(arr)
    ref.update(ref_out(ncol, nlay))
    run_ref(kd, lut, ints, ref)
    for n in names:
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < FLUX_TOL, n
    mx, _ = report(hr, out[hr] * 86400, ref[hr] * 86400, "K/day")
    assert mx < HR_TOL
