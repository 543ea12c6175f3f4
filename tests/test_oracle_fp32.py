"""The FP32 build of the oracle (liboracle32.so), timed by bench.py as the
FP32 CPU baseline.

It is the same source with the per-g-point optics and transport in FP32
(indices, interpolation weights, cloud optics by band, McICA and the band /
broadband sums stay FP64).  It is never a checker; these tests pin that it
runs the same algorithm and really computes in FP32:
- LW all-sky fluxes within 0.01 W/m2 and heating rates within 0.01 K/day of
  the FP64 build, and not bit-identical to it;
- SW: the reference's two-stream expressions cancel in FP32, mostly in the
  direct-beam reflectance and transmittance brackets (evaluating those alone
  in FP64 takes a 128-column sample's SW error from 4.8 to 1.5 W/m2), which
  the GPU kernel regroups; so its SW error is large, and the test records its
  size (below 20 W/m2 on this sample) and that every flux is finite."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def runs():
    from ccpprad import synthetic as syn, capi
    from oracle import oracle as O
    try:
        O.lib32()
    except RuntimeError as e:
        pytest.skip(str(e))
    ns, nl = 64, 127
    kd_lw, kd_sw = syn.make_kdist(True), syn.make_kdist(False)
    lut_lw, lut_sw = syn.make_cloud_lut(True), syn.make_cloud_lut(False)
    pr = syn.make_profiles(ns, nl, seed=20261203, clouds=True, kd=kd_lw)
    arr = capi.profile_fields(pr)
    arr["idx"] = np.arange(1, ns + 1, dtype=np.int32)
    out = {}
    for fp32 in (False, True):
        lw = dict(arr)
        lw.update(O.lw_outputs(ns, nl))
        sw = dict(arr)
        sw.update(O.sw_outputs(ns, nl))
        O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(ns, nl, blksz=16, nangle=3), lw, fp32=fp32)
        O.sw_main_run(kd_sw, lut_sw, capi.sw_default_ints(ns, ns, nl, blksz=16), sw, fp32=fp32)
        out[fp32] = (lw, sw)
    return out


def test_fp32_lw_matches_fp64(runs):
    (l64, _), (l32, _) = runs[False], runs[True]
    diffs = []
    for k in ("fluxlwUP_allsky", "fluxlwDOWN_allsky"):
        d = np.abs(l32[k] - l64[k]).max()
        diffs.append(d)
        assert d < 0.01, (k, d)
    assert np.abs(l32["htrlw"] - l64["htrlw"]).max() * 86400 < 0.01
    assert max(diffs) > 0.0        # the per-g-point work really ran in FP32


def test_fp32_sw_runs_with_its_cancellation(runs):
    (_, s64), (_, s32) = runs[False], runs[True]
    for k in ("fluxswUP_allsky", "fluxswDOWN_allsky", "htrsw"):
        assert np.isfinite(s32[k]).all(), k
    d = max(np.abs(s32[k] - s64[k]).max() for k in ("fluxswUP_allsky", "fluxswDOWN_allsky"))
    print("FP32 SW flux max-abs vs FP64:", d)
    assert 0.0 < d < 20.0
