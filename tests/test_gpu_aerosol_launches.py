"""Aerosol optics across launch boundaries.

Before each solver launch the library packs the launch's aerosol optics
(CCPP order (ncol, nlay, nband), FP64) into per-column FP32 records; the
SW pack follows the daylit index.  With 640 columns (10 blocks of 64), a
partial day for SW and launches of 128 columns, every launch packs a
different column range; the results must equal, bit for bit, those of one
launch over all columns, and the FP64 oracle within the plain-path bar
(0.05 W/m2, 0.01 K/day max-abs).  LW runs both solvers (no scattering with
3 angles, and the two-stream solver with cloud scattering).
"""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NCOL, NLAY, BLK = 640, 127, 64
FLUX_TOL, HR_TOL = 0.05, 0.01


def _aer(pre, nband, seed):
    rng = np.random.default_rng(seed)
    shp = (nband, NLAY, NCOL)
    return {pre + "tau": 10 ** rng.uniform(-5, -0.5, shp), pre + "ssa": rng.uniform(0.8, 1.0, shp),
            pre + "g": rng.uniform(0.5, 0.8, shp)}


def _run(s, ints, arr, outputs, launch):
    from ccpprad import device
    assert s.set_launch_columns(launch) == launch
    dev = device.to_device(arr)
    dev.update(outputs(NCOL, NLAY))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    return device.to_host(dev)


def test_sw_aerosols_across_launches(kd_sw, lut_sw):
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    arr = capi.profile_fields(syn.make_profiles(NCOL, NLAY, seed=31, clouds=True, kd=kd_sw))
    arr.update(_aer("aersw_", kd_sw.nband, 32))
    nday = 448   # 7 blocks of daylit columns
    idx = np.sort(np.random.default_rng(33).choice(NCOL, nday, replace=False)) + 1
    arr["idx"] = idx.astype(np.int32)
    ints = capi.sw_default_ints(NCOL, nday, NLAY, blksz=BLK, clrsky=True)
    s = lib.SWScheme(kd_sw, lut_sw, NLAY, blksz=BLK)
    small = _run(s, ints, arr, device.sw_outputs, 128)
    whole = _run(s, ints, arr, device.sw_outputs, nday)
    for k in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky", "htrsw", "htrswc",
              "scmpsw"):
        assert np.array_equal(small[k], whole[k]), k
    ref = dict(arr)
    ref.update(O.sw_outputs(NCOL, NLAY))
    O.sw_main_run(kd_sw, lut_sw, ints, ref)
    cols = idx - 1
    for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky"):
        mx, _ = report(n, small[n][:, cols], ref[n][:, cols], "W/m2")
        assert mx < FLUX_TOL, n
    for n in ("htrsw", "htrswc"):
        mx, _ = report(n, small[n] * 86400, ref[n] * 86400, "K/day")
        assert mx < HR_TOL, n


@pytest.mark.parametrize("scat", [False, True])
def test_lw_aerosols_across_launches(kd_lw, lut_lw, scat):
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    arr = capi.profile_fields(syn.make_profiles(NCOL, NLAY, seed=34, clouds=True, kd=kd_lw))
    arr.update(_aer("aerlw_", kd_lw.nband, 35))
    nang = 1 if scat else 3
    ints = capi.lw_default_ints(NCOL, NLAY, blksz=BLK, nangle=nang, clrsky=not scat, scat=scat)
    s = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=BLK)
    small = _run(s, ints, arr, device.lw_outputs, 128)
    whole = _run(s, ints, arr, device.lw_outputs, NCOL)
    names = ("fluxlwUP_allsky", "fluxlwDOWN_allsky") + (() if scat else ("fluxlwUP_clrsky", "fluxlwDOWN_clrsky"))
    for k in names + ("htrlw",):
        assert np.array_equal(small[k], whole[k]), k
    ref = dict(arr)
    ref.update(O.lw_outputs(NCOL, NLAY))
    O.lw_main_run(kd_lw, lut_lw, ints, ref)
    for n in names:
        mx, _ = report(n, small[n], ref[n], "W/m2")
        assert mx < FLUX_TOL, n
    mx, _ = report("htrlw", small["htrlw"] * 86400, ref["htrlw"] * 86400, "K/day")
    assert mx < HR_TOL
