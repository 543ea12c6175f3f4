"""GPU parity of the paths the plain all-sky tests leave out, against the
FP64 oracle through the C-ABI:

- several launches per call (the library's columns per launch set small,
  and a 131,072-column call at the default 65,536), compared with the
  oracle and bit for bit with a one-launch run of the same call;
- precipitation optics (rain and snow, snow radius on both sides of the
  10 um threshold, cloud fractions on both sides of eps and ftiny),
  convective and PBL sub-grid clouds, isubc = 2 seeds, every SW overlap
  flag and LW exponential-random overlap.

Tolerances (the plain-path bar): fluxes 0.05 W/m2 and heating rate
0.01 K/day max-abs against FP64."""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FLUX_TOL = 0.05      # W/m2
HR_TOL = 0.01        # K/day
LW_FLUX = ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky")
SW_FLUX = ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky")


def _fields(kd, ncol, nlay, seed, precip=False, sgs=False):
    """Scheme inputs of config[1]-style profiles; precip adds rain and snow
    water paths in the cloudy layers (snow radius 5-60 um, so on both sides
    of the reference's 10 um threshold) and a few layers whose cloud
    fraction sits just above or below eps = 1e-6 and ftiny = 1e-12; sgs adds
    convective and PBL cloud water paths with radii inside the LUT range."""
    from ccpprad import synthetic as syn, capi
    pr = syn.make_profiles(ncol, nlay, seed=seed, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    rng = np.random.default_rng(seed + 1)
    shape = (nlay, ncol)
    cloudy = arr["cld_frac"] > 0.0
    if precip:
        # cloud fractions around the thresholds in a few clear layers above the cloud deck
        for v in (2e-6, 5e-7, 2e-12, 5e-13):
            l = int(rng.integers(nlay // 2, nlay - 2))
            cols = rng.choice(ncol, size=max(1, ncol // 8), replace=False)
            arr["cld_frac"][l, cols] = v
            arr["cld_lwp"][l, cols] = 5.0
            arr["cld_reliq"][l, cols] = 8.0
        cloudy = arr["cld_frac"] > 0.0
        arr["cld_rwp"] = np.where(cloudy, 10 ** rng.uniform(-1.0, 2.0, shape), 0.0)
        arr["cld_swp"] = np.where(cloudy & (rng.uniform(size=shape) < 0.7), 10 ** rng.uniform(-1.0, 2.0, shape), 0.0)
        arr["cld_resnow"] = rng.uniform(5.0, 60.0, shape)
    if sgs:
        for pre in ("cld_cnv_", "cld_pbl_"):
            on = cloudy & (rng.uniform(size=shape) < 0.5)
            arr[pre + "lwp"] = np.where(on, 20.0 * np.exp(rng.normal(0.0, 1.0, shape)), 0.0)
            arr[pre + "iwp"] = np.where(on, 10.0 * np.exp(rng.normal(0.0, 1.0, shape)), 0.0)
            arr[pre + "reliq"] = rng.uniform(4.0, 18.0, shape)
            arr[pre + "reice"] = rng.uniform(15.0, 120.0, shape)
    arr["cloud_overlap_param"] = rng.uniform(0.0, 1.0, shape)
    arr["icseed_lw"] = rng.integers(1, 2 ** 31 - 1, ncol).astype(np.int32)
    arr["icseed_sw"] = rng.integers(1, 2 ** 31 - 1, ncol).astype(np.int32)
    return {k: np.ascontiguousarray(v) for k, v in arr.items()}


def _lw(kd, lut, arr, ints, launch=None):
    from ccpprad import lib, device
    from oracle import oracle as O
    ncol, nlay = ints["nCol"], ints["nLay"]
    ref = dict(arr)
    ref.update(O.lw_outputs(ncol, nlay))
    O.lw_main_run(kd, lut, ints, ref)
    s = lib.LWScheme(kd, lut, nlay, blksz=ints["rrtmgp_phys_blksz"])
    outs = []
    for n in ([launch, ncol] if launch else [None]):
        if n:
            s.set_launch_columns(n)
        dev = device.to_device(arr)
        dev.update(device.lw_outputs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
        outs.append(device.to_host(dev))
    return ref, outs


def _sw(kd, lut, arr, ints, launch=None):
    from ccpprad import lib, device
    from oracle import oracle as O
    ncol, nlay = ints["nCol"], ints["nLay"]
    ref = dict(arr)
    ref.update(O.sw_outputs(ncol, nlay))
    O.sw_main_run(kd, lut, ints, ref)
    s = lib.SWScheme(kd, lut, nlay, blksz=ints["rrtmgp_phys_blksz"])
    outs = []
    for n in ([launch, ints["nDay"]] if launch else [None]):
        if n:
            s.set_launch_columns(n)
        dev = device.to_device(arr)
        dev.update(device.sw_outputs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
        outs.append(device.to_host(dev))
    return ref, outs


def _check(ref, out, fluxes, hrs):
    for n in fluxes:
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < FLUX_TOL, n
    for n in hrs:
        mx, _ = report(n, out[n] * 86400, ref[n] * 86400, "K/day")
        assert mx < HR_TOL, n


def _same(a, b, names):
    for n in names:
        assert np.array_equal(a[n], b[n]), f"{n}: the multi-launch run differs from the one-launch run"


@pytest.mark.parametrize("isubc", [1, 2])
def test_lw_multi_launch(kd_lw, lut_lw, isubc):
    """256 columns, blocks of 16, 32 columns per launch (8 launches), all sky
    with precipitation and sub-grid clouds, clear sky on: every column
    against the oracle, and bit for bit against one launch (column offsets,
    McICA seeds by global column, work buffers reused across launches)."""
    from ccpprad import capi
    ncol, nlay = 256, 127
    arr = _fields(kd_lw, ncol, nlay, seed=301, precip=True, sgs=True)
    ints = capi.lw_default_ints(ncol, nlay, blksz=16, nangle=3, clrsky=True, isubc=isubc)
    ints.update(doGP_sgs_cnv=1, doGP_sgs_pbl=1)
    ref, (multi, one) = _lw(kd_lw, lut_lw, arr, ints, launch=32)
    _check(ref, multi, LW_FLUX, ("htrlw", "htrlwc"))
    _same(multi, one, LW_FLUX + ("htrlw", "htrlwc"))


@pytest.mark.parametrize("clrsky", [True, False])
def test_sw_multi_launch(kd_sw, lut_sw, clrsky):
    """256 columns of which 192 daylit (scattered), blocks of 16, 32 daylit
    columns per launch (6 launches), precipitation and sub-grid clouds:
    fluxes, heating rates, scmpsw (per-block resets) and cldtausw (last
    block only) against the oracle, and bit for bit against one launch."""
    from ccpprad import capi
    ncol, nlay, nday = 256, 127, 192
    arr = _fields(kd_sw, ncol, nlay, seed=302, precip=True, sgs=True)
    rng = np.random.default_rng(5)
    arr["idx"] = np.sort(rng.choice(ncol, nday, replace=False)).astype(np.int32) + 1
    ints = capi.sw_default_ints(ncol, nday, nlay, blksz=16, clrsky=clrsky)
    ints.update(doGP_sgs_cnv=1, doGP_sgs_pbl=1)
    ref, (multi, one) = _sw(kd_sw, lut_sw, arr, ints, launch=32)
    _check(ref, multi, SW_FLUX, ("htrsw",) + (("htrswc",) if clrsky else ()))
    mx, _ = report("scmpsw", multi["scmpsw"], ref["scmpsw"], "W/m2")
    assert mx < FLUX_TOL
    mx, _ = report("cldtausw", multi["cldtausw"], ref["cldtausw"], "")
    assert mx < 1e-4 * max(1.0, float(np.abs(ref["cldtausw"]).max()))
    assert np.abs(ref["cldtausw"]).max() > 0.0
    _same(multi, one, SW_FLUX + ("htrsw", "scmpsw", "cldtausw"))


@pytest.mark.parametrize("isubc", [1, 2])
def test_lw_precip_and_subgrid_clouds(kd_lw, lut_lw, isubc):
    """LW rain and snow depths (rrtmgp_lw_main.F90:404-431) and convective
    and PBL cloud optics (:372-400), exponential-random overlap."""
    from ccpprad import capi
    ncol, nlay = 128, 127
    arr = _fields(kd_lw, ncol, nlay, seed=303, precip=True, sgs=True)
    ints = capi.lw_default_ints(ncol, nlay, blksz=32, nangle=3, clrsky=True, isubc=isubc, iovr="iovr_exprand")
    ints.update(doGP_sgs_cnv=1, doGP_sgs_pbl=1)
    ref, (out,) = _lw(kd_lw, lut_lw, arr, ints)
    _check(ref, out, LW_FLUX, ("htrlw", "htrlwc"))
    # the precipitation changes the all-sky result (it is not silently dropped)
    arr0 = dict(arr)
    arr0["cld_rwp"] = np.zeros_like(arr["cld_rwp"])
    arr0["cld_swp"] = np.zeros_like(arr["cld_swp"])
    ref0, (out0,) = _lw(kd_lw, lut_lw, arr0, ints)
    assert np.abs(out0["fluxlwDOWN_allsky"] - out["fluxlwDOWN_allsky"]).max() > 0.1


def test_lw_scattering_with_precip(kd_lw, lut_lw):
    """doGP_lwscat with precipitation and sub-grid clouds (the two-stream
    cloud record carries the precipitation depth)."""
    from ccpprad import capi
    ncol, nlay = 128, 127
    arr = _fields(kd_lw, ncol, nlay, seed=304, precip=True, sgs=True)
    ints = capi.lw_default_ints(ncol, nlay, blksz=32, nangle=1, clrsky=False, scat=True)
    ints.update(doGP_sgs_cnv=1, doGP_sgs_pbl=1)
    ref, (out,) = _lw(kd_lw, lut_lw, arr, ints)
    _check(ref, out, LW_FLUX[:2], ("htrlw",))


@pytest.mark.parametrize("isubc", [1, 2])
def test_sw_precip_and_subgrid_clouds(kd_sw, lut_sw, isubc):
    """SW rain and snow optics with their delta-scaling and single-precision
    literals (rrtmgp_sw_main.F90:436-471), convective and PBL clouds."""
    from ccpprad import capi
    ncol, nlay = 128, 127
    arr = _fields(kd_sw, ncol, nlay, seed=305, precip=True, sgs=True)
    arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
    ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=32, clrsky=True, isubc=isubc)
    ints.update(doGP_sgs_cnv=1, doGP_sgs_pbl=1)
    ref, (out,) = _sw(kd_sw, lut_sw, arr, ints)
    _check(ref, out, SW_FLUX, ("htrsw", "htrswc"))
    mx, _ = report("scmpsw", out["scmpsw"], ref["scmpsw"], "W/m2")
    assert mx < FLUX_TOL
    arr0 = dict(arr)
    arr0["cld_rwp"] = np.zeros_like(arr["cld_rwp"])
    arr0["cld_swp"] = np.zeros_like(arr["cld_swp"])
    ref0, (out0,) = _sw(kd_sw, lut_sw, arr0, ints)
    assert np.abs(out0["fluxswUP_allsky"] - out["fluxswUP_allsky"]).max() > 0.1


@pytest.mark.parametrize("iovr", ["iovr_rand", "iovr_maxrand", "iovr_max", "iovr_dcorr", "iovr_exp",
                                  "iovr_exprand"])
def test_sw_overlap_variants(kd_sw, lut_sw, iovr):
    """Every overlap flag of the SW McICA sampling (rrtmgp_sw_main.F90:488-523),
    with isubc 2 seeds for the decorrelation-length one."""
    from ccpprad import capi
    ncol, nlay, nday = 96, 127, 64
    arr = _fields(kd_sw, ncol, nlay, seed=306)
    arr["idx"] = np.sort(np.random.default_rng(7).choice(ncol, nday, replace=False)).astype(np.int32) + 1
    ints = capi.sw_default_ints(ncol, nday, nlay, blksz=32, iovr=iovr, isubc=2 if iovr == "iovr_dcorr" else 1)
    ref, (out,) = _sw(kd_sw, lut_sw, arr, ints)
    _check(ref, out, SW_FLUX[:2], ("htrsw",))


def test_lw_exprand_overlap(kd_lw, lut_lw):
    from ccpprad import capi
    ncol, nlay = 96, 127
    arr = _fields(kd_lw, ncol, nlay, seed=307)
    ints = capi.lw_default_ints(ncol, nlay, blksz=32, nangle=2, iovr="iovr_exprand")
    ref, (out,) = _lw(kd_lw, lut_lw, arr, ints)
    _check(ref, out, LW_FLUX[:2], ("htrlw",))
