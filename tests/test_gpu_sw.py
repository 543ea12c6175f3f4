"""GPU parity of the SW path against the FP64 oracle, through the C-ABI."""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FLUX_TOL = 0.05
HR_TOL = 0.01


def _run_both(kd, lut, ncol, nlay, clouds, clrsky, blksz=None, nday=None, iovr="iovr_maxrand", seed=20261203,
              aer=False, top_at_1=False):
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    pr = syn.make_profiles(ncol, nlay, seed=seed, clouds=clouds, kd=kd)
    arr = capi.profile_fields(pr)
    if aer:
        # aerosol optics by band, Fortran (ncol, nlay, nband)
        rng = np.random.default_rng(seed + 9)
        shp = (kd.nband, nlay, ncol)
        arr["aersw_tau"] = 10 ** rng.uniform(-5, -1, shp)
        arr["aersw_ssa"] = rng.uniform(0.8, 1.0, shp)
        arr["aersw_g"] = rng.uniform(0.5, 0.8, shp)
    if top_at_1:
        from test_gpu_lw import _flip
        arr = _flip(arr, nlay)
    nday = ncol if nday is None else nday
    rng = np.random.default_rng(seed + 5)
    idx = np.sort(rng.choice(ncol, nday, replace=False)) + 1 if nday < ncol else np.arange(1, ncol + 1)
    arr["idx"] = idx.astype(np.int32)
    ints = capi.sw_default_ints(ncol, nday, nlay, blksz=blksz, clrsky=clrsky, iovr=iovr, top_at_1=top_at_1)
    ref = dict(arr)
    ref.update(O.sw_outputs(ncol, nlay))
    O.sw_main_run(kd, lut, ints, ref)
    s = lib.SWScheme(kd, lut, nlay, blksz=ints["rrtmgp_phys_blksz"])
    dev = device.to_device(arr)
    dev.update(device.sw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    return ref, device.to_host(dev), idx - 1


def test_sw_gas_optics_parity(kd_sw):
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    ncol, nlay = 96, 127
    pr = syn.make_profiles(ncol, nlay, clouds=False, kd=kd_sw)
    tau_o, ray_o = O.gas_optics(kd_sw, pr)
    s = lib.SWScheme(kd_sw, None, nlay)
    arr = capi.profile_fields(pr)
    d = device.to_device({k: arr[k] for k in ("p_lay", "p_lev", "t_lay")})
    v = torch.from_numpy(O.vmr_stack(pr)).cuda()
    ta = torch.zeros((ncol, nlay, kd_sw.ngpt), dtype=torch.float32, device="cuda")
    tr = torch.zeros_like(ta)
    s.gas_optics_dev(ncol, nlay, d["p_lay"], d["p_lev"], d["t_lay"], v, ta, tr)
    torch.cuda.synchronize()
    rel = np.abs(ta.cpu().numpy() - tau_o) / np.maximum(np.abs(tau_o), 1e-30)
    relr = np.abs(tr.cpu().numpy() - ray_o) / np.maximum(np.abs(ray_o), 1e-30)
    print("tau rel", rel.max(), "ray rel", relr.max())
    assert rel.max() < 2e-5 and relr.max() < 2e-5


def _check(ref, out, cols, clr=True):
    names = ["fluxswUP_allsky", "fluxswDOWN_allsky"] + (["fluxswUP_clrsky", "fluxswDOWN_clrsky"] if clr else [])
    for n in names:
        mx, _ = report(n, out[n][:, cols], ref[n][:, cols], "W/m2")
        assert mx < FLUX_TOL, n
    mx, _ = report("htrsw", out["htrsw"] * 86400, ref["htrsw"] * 86400, "K/day")
    assert mx < HR_TOL
    mx, _ = report("scmpsw", out["scmpsw"], ref["scmpsw"], "W/m2")
    assert mx < FLUX_TOL
    mx, _ = report("cldtausw", out["cldtausw"], ref["cldtausw"], "")
    assert mx < 1e-3 * max(1.0, np.abs(ref["cldtausw"]).max())


def test_sw_all_sky(kd_sw, lut_sw):
    ref, out, cols = _run_both(kd_sw, lut_sw, 256, 127, clouds=True, clrsky=True, blksz=64)
    _check(ref, out, cols)


def test_sw_all_sky_no_clrsky_partial_day(kd_sw, lut_sw):
    """nDay < nCol (night columns untouched), doSWclrsky false: exercises the
    reference's per-block reset of scmpsw and cldtausw."""
    ref, out, cols = _run_both(kd_sw, lut_sw, 192, 127, clouds=True, clrsky=False, blksz=32, nday=128)
    _check(ref, out, cols, clr=False)


def test_sw_clear_sky(kd_sw, lut_sw):
    ref, out, cols = _run_both(kd_sw, lut_sw, 128, 64, clouds=False, clrsky=True)
    _check(ref, out, cols)


def test_sw_aerosols(kd_sw, lut_sw):
    """Aerosol tau, ssa, g by band incremented into the clear and all sky
    (increment_2stream_by_2stream)."""
    ref, out, cols = _run_both(kd_sw, lut_sw, 128, 127, clouds=True, clrsky=True, blksz=64, aer=True)
    _check(ref, out, cols)


def test_sw_top_at_1(kd_sw, lut_sw):
    """The same kind of atmosphere stored top first (top_at_1, iSFC = nLay + 1),
    all and clear sky with a partial day, against the oracle; and the clear-sky
    fluxes (no McICA draws) equal the surface-first run's with the levels
    reversed, to the flux bar."""
    ref, out, cols = _run_both(kd_sw, lut_sw, 128, 127, clouds=True, clrsky=True, blksz=32, nday=96,
                               top_at_1=True)
    _check(ref, out, cols)
    ref0, out0, cols0 = _run_both(kd_sw, lut_sw, 128, 127, clouds=True, clrsky=True, blksz=32, nday=96)
    assert np.array_equal(cols, cols0)
    for n in ("fluxswUP_clrsky", "fluxswDOWN_clrsky"):
        mx, _ = report(n + " flipped", out[n][::-1][:, cols], out0[n][:, cols], "W/m2")
        assert mx < FLUX_TOL, n


def test_sw_no_daylit_columns(kd_sw, lut_sw):
    """nDay = 0 (rrtmgp_sw_main.F90:676-682): the level fluxes and scmpsw are
    zeroed, cldtausw keeps its values; the same through the host entry (the
    Fortran path) and as the oracle."""
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    ncol, nlay = 64, 127
    pr = syn.make_profiles(ncol, nlay, clouds=True, kd=kd_sw)
    arr = capi.profile_fields(pr)
    arr["idx"] = np.zeros(1, dtype=np.int32)
    ints = capi.sw_default_ints(ncol, 0, nlay, blksz=16)
    s = lib.SWScheme(kd_sw, lut_sw, nlay, blksz=16)
    dev = device.to_device(arr)
    outs = device.sw_outputs(ncol, nlay)
    for v in outs.values():
        v.fill_(5.0)
    dev.update(outs)
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    got = device.to_host(dev)
    host = dict(arr)
    host.update({k: np.full(v.shape, 5.0) for k, v in O.sw_outputs(ncol, nlay).items()})
    s.run(ints, host)
    ref = dict(arr)
    ref.update({k: np.full(v.shape, 5.0) for k, v in O.sw_outputs(ncol, nlay).items()})
    O.sw_main_run(kd_sw, lut_sw, ints, ref)
    for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky", "scmpsw"):
        assert not got[n].any() and not host[n].any() and not ref[n].any(), n
    for d in (got, host, ref):
        assert (d["cldtausw"] == 5.0).all()


@pytest.mark.parametrize("aer", [False, True])
def test_sw_clear_blocks_with_clrsky(kd_sw, lut_sw, aer):
    """Blocks without clouds next to cloudy ones, with clear-sky fluxes on: the
    clear-sky pass skips the cloud-free blocks, whose clear-sky values the
    all-sky pass wrote; all against the oracle, and all sky equal to clear
    sky bit for bit in those blocks."""
    from ccpprad import synthetic as syn, capi, lib, device
    from oracle import oracle as O
    ncol, nlay, blk = 256, 127, 64
    arr = capi.profile_fields(syn.make_profiles(ncol, nlay, seed=41, clouds=True, kd=kd_sw))
    clear = np.r_[blk:2 * blk, 3 * blk:4 * blk]
    arr["cld_frac"][:, clear] = 0.0
    if aer:
        rng = np.random.default_rng(42)
        shp = (kd_sw.nband, nlay, ncol)
        arr.update(aersw_tau=10 ** rng.uniform(-5, -1, shp), aersw_ssa=rng.uniform(0.8, 1.0, shp),
                   aersw_g=rng.uniform(0.5, 0.8, shp))
    arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
    ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=blk, clrsky=True)
    ref = dict(arr)
    ref.update(O.sw_outputs(ncol, nlay))
    O.sw_main_run(kd_sw, lut_sw, ints, ref)
    s = lib.SWScheme(kd_sw, lut_sw, nlay, blksz=blk)
    dev = device.to_device(arr)
    dev.update(device.sw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    _check(ref, out, np.arange(ncol))
    mx, _ = report("htrswc", out["htrswc"] * 86400, ref["htrswc"] * 86400, "K/day")
    assert mx < HR_TOL
    for a, c in (("fluxswUP_allsky", "fluxswUP_clrsky"), ("fluxswDOWN_allsky", "fluxswDOWN_clrsky"),
                 ("htrsw", "htrswc")):
        assert np.array_equal(out[a][:, clear], out[c][:, clear]), a
    cloudy = np.setdiff1d(np.arange(ncol), clear)
    assert not np.array_equal(out["fluxswUP_allsky"][:, cloudy], out["fluxswUP_clrsky"][:, cloudy])
