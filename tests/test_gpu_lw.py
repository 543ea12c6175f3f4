"""GPU parity of the LW path against the FP64 oracle, through the C-ABI.

Tolerances (the plain-path bar): fluxes 0.05 W/m2 and heating rate
0.01 K/day max-abs against FP64; gas-optics tau relative 2e-5."""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FLUX_TOL = 0.05      # W/m2
HR_TOL = 0.01        # K/day


def _setup(kd, lut, ncol, nlay, clouds, seed=20261203):
    from ccpprad import synthetic as syn, capi
    pr = syn.make_profiles(ncol, nlay, seed=seed, clouds=clouds, kd=kd)
    return pr, capi.profile_fields(pr)


def test_lw_gas_optics_parity(kd_lw, lut_lw):
    from ccpprad import lib, device
    from oracle import oracle as O
    ncol, nlay = 96, 127
    pr, arr = _setup(kd_lw, lut_lw, ncol, nlay, False)
    tau_o, pf_o = O.gas_optics(kd_lw, pr)
    s = lib.LWScheme(kd_lw, lut_lw, nlay)
    v = torch.from_numpy(O.vmr_stack(pr)).cuda()
    d = device.to_device({k: arr[k] for k in ("p_lay", "p_lev", "t_lay")})
    tau = torch.zeros((ncol, nlay, kd_lw.ngpt), dtype=torch.float32, device="cuda")
    pf = torch.zeros_like(tau)
    s.gas_optics_dev(ncol, nlay, d["p_lay"], d["p_lev"], d["t_lay"], v, tau, pf)
    torch.cuda.synchronize()
    tau_g, pf_g = tau.cpu().numpy().astype(np.float64), pf.cpu().numpy().astype(np.float64)
    rel = np.abs(tau_g - tau_o) / np.maximum(np.abs(tau_o), 1e-30)
    print("tau rel max", rel.max(), "pfrac abs max", np.abs(pf_g - pf_o).max())
    assert rel.max() < 2e-5
    assert np.abs(pf_g - pf_o).max() < 2e-6


def _aerosols(ncol, nlay, nband, seed, sw=False):
    """Aerosol optics by band, Fortran (ncol, nlay, nband): tau up to ~0.1 per
    layer, ssa 0.8-1, g 0.5-0.8 (LW uses tau only)."""
    rng = np.random.default_rng(seed)
    pre = "aersw_" if sw else "aerlw_"
    f = lambda a: np.ascontiguousarray(a)
    return {pre + "tau": f(10 ** rng.uniform(-5, -1, (nband, nlay, ncol))),
            pre + "ssa": f(rng.uniform(0.8, 1.0, (nband, nlay, ncol))),
            pre + "g": f(rng.uniform(0.5, 0.8, (nband, nlay, ncol)))}


def _flip(arr, nlay):
    """Reverse the vertical axis of every (nlay or nlay+1, ...) field: the
    same atmosphere stored top first (top_at_1 true)."""
    out = {}
    for k, v in arr.items():
        if v.ndim >= 2 and v.shape[0] in (nlay, nlay + 1):
            v = np.ascontiguousarray(v[::-1])
        elif v.ndim == 3 and v.shape[1] == nlay:
            v = np.ascontiguousarray(v[:, ::-1])
        out[k] = v
    return out


def _run_both(kd, lut, ncol, nlay, clouds, nangle, clrsky, jac, blksz=None, iovr="iovr_maxrand", scat=False,
              aer=False, top_at_1=False):
    from ccpprad import lib, device, capi
    from oracle import oracle as O
    pr, arr = _setup(kd, lut, ncol, nlay, clouds)
    if aer:
        arr.update(_aerosols(ncol, nlay, kd.nband, seed=11))
    if top_at_1:
        arr = _flip(arr, nlay)
    ints = capi.lw_default_ints(ncol, nlay, blksz=blksz, nangle=nangle, clrsky=clrsky, jacobian=jac, iovr=iovr,
                                scat=scat, top_at_1=top_at_1)
    ref = dict(arr)
    ref.update(O.lw_outputs(ncol, nlay))
    O.lw_main_run(kd, lut, ints, ref)
    s = lib.LWScheme(kd, lut, nlay, blksz=ints["rrtmgp_phys_blksz"])
    dev = device.to_device(arr)
    dev.update(device.lw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    return ref, out


def _check(ref, out, names, hr_names=("htrlw",)):
    for n in names:
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < FLUX_TOL, n
    for n in hr_names:
        mx, _ = report(n, out[n] * 86400, ref[n] * 86400, "K/day")
        assert mx < HR_TOL, n


def test_lw_clear_sky_one_angle(kd_lw, lut_lw):
    """128 columns x 64 layers, clear sky, 1 angle (optimal angles for the
    clear-sky call), Jacobian on."""
    ref, out = _run_both(kd_lw, lut_lw, 128, 64, clouds=False, nangle=1, clrsky=True, jac=True)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
                      "fluxlwUP_jac", "fluxlwUP_radtime", "fluxlwDOWN_radtime"], ("htrlw", "htrlwc"))


def test_lw_all_sky_three_angles(kd_lw, lut_lw):
    """All-sky, 127 layers, 3 Gauss angles, McICA max-random overlap."""
    ref, out = _run_both(kd_lw, lut_lw, 256, 127, clouds=True, nangle=3, clrsky=True, jac=False, blksz=64)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky"],
           ("htrlw", "htrlwc"))


@pytest.mark.parametrize("iovr", ["iovr_rand", "iovr_max", "iovr_exp", "iovr_dcorr"])
def test_lw_overlap_variants(kd_lw, lut_lw, iovr):
    ref, out = _run_both(kd_lw, lut_lw, 64, 127, clouds=True, nangle=2, clrsky=False, jac=False, iovr=iovr)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky"])


def test_lw_host_entry_matches_device_entry(kd_lw, lut_lw):
    """ccpprad_lw_run (host arrays, the Fortran drop-in path) equals the
    device entry bit for bit.  The outputs written in every column (fluxes,
    radtime copies, heating rate) are not copied in: they start as NaN on the
    host and must come back complete; the Jacobian, not requested, keeps its
    host values.  Two calls with different profiles on one context, so a
    staging buffer's contents from the first call cannot pass for the
    second's."""
    from ccpprad import lib, device, capi
    from oracle import oracle as O
    ncol, nlay = 64, 127
    ints = capi.lw_default_ints(ncol, nlay, nangle=3)
    s = lib.LWScheme(kd_lw, lut_lw, nlay, blksz=ncol)
    for seed in (3, 4):
        pr, arr = _setup(kd_lw, lut_lw, ncol, nlay, True, seed=seed)
        host = dict(arr)
        outs = O.lw_outputs(ncol, nlay)
        sentinel = {}
        for k in outs:
            outs[k][:] = np.nan
        outs["fluxlwUP_jac"][:] = 7.0
        sentinel["fluxlwUP_jac"] = outs["fluxlwUP_jac"].copy()
        host.update(outs)
        s.run(ints, host)
        dev = device.to_device(arr)
        dev.update(device.lw_outputs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
        out = device.to_host(dev)
        for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
                  "fluxlwUP_radtime", "fluxlwDOWN_radtime", "htrlw"):
            assert not np.isnan(host[n]).any(), n
            assert np.array_equal(out[n], host[n]), n
        assert np.array_equal(host["fluxlwUP_jac"], sentinel["fluxlwUP_jac"])


@pytest.mark.parametrize("nangle,jac", [(1, False), (2, True), (3, False), (4, True)])
def test_lw_aerosols(kd_lw, lut_lw, nangle, jac):
    """Aerosol depths by band added to the clear and all sky (1scl += 1scl),
    for each angle count, with and without the surface Jacobian (the
    solver's aerosol instantiations; one angle also runs the clear sky at
    the optimal angle)."""
    ref, out = _run_both(kd_lw, lut_lw, 128, 127, clouds=True, nangle=nangle, clrsky=True, jac=jac, blksz=64,
                         aer=True)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky"]
           + (["fluxlwUP_jac"] if jac else []), ("htrlw", "htrlwc"))


def test_lw_top_at_1(kd_lw, lut_lw):
    """The same profiles stored top first (top_at_1 true), all sky."""
    ref, out = _run_both(kd_lw, lut_lw, 128, 127, clouds=True, nangle=3, clrsky=True, jac=True, blksz=64,
                         top_at_1=True)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
                      "fluxlwUP_jac"], ("htrlw", "htrlwc"))


def test_lw_scattering_all_sky(kd_lw, lut_lw):
    """doGP_lwscat: all-sky fluxes from the two-stream solver with cloud
    scattering (the clear sky still from the no-scattering solver), 127
    layers, McICA max-random, with aerosols."""
    ref, out = _run_both(kd_lw, lut_lw, 256, 127, clouds=True, nangle=3, clrsky=True, jac=False, blksz=64,
                         scat=True, aer=True)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
                      "fluxlwUP_radtime", "fluxlwDOWN_radtime"], ("htrlw", "htrlwc"))


def test_lw_scattering_top_at_1_and_clear(kd_lw, lut_lw):
    """doGP_lwscat with the profiles stored top first, and a clear column set
    (no clouds: the two-stream solver with ssa = 0)."""
    ref, out = _run_both(kd_lw, lut_lw, 128, 127, clouds=True, nangle=1, clrsky=False, jac=False, blksz=32,
                         scat=True, top_at_1=True)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky"])
    ref, out = _run_both(kd_lw, lut_lw, 64, 64, clouds=False, nangle=1, clrsky=False, jac=False, scat=True)
    _check(ref, out, ["fluxlwUP_allsky", "fluxlwDOWN_allsky"])


def test_lw_scattering_rejects_jacobian(kd_lw, lut_lw):
    from ccpprad import lib, device, capi
    ncol, nlay = 32, 32
    pr, arr = _setup(kd_lw, lut_lw, ncol, nlay, False)
    ints = capi.lw_default_ints(ncol, nlay, jacobian=True, scat=True)
    s = lib.LWScheme(kd_lw, lut_lw, nlay, blksz=ncol)
    dev = device.to_device(arr)
    dev.update(device.lw_outputs(ncol, nlay))
    with pytest.raises(capi.CcppradError, match="Jacobian"):
        s.run_dev(ints, dev)
