"""Gas optics and fluxes with k-distributions at the device limits (16
flavors, 32 minor intervals per regime, a g-point inside 8 of them: the
tables of test_table_limits.py) against the FP64 oracle, and the per-g-point
limit (a ninth interval over one g-point) as an init error.  Tolerances as
for the synthetic tables: tau 2e-5 relative, Planck fraction 2e-6, fluxes
0.05 W/m2, heating rates 0.01 K/day."""
import numpy as np
import pytest

from test_table_limits import MAX_MINOR_G, table_at_limits
from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _profiles(kd, ncol, nlay, seed):
    from ccpprad import synthetic as syn, capi
    pr = syn.make_profiles(ncol, nlay, seed=seed, clouds=True, kd=kd)
    return pr, capi.profile_fields(pr)


def test_lw_at_the_limits(lut_lw):
    from ccpprad import lib, capi, device
    from oracle import oracle as O
    kd = table_at_limits(True)
    ncol, nlay = 96, 127
    pr, arr = _profiles(kd, ncol, nlay, 41)
    tau_o, pf_o = O.gas_optics(kd, pr)
    s = lib.LWScheme(kd, lut_lw, nlay, blksz=32)
    v = torch.from_numpy(O.vmr_stack(pr)).cuda()
    d = device.to_device({k: arr[k] for k in ("p_lay", "p_lev", "t_lay")})
    tau = torch.zeros((ncol, nlay, kd.ngpt), dtype=torch.float32, device="cuda")
    pf = torch.zeros_like(tau)
    s.gas_optics_dev(ncol, nlay, d["p_lay"], d["p_lev"], d["t_lay"], v, tau, pf)
    torch.cuda.synchronize()
    tg, pg = tau.cpu().numpy().astype(np.float64), pf.cpu().numpy().astype(np.float64)
    rel = np.abs(tg - tau_o) / np.maximum(np.abs(tau_o), 1e-30)
    print("tau rel max", rel.max(), "pfrac abs max", np.abs(pg - pf_o).max())
    assert rel.max() < 2e-5 and np.abs(pg - pf_o).max() < 2e-6
    ints = capi.lw_default_ints(ncol, nlay, blksz=32, nangle=3, clrsky=True)
    dev = device.to_device(arr)
    dev.update(device.lw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    ref = dict(arr)
    ref.update(O.lw_outputs(ncol, nlay))
    O.lw_main_run(kd, lut_lw, ints, ref)
    for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky"):
        mx, _ = report(n, dev[n].cpu().numpy(), ref[n], "W/m2")
        assert mx < 0.05, n
    mx, _ = report("htrlw", dev["htrlw"].cpu().numpy() * 86400, ref["htrlw"] * 86400, "K/day")
    assert mx < 0.01


def test_sw_at_the_limits(lut_sw):
    from ccpprad import lib, capi, device
    from oracle import oracle as O
    kd = table_at_limits(False)
    ncol, nlay = 96, 127
    pr, arr = _profiles(kd, ncol, nlay, 42)
    arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
    s = lib.SWScheme(kd, lut_sw, nlay, blksz=32)
    ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=32, clrsky=True)
    dev = device.to_device(arr)
    dev.update(device.sw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    ref = dict(arr)
    ref.update(O.sw_outputs(ncol, nlay))
    O.sw_main_run(kd, lut_sw, ints, ref)
    for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky"):
        mx, _ = report(n, dev[n].cpu().numpy(), ref[n], "W/m2")
        assert mx < 0.05, n
    mx, _ = report("htrsw", dev["htrsw"].cpu().numpy() * 86400, ref["htrsw"] * 86400, "K/day")
    assert mx < 0.01


def test_nine_intervals_over_one_gpoint_is_an_init_error(lut_lw):
    from ccpprad import lib, capi
    kd = table_at_limits(True, per_g=MAX_MINOR_G + 1)
    with pytest.raises(capi.CcppradError, match="minor absorbers"):
        lib.LWScheme(kd, lut_lw, 127)
