"""GPU parity of the Pade cloud optics (doGP_cldoptics_PADE) against the FP64
oracle, all sky: LW without and with cloud scattering, and SW (flux 0.05
W/m2, heating rate 0.01 K/day).  The device evaluates the rational
approximants in FP32 from coefficients staged in LDS."""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["lw", "lw_scat", "sw"])
def test_pade_cloud_optics_all_sky(kd_lw, kd_sw, scheme):
    from ccpprad import capi, device, lib, synthetic as syn
    from oracle import oracle as O
    ncol, nlay = 128, 127
    lw = scheme != "sw"
    kd = kd_lw if lw else kd_sw
    lut = syn.make_cloud_pade(lw)
    pr = syn.make_profiles(ncol, nlay, seed=29, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    if lw:
        ints = capi.lw_default_ints(ncol, nlay, blksz=ncol, nangle=3, clrsky=True, scat=scheme == "lw_scat")
        s, outs, ref_out, run_ref = lib.LWScheme(kd, lut, nlay, blksz=ncol), device.lw_outputs, O.lw_outputs, \
            O.lw_main_run
        names, hr = ("fluxlwUP_allsky", "fluxlwDOWN_allsky"), "htrlw"
    else:
        arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
        ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=ncol, clrsky=True)
        s, outs, ref_out, run_ref = lib.SWScheme(kd, lut, nlay, blksz=ncol), device.sw_outputs, O.sw_outputs, \
            O.sw_main_run
        names, hr = ("fluxswUP_allsky", "fluxswDOWN_allsky"), "htrsw"
    print(f"Pade cloud optics: {scheme} {ncol} columns x {nlay} layers", flush=True)
    dev = device.to_device(arr)
    dev.update(outs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    ref = dict(arr)
    ref.update(ref_out(ncol, nlay))
    run_ref(kd, lut, ints, ref)
    # the clouds change the fluxes (the Pade path is exercised)
    clr = names[0].replace("allsky", "clrsky")
    assert np.max(np.abs(ref[names[0]] - ref[clr])) > 1.0
    for n in names:
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < 0.05, n
    mx, _ = report(hr, out[hr] * 86400, ref[hr] * 86400, "K/day")
    assert mx < 0.01


def test_pade_from_netcdf_matches_struct(tmp_path, kd_lw):
    """The Pade coefficients read from a netCDF cloud file give the same LW
    fluxes, bit for bit, as the same coefficients passed in the struct."""
    import dataclasses
    from ccpprad import capi, device, lib, synthetic as syn
    from nc_writer import write_kdist_nc, write_pade_nc
    lut = syn.make_cloud_pade(True)
    lut = dataclasses.replace(lut, icergh=lut.nrghice)   # the file path selects the last category
    write_kdist_nc(tmp_path / "k.nc", kd_lw)
    write_pade_nc(tmp_path / "c.nc", lut)
    ncol, nlay = 64, 127
    print(f"Pade cloud file: LW {ncol} columns", flush=True)
    arr = capi.profile_fields(syn.make_profiles(ncol, nlay, seed=31, clouds=True, kd=kd_lw))
    ints = capi.lw_default_ints(ncol, nlay, blksz=ncol, nangle=3)
    res = []
    for s in (lib.LWScheme(kd_lw, lut, nlay, blksz=ncol),
              lib.LWScheme.from_files(tmp_path / "k.nc", tmp_path / "c.nc", nlay, blksz=ncol)):
        dev = device.to_device(arr)
        dev.update(device.lw_outputs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
        res.append(device.to_host(dev))
    for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "htrlw"):
        assert np.array_equal(res[0][n], res[1][n]), n


@pytest.mark.parametrize("is_lw", [True, False])
def test_cloud_file_with_both_forms(tmp_path, kd_lw, kd_sw, is_lw):
    """A cloud file with lookup tables and Pade coefficients (as the
    reference's files): the Pade entry (doGP_cldoptics_PADE) gives the Pade
    struct's results, the plain entry the lookup tables', bit for bit."""
    import dataclasses
    from ccpprad import capi, device, lib, synthetic as syn
    from nc_writer import write_kdist_nc, write_pade_nc
    kd = kd_lw if is_lw else kd_sw
    both = syn.make_cloud_pade(is_lw)
    both = dataclasses.replace(both, icergh=both.nrghice)
    write_kdist_nc(tmp_path / "k.nc", kd)
    write_pade_nc(tmp_path / "c.nc", both, with_lut=True)
    lut_only = dataclasses.replace(syn.make_cloud_lut(is_lw), icergh=both.nrghice)
    ncol, nlay = 64, 127
    print(f"cloud file with both forms: {'LW' if is_lw else 'SW'} {ncol} columns", flush=True)
    arr = capi.profile_fields(syn.make_profiles(ncol, nlay, seed=37, clouds=True, kd=kd))
    cls = lib.LWScheme if is_lw else lib.SWScheme
    if is_lw:
        ints, outs, names = capi.lw_default_ints(ncol, nlay, blksz=ncol, nangle=3), device.lw_outputs, \
            ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "htrlw")
    else:
        arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
        ints, outs, names = capi.sw_default_ints(ncol, ncol, nlay, blksz=ncol), device.sw_outputs, \
            ("fluxswUP_allsky", "fluxswDOWN_allsky", "htrsw")

    def run(s):
        dev = device.to_device(arr)
        dev.update(outs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
        return device.to_host(dev)

    f_pade = run(cls.from_files(tmp_path / "k.nc", tmp_path / "c.nc", nlay, blksz=ncol, pade=True))
    f_lut = run(cls.from_files(tmp_path / "k.nc", tmp_path / "c.nc", nlay, blksz=ncol))
    s_pade = run(cls(kd, both, nlay, blksz=ncol))
    s_lut = run(cls(kd, lut_only, nlay, blksz=ncol))
    for n in names:
        assert np.array_equal(f_pade[n], s_pade[n]), n
        assert np.array_equal(f_lut[n], s_lut[n]), n
    assert not np.array_equal(f_pade[names[0]], f_lut[names[0]])
