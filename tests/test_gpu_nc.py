"""GPU: schemes initialised from the reference's netCDF table files (read by
the library's native classic-format reader and reduced to the active gases)
give the same fluxes and heating rates, bit for bit, as schemes built
directly from the tables the files were written from.  The files carry two
extra gases and minor intervals on them, so the reduction has to drop and
repack them and renumber the gases; flavors may come out in another order,
which leaves every result unchanged."""
import numpy as np
import pytest

from nc_writer import to_cdf5, write_kdist_nc, write_lut_nc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cdf", [2, 5])
@pytest.mark.parametrize("is_lw", [True, False])
def test_netcdf_tables_match_direct_tables(tmp_path, is_lw, cdf):
    from ccpprad import capi, device, lib, synthetic as syn
    kd, lut = syn.make_kdist(is_lw), syn.make_cloud_lut(is_lw)
    lut.icergh = lut.nrghice           # the netCDF path selects the last roughness category
    write_kdist_nc(tmp_path / "k.nc", kd)
    write_lut_nc(tmp_path / "c.nc", lut)
    if cdf == 5:                       # the format `nccopy -k cdf5` writes
        for n in ("k", "c"):
            to_cdf5(tmp_path / f"{n}.nc", tmp_path / f"{n}5.nc")
            (tmp_path / f"{n}5.nc").replace(tmp_path / f"{n}.nc")
    ncol, nlay = 128, 127
    print(f"netCDF tables: {'LW' if is_lw else 'SW'} {ncol} columns x {nlay} layers", flush=True)
    pr = syn.make_profiles(ncol, nlay, seed=17, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    if is_lw:
        cls, ints, outs = lib.LWScheme, capi.lw_default_ints(ncol, nlay, blksz=ncol, nangle=3, clrsky=True), \
            device.lw_outputs
    else:
        arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
        cls, ints, outs = lib.SWScheme, capi.sw_default_ints(ncol, ncol, nlay, blksz=ncol, clrsky=True), \
            device.sw_outputs
    direct = cls(kd, lut, nlay, blksz=ncol)
    from_nc = cls.from_files(tmp_path / "k.nc", tmp_path / "c.nc", nlay, blksz=ncol)
    assert from_nc.ngpt == kd.ngpt and from_nc.nband == kd.nband
    res = []
    for s in (direct, from_nc):
        dev = device.to_device(arr)
        dev.update(outs(ncol, nlay))
        s.run_dev(ints, dev)
        torch.cuda.synchronize()
        res.append(device.to_host(dev))
    names = [k for k in outs(1, 2)]
    for n in names:
        assert np.array_equal(res[0][n], res[1][n]), n
    assert np.any(res[0][names[0]] != 0.0)


def test_netcdf_minor_scaled_by_inactive_gas_matches_oracle(tmp_path):
    """A file whose extra o3 interval is scaled by n2, which the run does not
    use: the library keeps the interval without its scaling, as the oracle's
    reduction does; all-sky LW against the oracle run on that reduction."""
    from ccpprad import capi, device, lib, synthetic as syn
    from oracle import kdist_nc, oracle as O
    from util import report
    kd, lut = syn.make_kdist(True), syn.make_cloud_lut(True)
    lut.icergh = lut.nrghice
    write_kdist_nc(tmp_path / "k.nc", kd, scaled_by_inactive=True)
    write_lut_nc(tmp_path / "c.nc", lut)
    kr = kdist_nc.load_kdist(tmp_path / "k.nc", True)
    ncol, nlay = 64, 127
    print(f"netCDF tables, o3 interval scaled by inactive n2: LW {ncol} columns", flush=True)
    pr = syn.make_profiles(ncol, nlay, seed=19, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    ints = capi.lw_default_ints(ncol, nlay, blksz=ncol, nangle=3)
    ref = dict(arr)
    ref.update(O.lw_outputs(ncol, nlay))
    O.lw_main_run(kr, lut, ints, ref)
    s = lib.LWScheme.from_files(tmp_path / "k.nc", tmp_path / "c.nc", nlay, blksz=ncol)
    dev = device.to_device(arr)
    dev.update(device.lw_outputs(ncol, nlay))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    out = device.to_host(dev)
    for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky"):
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < 0.05, n
    mx, _ = report("htrlw", out["htrlw"] * 86400, ref["htrlw"] * 86400, "K/day")
    assert mx < 0.01
