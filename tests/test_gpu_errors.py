"""Error paths through the C-ABI and the Fortran cap: each is forced from the
host, every kernel still exits normally, and the call returns a nonzero
status with a message (the Fortran modules map it to errflg = 1; the
reference prints rte-rrtmgp errors and keeps errflg = 0,
radiation_tools.F90:89-98, a documented deviation)."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _lw_setup(kd, lut, ncol=64, nlay=127, blksz=None):
    from ccpprad import synthetic as syn, capi, lib
    pr = syn.make_profiles(ncol, nlay, seed=91, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    ints = capi.lw_default_ints(ncol, nlay, blksz=blksz, nangle=3)
    s = lib.LWScheme(kd, lut, nlay, blksz=ints["rrtmgp_phys_blksz"])
    return s, arr, ints


def _sw_setup(kd, lut, ncol=64, nlay=127, blksz=None):
    from ccpprad import synthetic as syn, capi, lib
    pr = syn.make_profiles(ncol, nlay, seed=92, clouds=True, kd=kd)
    arr = capi.profile_fields(pr)
    arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
    ints = capi.sw_default_ints(ncol, ncol, nlay, blksz=blksz)
    s = lib.SWScheme(kd, lut, nlay, blksz=ints["rrtmgp_phys_blksz"])
    return s, arr, ints


def _run(s, arr, ints, lw=True):
    from ccpprad import device
    dev = device.to_device(arr)
    nc, nl = ints["nCol"], ints["nLay"]
    dev.update(device.lw_outputs(nc, nl) if lw else device.sw_outputs(nc, nl))
    s.run_dev(ints, dev)
    torch.cuda.synchronize()
    return dev


@pytest.mark.parametrize("field,value", [("t_lay", 400.0), ("t_lay", 100.0), ("p_lay", 2.0e5)])
@pytest.mark.parametrize("lw", [True, False])
def test_table_range_is_status_3(kd_lw, kd_sw, lut_lw, lut_sw, lw, field, value):
    """A layer outside the k-distribution's p/T range: status 3 with the
    gas_optics message (the kernels clamp their table indices)."""
    from ccpprad import capi
    s, arr, ints = (_lw_setup(kd_lw, lut_lw) if lw else _sw_setup(kd_sw, lut_sw))
    arr[field] = arr[field].copy()
    arr[field][40, 7] = value
    with pytest.raises(capi.CcppradError, match="outside the table range"):
        _run(s, arr, ints, lw)
    # the context stays usable: a good call after the bad one succeeds
    arr[field][40, 7] = arr[field][41, 7]
    _run(s, arr, ints, lw)


@pytest.mark.parametrize("lw", [True, False])
def test_radius_outside_lut_is_status_3(kd_lw, kd_sw, lut_lw, lut_sw, lw):
    from ccpprad import capi
    s, arr, ints = (_lw_setup(kd_lw, lut_lw) if lw else _sw_setup(kd_sw, lut_sw))
    l, c = np.argwhere(arr["cld_frac"] > 0.0)[0]
    arr["cld_reliq"] = arr["cld_reliq"].copy()
    arr["cld_reliq"][l, c] = 40.0          # the liquid LUT ends at 21.5 um
    with pytest.raises(capi.CcppradError, match="effective radius outside"):
        _run(s, arr, ints, lw)


@pytest.mark.parametrize("lw", [True, False])
def test_block_size_must_divide_columns(kd_lw, kd_sw, lut_lw, lut_sw, lw):
    from ccpprad import capi
    s, arr, ints = (_lw_setup(kd_lw, lut_lw, blksz=24) if lw else _sw_setup(kd_sw, lut_sw, blksz=24))
    with pytest.raises(capi.CcppradError, match="multiple of rrtmgp_phys_blksz"):
        _run(s, arr, ints, lw)


@pytest.mark.parametrize("lw", [True, False])
def test_nlay_must_match_init(kd_lw, kd_sw, lut_lw, lut_sw, lw):
    from ccpprad import capi, lib, synthetic as syn
    nlay = 64
    s = lib.LWScheme(kd_lw, lut_lw, 127) if lw else lib.SWScheme(kd_sw, lut_sw, 127)
    pr = syn.make_profiles(32, nlay, seed=93, clouds=True, kd=kd_lw if lw else kd_sw)
    arr = capi.profile_fields(pr)
    arr["idx"] = np.arange(1, 33, dtype=np.int32)
    ints = capi.lw_default_ints(32, nlay) if lw else capi.sw_default_ints(32, 32, nlay)
    with pytest.raises(capi.CcppradError, match="differs from init"):
        _run(s, arr, ints, lw)


@pytest.mark.parametrize("missing", ["p_lay", "t_lev", "cld_frac"])
def test_null_required_array(kd_lw, lut_lw, missing):
    from ccpprad import capi
    s, arr, ints = _lw_setup(kd_lw, lut_lw)
    del arr[missing]
    with pytest.raises(capi.CcppradError, match="NULL"):
        _run(s, arr, ints, True)


@pytest.mark.parametrize("missing", ["aersw_ssa", "aersw_g"])
def test_sw_aerosols_all_or_none(kd_sw, lut_sw, missing):
    """SW aerosol optics come as the three fields together: one left out is
    an argument error, not a read through a NULL pointer."""
    from ccpprad import capi
    s, arr, ints = _sw_setup(kd_sw, lut_sw)
    shp = (kd_sw.nband, ints["nLay"], ints["nCol"])
    arr.update(aersw_tau=np.full(shp, 0.01), aersw_ssa=np.full(shp, 0.9), aersw_g=np.full(shp, 0.6))
    del arr[missing]
    with pytest.raises(capi.CcppradError, match="aersw_tau, aersw_ssa and aersw_g"):
        _run(s, arr, ints, False)


def test_sw_null_output(kd_sw, lut_sw):
    from ccpprad import capi, device
    s, arr, ints = _sw_setup(kd_sw, lut_sw)
    dev = device.to_device(arr)
    out = device.sw_outputs(ints["nCol"], ints["nLay"])
    del out["scmpsw"]
    dev.update(out)
    with pytest.raises(capi.CcppradError, match="NULL"):
        s.run_dev(ints, dev)


def test_mcica_mask_entry_errors(kd_lw, lut_lw):
    from ccpprad import capi
    s, arr, ints = _lw_setup(kd_lw, lut_lw)
    words = torch.zeros((64, 127, 4), dtype=torch.int32, device="cuda")
    dev = {"cld_frac": torch.from_numpy(arr["cld_frac"]).cuda()}
    bad = dict(ints, isubc_lw=2)
    with pytest.raises(capi.CcppradError, match="icseed"):
        s.mcica_mask_dev(bad, dev, words)
    bad = dict(ints, iovr=capi.IOVR["iovr_exp"])
    with pytest.raises(capi.CcppradError, match="cloud_overlap_param"):
        s.mcica_mask_dev(bad, dev, words)


def test_fortran_cap_maps_errors_to_errflg(tmp_path, kd_lw, kd_sw, lut_lw, lut_sw):
    """Through the Fortran scheme modules: a block size that does not divide
    nCol gives errflg = 1 and the library's message in errmsg, for LW and SW."""
    from ccpprad import synthetic as syn, capi, tables, build
    from test_gpu_fortran import _write_input
    drv = build.F_DRIVER
    assert os.path.exists(drv), "Fortran driver not built"
    ncol, nlay, blksz = 64, 127, 24
    tables.write_kdist(tmp_path / "lw_gas.tab", kd_lw)
    tables.write_lut(tmp_path / "lw_cloud.tab", lut_lw)
    tables.write_kdist(tmp_path / "sw_gas.tab", kd_sw)
    tables.write_lut(tmp_path / "sw_cloud.tab", lut_sw)
    pr = syn.make_profiles(ncol, nlay, seed=77, clouds=True, kd=kd_lw)
    arr = capi.profile_fields(pr)
    arr["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
    _write_input(tmp_path / "in.bin", pr, arr, ncol, blksz)
    r = subprocess.run([drv, str(tmp_path) + "/", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    err = np.frombuffer(open(tmp_path / "out.bin", "rb").read()[:8], dtype=np.int32)
    assert (err == 1).all(), err
    assert "lw run:" in r.stdout and "sw run:" in r.stdout and "multiple of rrtmgp_phys_blksz" in r.stdout


def test_nrghice_selects_the_ice_roughness(tmp_path, kd_lw, lut_lw):
    """The namelist's nrghice picks that ice-roughness category of the cloud
    file (rrtmgp_lw_cloud_optics.F90:171-173,377): nrghice = 1 equals the
    in-memory LUT at category 1 bit for bit and differs from category 3;
    beyond the file's count it is an error."""
    import dataclasses
    from ccpprad import lib, tables, capi
    tables.write_kdist(tmp_path / "k.tab", kd_lw)
    tables.write_lut(tmp_path / "c.tab", lut_lw)
    s_file1 = lib.LWScheme.from_files(tmp_path / "k.tab", tmp_path / "c.tab", 127, blksz=64, nrghice=1)
    s_file3 = lib.LWScheme.from_files(tmp_path / "k.tab", tmp_path / "c.tab", 127, blksz=64, nrghice=3)
    s_mem1 = lib.LWScheme(kd_lw, dataclasses.replace(lut_lw, icergh=1), 127, blksz=64)
    _, arr, ints = _lw_setup(kd_lw, lut_lw)
    o1 = _run(s_file1, arr, ints)["fluxlwUP_allsky"].cpu().numpy()
    o3 = _run(s_file3, arr, ints)["fluxlwUP_allsky"].cpu().numpy()
    m1 = _run(s_mem1, arr, ints)["fluxlwUP_allsky"].cpu().numpy()
    assert np.array_equal(o1, m1)
    assert not np.array_equal(o1, o3)
    with pytest.raises(capi.CcppradError, match="nrghice"):
        lib.LWScheme.from_files(tmp_path / "k.tab", tmp_path / "c.tab", 127, blksz=64, nrghice=4)
