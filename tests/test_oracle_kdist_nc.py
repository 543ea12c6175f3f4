"""The k-distribution and cloud tables in the reference's netCDF layout.

Oracle pins: the oracle's restatement of rte-rrtmgp's load reduction, applied
to a file written from a synthetic table with two extra gases (one shifting
the others' file positions) and minor intervals on them interleaved among the
real ones, gives back the original table: the same gas pair at every
g-point and layer, the same vmr_ref columns, the same minor intervals with
their contributors repacked.  A key species that is not active is an error.

Library (no GPU needed): the native classic-format reader parses the same
files as CDF-1, CDF-2 and CDF-5 up to the device upload, and refuses netCDF-4
(HDF5) containers, truncated headers, missing variables and inactive key
species with a message."""
import ctypes as C

import numpy as np
import pytest

from nc_writer import to_cdf5, write_kdist_nc, write_lut_nc


def _pairs(kd):
    return kd.flavor[kd.gpoint_flavor - 1]          # (ngpt, 2 layers, 2 gases)


@pytest.mark.parametrize("is_lw", [True, False])
def test_oracle_reduction_round_trip(tmp_path, is_lw):
    from ccpprad import synthetic as syn
    from oracle import kdist_nc
    kd = syn.make_kdist(is_lw)
    write_kdist_nc(tmp_path / "k.nc", kd)
    r = kdist_nc.load_kdist(tmp_path / "k.nc", is_lw)
    np.testing.assert_array_equal(_pairs(r), _pairs(kd))
    assert r.nflav <= kd.nflav
    for k in ("band_lims_gpt", "band_lims_wvn", "press_ref", "temp_ref", "vmr_ref", "kmajor"):
        np.testing.assert_array_equal(getattr(r, k), getattr(kd, k), err_msg=k)
    assert r.press_ref_trop == kd.press_ref_trop
    for rg in ("minor_lower", "minor_upper"):
        a, b = getattr(r, rg), getattr(kd, rg)
        assert a["n"] == b["n"] and a["ncontrib"] == b["ncontrib"], rg
        for k in ("limits_gpt", "scales_with_density", "idx_scaling", "scale_by_complement", "idx_minor",
                  "kminor_start", "kminor"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{rg} {k}")
    if is_lw:
        for k in ("planck_frac", "totplnk", "optimal_angle_fit"):
            np.testing.assert_array_equal(getattr(r, k), getattr(kd, k), err_msg=k)
    else:
        for k in ("krayl_lower", "krayl_upper", "solar_src"):
            np.testing.assert_array_equal(getattr(r, k), getattr(kd, k), err_msg=k)


def test_oracle_minor_scaled_by_an_inactive_gas(tmp_path):
    """An interval whose gas is active but whose scaling gas is not stays,
    with no scaling gas (index 0), and its contributors are appended."""
    from ccpprad import synthetic as syn
    from oracle import kdist_nc
    kd = syn.make_kdist(True)
    write_kdist_nc(tmp_path / "k.nc", kd, scaled_by_inactive=True)
    r = kdist_nc.load_kdist(tmp_path / "k.nc", True)
    a, b = r.minor_lower, kd.minor_lower
    assert a["n"] == b["n"] + 1 and a["ncontrib"] == b["ncontrib"] + 4
    assert a["idx_minor"][-1] == 3 and a["idx_scaling"][-1] == 0 and a["scales_with_density"][-1] == 1
    assert list(a["limits_gpt"][-1]) == [1, 4] and a["kminor_start"][-1] == b["ncontrib"] + 1
    np.testing.assert_array_equal(a["kminor"][:, :, :b["ncontrib"]], b["kminor"])


def test_oracle_inactive_key_species_is_an_error(tmp_path):
    from ccpprad import synthetic as syn
    from oracle import kdist_nc
    write_kdist_nc(tmp_path / "k.nc", syn.make_kdist(True))
    with pytest.raises(ValueError, match="not active"):
        kdist_nc.load_kdist(tmp_path / "k.nc", True, active=("co2", "o3", "n2o", "ch4", "o2"))


def _init(L, kfile, cfile, names=None, lw=True):
    from ccpprad import capi
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    arr = None
    if names is not None:
        arr = (C.c_char_p * len(names))(*[s.encode() for s in names])
    fn = L.ccpprad_lw_init_files if lw else L.ccpprad_sw_init_files
    rc = fn(str(kfile).encode(), str(cfile).encode() if cfile else b"", 0, arr, len(names or []), 127, 1,
            C.byref(ctx), buf, n)
    return rc, buf.value.decode(errors="replace")


@pytest.mark.parametrize("version", [1, 2, 5])
@pytest.mark.parametrize("is_lw", [True, False])
def test_library_parses_netcdf_tables(tmp_path, version, is_lw):
    import torch
    from ccpprad import lib, synthetic as syn
    if torch.cuda.is_available():
        pytest.skip("covered by the GPU tests")
    L = lib.load()
    write_kdist_nc(tmp_path / "k.nc", syn.make_kdist(is_lw), version=min(version, 2))
    write_lut_nc(tmp_path / "c.nc", syn.make_cloud_lut(is_lw), version=min(version, 2))
    if version == 5:
        for n in ("k", "c"):
            to_cdf5(tmp_path / f"{n}.nc", tmp_path / f"{n}5.nc")
            (tmp_path / f"{n}5.nc").replace(tmp_path / f"{n}.nc")
        assert (tmp_path / "k.nc").read_bytes()[:4] == b"CDF\x05"
    rc, msg = _init(L, tmp_path / "k.nc", tmp_path / "c.nc", syn.GAS_NAMES, lw=is_lw)
    # no GPU here: the tables parse and reduce, then the device upload fails
    assert rc != 0
    for bad in ("lacks", "malformed", "netCDF", "key species", "other spectral range", "not in the k-distribution"):
        assert bad not in msg, msg


def test_library_rejects_bad_netcdf(tmp_path):
    from ccpprad import lib, synthetic as syn
    L = lib.load()
    h5 = tmp_path / "h5.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    rc, msg = _init(L, h5, None, syn.GAS_NAMES)
    assert rc != 0 and "HDF5" in msg and "nccopy" in msg
    write_kdist_nc(tmp_path / "k.nc", syn.make_kdist(True))
    raw = (tmp_path / "k.nc").read_bytes()
    (tmp_path / "t.nc").write_bytes(raw[:200])
    rc, msg = _init(L, tmp_path / "t.nc", None, syn.GAS_NAMES)
    assert rc != 0 and "malformed" in msg
    rc, msg = _init(L, tmp_path / "k.nc", None, ["co2", "o3", "n2o", "ch4", "o2"])
    assert rc != 0 and "key species 'h2o'" in msg
    rc, msg = _init(L, tmp_path / "k.nc", None, syn.GAS_NAMES, lw=False)
    assert rc != 0 and "other spectral range" in msg
    # a file without kmajor
    from scipy.io import netcdf_file
    src = netcdf_file(tmp_path / "k.nc", "r", mmap=False)
    dst = netcdf_file(tmp_path / "m.nc", "w", version=2)
    for d, v in src.dimensions.items():
        dst.createDimension(d, v)
    for name, v in src.variables.items():
        if name == "kmajor":
            continue
        w = dst.createVariable(name, v.typecode(), v.dimensions)
        w[...] = v[...]
    src.close()
    dst.close()
    rc, msg = _init(L, tmp_path / "m.nc", None, syn.GAS_NAMES)
    assert rc != 0 and "lacks kmajor" in msg
