"""End to end through the Fortran scheme modules: a cap-like Fortran driver
calls rrtmgp_lw_main_init/_run and rrtmgp_sw_main_init/_run with the
reference's argument lists; results are compared with the FP64 oracle."""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write_input(path, pr, arr, nday, blksz):
    with open(path, "wb") as f:
        np.array([pr.ncol, pr.nlay, nday, blksz], dtype=np.int32).tofile(f)
        for k in ("p_lay", "p_lev", "t_lay", "t_lev", "tsfg", "semis", "vmr_o2", "vmr_h2o", "vmr_o3",
                  "vmr_ch4", "vmr_n2o", "vmr_co2", "cld_frac", "cld_lwp", "cld_reliq", "cld_iwp",
                  "cld_reice", "coszen", "sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir",
                  "sfc_alb_uvvis_dif"):
            np.ascontiguousarray(arr[k], dtype=np.float64).tofile(f)
        np.ascontiguousarray(arr["idx"], dtype=np.int32).tofile(f)


def test_fortran_cap_driver_matches_oracle(tmp_path, kd_lw, kd_sw, lut_lw, lut_sw):
    from ccpprad import synthetic as syn, capi, tables, build
    from oracle import oracle as O
    drv = build.F_DRIVER
    if not os.path.exists(drv):
        pytest.fail("Fortran driver not built")
    ncol, nlay, blksz = 64, 127, 32
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
    raw = open(tmp_path / "out.bin", "rb").read()
    err = np.frombuffer(raw[:8], dtype=np.int32)
    assert (err == 0).all(), err
    nv = nlay + 1
    vals = np.frombuffer(raw[8:], dtype=np.float64)
    o = 0
    out = {}
    for k, n in (("lup", ncol * nv), ("ldn", ncol * nv), ("sup", ncol * nv), ("sdn", ncol * nv),
                 ("scmpsw", 6 * ncol), ("cldtau", ncol * nlay)):
        out[k] = vals[o:o + n]
        o += n
    # the driver passes the namelist nrghice = 3, which selects the cloud
    # file's third ice-roughness category
    lut_lw3, lut_sw3 = dataclasses.replace(lut_lw, icergh=3), dataclasses.replace(lut_sw, icergh=3)
    lw = dict(arr)
    lw.update(O.lw_outputs(ncol, nlay))
    O.lw_main_run(kd_lw, lut_lw3, capi.lw_default_ints(ncol, nlay, blksz=blksz, nangle=3), lw)
    sw = dict(arr)
    sw.update(O.sw_outputs(ncol, nlay))
    O.sw_main_run(kd_sw, lut_sw3, capi.sw_default_ints(ncol, ncol, nlay, blksz=blksz), sw)
    d = lambda a, b: float(np.abs(a.reshape(-1) - b.reshape(-1)).max())
    print("lw up", d(out["lup"], lw["fluxlwUP_allsky"]), "sw dn", d(out["sdn"], sw["fluxswDOWN_allsky"]))
    assert d(out["lup"], lw["fluxlwUP_allsky"]) < 0.05
    assert d(out["ldn"], lw["fluxlwDOWN_allsky"]) < 0.05
    assert d(out["sup"], sw["fluxswUP_allsky"]) < 0.05
    assert d(out["sdn"], sw["fluxswDOWN_allsky"]) < 0.05
    assert d(out["scmpsw"], sw["scmpsw"]) < 0.05
