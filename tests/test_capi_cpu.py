"""CPU checks of the C-ABI library: it builds, loads without a GPU and
exports every entry point include/ccpprad.h declares."""
import ctypes as C
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(REPO, "include", "ccpprad.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ccpprad_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from ccpprad import lib, build
    lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", build.HIP_LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ccpprad_\w+)", out))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    assert set(lib.EXPORTS) <= exported


def test_version_string_without_gpu():
    from ccpprad import lib
    assert b"gfx950" in lib.load().ccpprad_version()


def test_fortran_modules_built_with_reference_names():
    """The Fortran layer exports the reference's scheme entry points."""
    from ccpprad import build
    if not os.path.exists(build.F_LIB):
        pytest.skip("flang not available")
    out = subprocess.run(["nm", "-D", "--defined-only", build.F_LIB], capture_output=True, text=True).stdout
    for sym in ("rrtmgp_lw_main_init", "rrtmgp_lw_main_run", "rrtmgp_sw_main_init", "rrtmgp_sw_main_run",
                "compute_heating_rate", "gfs_rrtmgp_post_run", "gfs_rrtmgp_cloud_overlap_run"):
        assert sym in out, sym
    # the device release is no CCPP phase name (the unchanged .meta has no table for it)
    assert "rrtmgp_lw_main_finalize" not in out and "rrtmgp_lw_main_release" in out


def test_table_file_loader_rejects_malformed_files(tmp_path):
    """ccpprad_lw_init_files parses on the host: a bad file is an error (no GPU needed)."""
    import ctypes as C
    from ccpprad import lib, capi
    L = lib.load()
    bad = tmp_path / "bad.tab"
    bad.write_bytes(b"NOTATABLE" + b"\0" * 64)
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    rc = L.ccpprad_lw_init_files(str(bad).encode(), b"", 0, None, 0, 127, 1, C.byref(ctx), buf, n)
    assert rc != 0 and b"malformed" in buf.value
    rc = L.ccpprad_lw_init_files(str(tmp_path / "missing.tab").encode(), b"", 0, None, 0, 127, 1,
                                 C.byref(ctx), buf, n)
    assert rc != 0 and b"cannot open" in buf.value


def test_table_file_round_trip_parses(tmp_path):
    """A written table parses fully: with no GPU the library reaches the
    device upload and reports that (not a parse error)."""
    import ctypes as C
    import torch
    from ccpprad import lib, capi, tables, synthetic as syn
    if torch.cuda.is_available():
        pytest.skip("covered by the GPU tests")
    L = lib.load()
    tables.write_kdist(tmp_path / "k.tab", syn.make_kdist(True))
    tables.write_lut(tmp_path / "c.tab", syn.make_cloud_lut(True))
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    rc = L.ccpprad_lw_init_files(str(tmp_path / "k.tab").encode(), str(tmp_path / "c.tab").encode(), 0, None, 0,
                                 127, 1, C.byref(ctx), buf, n)
    assert rc != 0 and b"malformed" not in buf.value and b"lacks" not in buf.value


def test_unknown_active_gas_is_an_init_error():
    """An active gas the k-distribution lacks is rejected at init, on the
    host, before any device work."""
    import ctypes as C
    from ccpprad import lib, capi, synthetic as syn
    L = lib.load()
    hk = capi.kdist_struct(syn.make_kdist(True))
    names = (C.c_char_p * 3)(b"h2o", b"co2", b"cfc11")
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    rc = L.ccpprad_lw_init(C.addressof(hk.struct), None, names, 3, 127, 1, C.byref(ctx), buf, n)
    assert rc != 0 and b"cfc11" in buf.value


def test_nrghice_beyond_the_file_is_an_init_error(tmp_path):
    """nrghice selects a category of the cloud file; beyond the file's count
    it is an error before any device work."""
    import ctypes as C
    from ccpprad import lib, capi, tables, synthetic as syn
    L = lib.load()
    tables.write_kdist(tmp_path / "k.tab", syn.make_kdist(True))
    tables.write_lut(tmp_path / "c.tab", syn.make_cloud_lut(True))
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    rc = L.ccpprad_lw_init_files(str(tmp_path / "k.tab").encode(), str(tmp_path / "c.tab").encode(), 4, None, 0,
                                 127, 1, C.byref(ctx), buf, n)
    assert rc != 0 and b"nrghice=4" in buf.value


def test_pade_bounds_that_need_the_regime_clamp_are_rejected():
    """Pade size-regime bounds with b2 - b3 > b1 would let rte's regime rule
    fall below 1 (where this build clamps): rejected at init, on the host."""
    import ctypes as C
    import dataclasses
    import numpy as np
    from ccpprad import lib, capi, synthetic as syn
    L = lib.load()
    hk = capi.kdist_struct(syn.make_kdist(True))
    lut = syn.make_cloud_pade(True)
    pade = dict(lut.pade)
    pade["pade_sizreg_extliq"] = np.array([2.5, 15.0, 5.0, 21.5])   # 15 - 5 > 2.5
    bad = dataclasses.replace(lut, pade=pade)
    hl = capi.lut_struct(bad)
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    rc = L.ccpprad_lw_init(C.addressof(hk.struct), C.addressof(hl.struct), None, 0, 127, 1, C.byref(ctx), buf, n)
    assert rc != 0 and b"b2 - b3 <= b1" in buf.value


REF_PHYSICS = "/root/reference/physics"
DROP_INS = [
    ("rrtmgp_lw_main.F90", "Radiation/RRTMGP/rrtmgp_lw_main.F90", "rrtmgp_lw_main_init"),
    ("rrtmgp_lw_main.F90", "Radiation/RRTMGP/rrtmgp_lw_main.F90", "rrtmgp_lw_main_run"),
    ("rrtmgp_sw_main.F90", "Radiation/RRTMGP/rrtmgp_sw_main.F90", "rrtmgp_sw_main_init"),
    ("rrtmgp_sw_main.F90", "Radiation/RRTMGP/rrtmgp_sw_main.F90", "rrtmgp_sw_main_run"),
    ("GFS_rrtmgp_post.F90", "Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_post.F90", "GFS_rrtmgp_post_run"),
    ("GFS_rrtmgp_cloud_overlap.F90", "Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_cloud_overlap.F90",
     "GFS_rrtmgp_cloud_overlap_run"),
]


@pytest.mark.parametrize("ours,theirs,name", DROP_INS)
def test_fortran_drop_ins_keep_the_reference_argument_lists(ours, theirs, name):
    """The scheme subroutines' dummy-argument lists, name for name and in
    order, against the reference's sources (read as text where they are
    mounted), so the reference's .meta tables bind them unchanged."""
    import re
    from ccpprad import build
    ref = os.path.join(REF_PHYSICS, theirs)
    if not os.path.exists(ref):
        pytest.skip("reference sources not mounted")

    def args(path):
        txt = open(path).read()
        m = re.search(r"subroutine\s+" + name + r"\s*\((.*?)\)", txt, re.S | re.I)
        assert m, (path, name)
        body = re.sub(r"!.*", "", m.group(1))
        return [a.lower() for a in re.findall(r"\w+", body)]

    assert args(os.path.join(build.FSRC, ours)) == args(ref)


def _post_call(ints, hr_ints, arrays):
    from ccpprad import lib, capi
    L = lib.load()
    a = capi.post_struct(ints, dict(fhlwr=3600.0, fhswr=3600.0, raddt=1800.0), arrays)
    h = capi.post_hr_struct(hr_ints, arrays)
    buf, n = capi.errbuf()
    rc = L.ccpprad_post(C.addressof(a.struct), C.addressof(h.struct), buf, n)
    return rc, buf.value


def test_post_host_entry_argument_errors():
    """ccpprad_post (GFS_rrtmgp_post_run on host arrays) checks its heating-rate
    arguments on the host before any device work: a requested clear-sky
    heating rate without its array, nDay outside 0..nCol, daylit indices
    outside 1..nCol or repeated; a call with neither LW nor SW does nothing."""
    import numpy as np
    nc, nl = 8, 4
    ints = dict(nCol=nc, nLev=nl, nDay=3, iSFC=1, iTOA=nl + 1, doLWrad=1, doSWrad=1, save_diag=0, naerodp=6,
                nfxr=45)
    z = lambda *s: np.zeros(s)
    arrays = dict(htrlw=z(nl, nc), htrsw=z(nl, nc), idxday=np.array([1, 4, 6], dtype=np.int32))
    rc, msg = _post_call(ints, dict(do_lw_clrsky_hr=1), arrays)
    assert rc == 1 and b"htrlwc is NULL" in msg
    rc, msg = _post_call(ints, dict(do_sw_clrsky_hr=1), arrays)
    assert rc == 1 and b"htrswc is NULL" in msg
    rc, msg = _post_call(dict(ints, nDay=9), {}, arrays)
    assert rc == 1 and b"nDay" in msg
    for bad in ([1, 4, 9], [0, 4, 6], [4, 4, 6]):
        rc, msg = _post_call(ints, {}, dict(arrays, idxday=np.array(bad, dtype=np.int32)))
        assert rc == 1 and b"idxday" in msg, bad
    rc, msg = _post_call(dict(ints, doLWrad=0), {}, dict(arrays, htrsw=None))
    assert rc == 1 and b"htrsw is NULL" in msg
    rc, _ = _post_call(dict(ints, doLWrad=0, doSWrad=0), {}, {})
    assert rc == 0


def test_cloud_overlap_host_entry_argument_errors():
    """ccpprad_cloud_overlap rejects bad dimensions and does nothing without
    radiation, on the host."""
    from ccpprad import lib, capi
    L = lib.load()
    for ints, want in ((dict(nCol=-1, nLev=4, doLWrad=1), 1), (dict(nCol=4, nLev=0, doLWrad=1), 1),
                       (dict(nCol=4, nLev=4, doLWrad=0, doSWrad=0), 0), (dict(nCol=0, nLev=4, doLWrad=1), 0)):
        a = capi.overlap_struct(ints, {}, {})
        buf, n = capi.errbuf()
        assert L.ccpprad_cloud_overlap(C.addressof(a.struct), buf, n) == want, ints
