"""The Fortran scheme modules GFS_rrtmgp_cloud_overlap and GFS_rrtmgp_post
(same names and argument lists as
physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_cloud_overlap.F90:31-138
and GFS_rrtmgp_post.F90:35-389), called by a cap-like Fortran driver on host
arrays, against the FP64 oracle.

Tolerances: the overlap outputs to 1e-15 relative with exact zeros (the
device's exp against the host's libm); the heating rates to 1e-13 relative
(one FP64 division chain each, contraction may differ); the bookkeeping
copies bit for bit and the fluxr sums to 1e-15 relative.  Outputs the
reference leaves alone keep their values bit for bit."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _drv():
    from ccpprad import build
    assert os.path.exists(build.CALLER_DRIVER), "caller_driver not built"
    return build.CALLER_DRIVER


def _call(tmp_path, mode, write):
    fin, fout = tmp_path / f"{mode}_in.bin", tmp_path / f"{mode}_out.bin"
    with open(fin, "wb") as f:
        write(f)
    r = subprocess.run([_drv(), mode, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    return open(fout, "rb").read(), r.stdout


def _i32(f, *v):
    np.array(v, dtype=np.int32).tofile(f)


def _f64(f, *arrs):
    for a in arrs:
        np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def test_fortran_cloud_overlap_matches_oracle(tmp_path):
    from oracle import oracle as O
    from test_oracle_overlap import CASES, CNV, CON_PI, IDCOR, IOVR, flags, outputs, overlap_case
    ncol, nlev = 300, 48
    inp = overlap_case(ncol, nlev, seed=31)
    cases = CASES[::3] + [CASES[-1]]

    def write(f):
        _i32(f, ncol, nlev, len(cases), 365, IOVR["iovr_dcorr"], IOVR["iovr_exp"], IOVR["iovr_exprand"],
             IDCOR["idcor_con"], IDCOR["idcor_hogan"], IDCOR["idcor_oreopoulos"], CNV["imfdeepcnv_gf"],
             CNV["imfdeepcnv_samf"])
        _f64(f, np.array([CON_PI, 2.5]))
        _f64(f, inp["lat"], inp["deltaZc"], inp["cld_frac"], inp["cld_cnv_frac"])
        for idcor, iovr, (cnv, iovr_c), jul in cases:
            _i32(f, idcor, iovr, iovr_c, cnv)
            _f64(f, np.array([jul]))

    raw, _ = _call(tmp_path, "overlap", write)
    per = 4 + 8 * (ncol + 3 * ncol * nlev)
    assert len(raw) == per * len(cases)
    for i, (idcor, iovr, (cnv, iovr_c), jul) in enumerate(cases):
        rec = raw[i * per:(i + 1) * per]
        assert np.frombuffer(rec[:4], dtype=np.int32)[0] == 0
        v = np.frombuffer(rec[4:], dtype=np.float64)
        got = dict(de_lgth=v[:ncol])
        for j, k in enumerate(("cloud_overlap_param", "cnv_cloud_overlap_param", "precip_overlap_param")):
            got[k] = v[ncol + j * ncol * nlev:ncol + (j + 1) * ncol * nlev].reshape(nlev, ncol)
        # flang's default REAL is 4 bytes: the reference's literals are single precision
        ints = flags(ncol, nlev, idcor, iovr, iovr_c, cnv, False)
        ref = outputs(ncol, nlev, fill=-7.0)
        O.cloud_overlap(ints, dict(julian=jul, con_pi=CON_PI, dcorr_con=2.5), {**inp, **ref})
        for k, r in ref.items():
            g = got[k]
            assert np.array_equal(g == 0.0, r == 0.0), (i, k)
            assert np.array_equal(g == -7.0, r == -7.0), (i, k)
            assert np.allclose(g, r, rtol=1e-15, atol=0), (i, k)
        if cnv not in CNV.values():
            assert (got["cnv_cloud_overlap_param"] == -7.0).all(), "written without a deep-convection scheme"


def _post_case(nc, nl, nday, seed, dolw=1, dosw=1, lwclr=1, swclr=1, save=1):
    from test_oracle_prepare import post_inputs
    ints, reals, arr, out = post_inputs(nc=nc, nl=nl, nday=nday, seed=seed)
    rng = np.random.default_rng(seed + 100)
    # level pressures that vary by column, fluxes, 1-based distinct daylit columns
    arr["p_lev"] = np.exp(np.linspace(np.log(1.0e5), np.log(1.0), nl + 1))[:, None] * rng.uniform(0.97, 1.03, nc)
    arr["cld_frac"] = rng.uniform(0.0, 1.0, (nl, nc))
    idx = np.sort(rng.choice(nc, nday, replace=False)).astype(np.int32) + 1
    ints.update(doLWrad=dolw, doSWrad=dosw, save_diag=save)
    hr0 = {k: rng.normal(0.0, 1.0, (nl, nc)) for k in ("htrlw", "htrsw", "htrlwu", "htrlwc", "htrswc")}
    out.pop("htrlwu")
    return ints, reals, arr, out, idx, hr0, dict(lwclr=lwclr, swclr=swclr)


SCALARS = ("sfcdlw", "sfculw", "tsflw", "nirbmdi", "nirdfdi", "visbmdi", "visdfdi", "nirbmui", "nirdfui",
           "visbmui", "visdfui", "sfcnsw", "sfcdsw")
RECS = (("topflw", 2), ("sfcflw", 4), ("topfsw", 3), ("sfcfsw", 4))
FLUX = ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky", "fluxswUP_allsky",
        "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky")


def _run_post(tmp_path, case):
    ints, reals, arr, out, idx, hr0, fl = case
    nc, nl, nday = ints["nCol"], ints["nLev"], ints["nDay"]

    def write(f):
        _i32(f, nc, nl, nday, ints["iSFC"], ints["iTOA"], ints["doLWrad"], ints["doSWrad"], fl["lwclr"],
             fl["swclr"], ints["save_diag"], 6, 45)
        _f64(f, np.array([reals["fhlwr"], reals["fhswr"], reals["raddt"]]))
        _f64(f, *(arr[k] for k in FLUX), arr["p_lev"])
        _f64(f, *(arr[k] for k in ("sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir", "sfc_alb_uvvis_dif",
                                   "tsfa", "coszen", "coszdg")))
        _f64(f, arr["aerodp"], arr["cldsa"], arr["cld_frac"], arr["cldtaulw"], arr["cldtausw"], arr["scmpsw"])
        np.ascontiguousarray(arr["mtopa"], dtype=np.int32).tofile(f)
        np.ascontiguousarray(arr["mbota"], dtype=np.int32).tofile(f)
        idx.tofile(f)
        _f64(f, out["fluxr"], *(hr0[k] for k in ("htrlw", "htrsw", "htrlwu", "htrlwc", "htrswc")))
        _f64(f, *(out[k] for k, _ in RECS))
        _f64(f, *(out[k] for k in SCALARS))

    raw, stdout = _call(tmp_path, "post", write)
    errflg = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    v = np.frombuffer(raw[4:], dtype=np.float64)
    got, o = {}, 0
    for k, shape in ([("fluxr", (45, nc))] + [(k, (nl, nc)) for k in ("htrlw", "htrsw", "htrlwu", "htrlwc", "htrswc")]
                     + [(k, (nc, m)) for k, m in RECS] + [(k, (nc,)) for k in SCALARS]):
        n = int(np.prod(shape))
        got[k] = v[o:o + n].reshape(shape)
        o += n
    assert o == v.size
    return errflg, got, stdout


def _expected(case):
    from oracle import oracle as O
    ints, reals, arr, out, idx, hr0, fl = case
    nc, nl, nday = ints["nCol"], ints["nLev"], ints["nDay"]
    exp = {k: v.copy() for k, v in hr0.items()}
    c = np.ascontiguousarray
    if ints["doLWrad"]:
        exp["htrlw"] = O.heating_rate(arr["fluxlwUP_allsky"], arr["fluxlwDOWN_allsky"], arr["p_lev"])
        if fl["lwclr"]:
            exp["htrlwc"] = O.heating_rate(arr["fluxlwUP_clrsky"], arr["fluxlwDOWN_clrsky"], arr["p_lev"])
    if ints["doSWrad"]:
        cols = idx - 1
        for k, up, dn, on in (("htrsw", "fluxswUP_allsky", "fluxswDOWN_allsky", True),
                              ("htrswc", "fluxswUP_clrsky", "fluxswDOWN_clrsky", fl["swclr"])):
            if not on:
                continue
            exp[k] = np.zeros((nl, nc))
            if nday:
                exp[k][:, cols] = O.heating_rate(c(arr[up][:, cols]), c(arr[dn][:, cols]), c(arr["p_lev"][:, cols]))
    book = {k: v.copy() for k, v in out.items()}
    book["htrlwu"] = hr0["htrlwu"].copy()
    a = dict(arr)
    a["htrlw"] = exp["htrlw"]
    a.update(book)
    O.post(ints, reals, a)
    exp["htrlwu"] = book.pop("htrlwu")
    exp.update(book)
    return exp


@pytest.mark.parametrize("kind", ["day", "dark", "lw_only", "sw_only_no_diag"])
def test_fortran_post_matches_oracle(tmp_path, kind):
    nc, nl = 700, 64
    case = {
        "day": lambda: _post_case(nc, nl, 420, 11),
        "dark": lambda: _post_case(nc, nl, 0, 12),
        "lw_only": lambda: _post_case(nc, nl, 300, 13, dosw=0, lwclr=0),
        "sw_only_no_diag": lambda: _post_case(nc, nl, 500, 14, dolw=0, swclr=0, save=0),
    }[kind]()
    errflg, got, _ = _run_post(tmp_path, case)
    assert errflg == 0
    exp = _expected(case)
    ints, _, _, _, _, hr0, fl = case
    for k, e in exp.items():
        g = got[k]
        if k.startswith("htr") and not (k == "htrlwu" and ints["doLWrad"]):
            assert np.allclose(g, e, rtol=1e-13, atol=0), (k, np.abs(g - e).max())
            assert np.array_equal(g == 0.0, e == 0.0), k
        elif k in ("fluxr", "nirbmui", "nirdfui", "visbmui", "visdfui", "sfcnsw"):
            assert np.allclose(g, e, rtol=1e-15, atol=0), k
        elif k == "htrlwu":
            assert np.array_equal(g, got["htrlw"]), k      # htrlwu = htrlw as computed
        else:
            assert np.array_equal(g, e), k
    # the fields the reference leaves alone are untouched
    if not ints["doLWrad"]:
        for k in ("htrlw", "htrlwc", "htrlwu"):
            assert np.array_equal(got[k], hr0[k]), k
    if not fl["lwclr"]:
        assert np.array_equal(got["htrlwc"], hr0["htrlwc"])
    if not ints["doSWrad"] or not fl["swclr"]:
        assert np.array_equal(got["htrswc"], hr0["htrswc"])


def test_fortran_post_heating_rate_equals_mo_heating_rates(tmp_path):
    """The LW heating rate out of GFS_rrtmgp_post_run is the one the
    host-pointer heating-rate entry gives, bit for bit."""
    from ccpprad import lib
    case = _post_case(400, 64, 200, 15)
    errflg, got, _ = _run_post(tmp_path, case)
    assert errflg == 0
    ints, _, arr, _, _, _, _ = case
    hr = lib.heating_rate(400, 64, arr["fluxlwUP_allsky"], arr["fluxlwDOWN_allsky"], arr["p_lev"])
    assert np.array_equal(got["htrlw"], hr)


def test_fortran_post_rejects_bad_idxday(tmp_path):
    case = list(_post_case(100, 16, 10, 16))
    idx = case[4].copy()
    idx[3] = 101            # beyond nCol
    case[4] = idx
    errflg, got, stdout = _run_post(tmp_path, tuple(case))
    assert errflg == 1
    assert "idxday" in stdout
    hr0 = case[5]
    for k in ("htrlw", "htrsw"):
        assert np.array_equal(got[k], hr0[k]), k     # nothing written
