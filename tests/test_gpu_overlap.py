"""GPU parity of the cloud-overlap scheme (GFS_rrtmgp_cloud_overlap_run) and of
GFS_rrtmgp_pre's layer geometry against the FP64 oracle, through the C-ABI.
Tolerances: the decorrelation lengths (formed without FMA, as the reference
rounds them) and the overlap parameters to 1e-15 relative (the device's exp
against the host's libm).  The layer thicknesses are differences of
logarithms, ln p_bot - ln p_top, so a last-bit difference between the two
log implementations grows by |ln p| / |d ln p|: deltaZ is held to 8 ulp
times that factor of the case (1.3e-12 here) relative.  The centre position
pfac = |ln p_bot - ln p_lay| / |d ln p| carries the same absolute error times
the layer's thickness, while deltaZc itself can be ~1 m where two centres sit
near the boundary between them: deltaZc's error is held to the same factor
times the two thicknesses it spans (the oracle's own error against an 80-bit
evaluation is 1.1e-13 of them).  deltaP is exact, tv 1e-15.  The zeros of
the exponential-random rule and the unwritten outputs are exact."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(d):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}


@pytest.mark.parametrize("r8", [False, True])
def test_cloud_overlap_matches_oracle(r8):
    from ccpprad import lib
    from oracle import oracle as O
    from test_oracle_overlap import CASES, CNV, CON_PI, flags, outputs, overlap_case
    ncol, nlev = 1000, 64
    inp = overlap_case(ncol, nlev, seed=9)
    dinp = _dev(inp)
    for idcor, iovr, (cnv, iovr_c), jul in CASES:
        print(f"overlap idcor={idcor} iovr={iovr} cnv={cnv} iovr_convcld={iovr_c} julian={jul} r8={r8}", flush=True)
        ints = flags(ncol, nlev, idcor, iovr, iovr_c, cnv, r8)
        reals = dict(julian=jul, con_pi=CON_PI, dcorr_con=2.5)
        ref = outputs(ncol, nlev)
        O.cloud_overlap(ints, reals, {**inp, **ref})
        dout = _dev(outputs(ncol, nlev))
        lib.cloud_overlap_dev(ints, reals, {**dinp, **dout})
        torch.cuda.synchronize()
        for k, r in ref.items():
            g = dout[k].cpu().numpy()
            if k == "cnv_cloud_overlap_param" and cnv not in CNV.values():
                assert np.all(np.isnan(g)), "written without a deep-convection scheme"
                continue
            assert np.array_equal(g == 0.0, r == 0.0), k
            assert np.allclose(g, r, rtol=1e-15, atol=0), (k, np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-300)))
        np.testing.assert_array_equal(dout["precip_overlap_param"].cpu().numpy(),
                                      dout["cloud_overlap_param"].cpu().numpy())


def test_cloud_overlap_no_radiation_writes_nothing():
    from ccpprad import lib
    from test_oracle_overlap import CON_PI, flags, outputs, overlap_case
    ncol, nlev = 100, 8
    ints = flags(ncol, nlev, 1, 5, 4, 2, True)
    ints["doSWrad"] = ints["doLWrad"] = 0
    dout = _dev(outputs(ncol, nlev))
    lib.cloud_overlap_dev(ints, dict(con_pi=CON_PI), {**_dev(overlap_case(ncol, nlev)), **dout})
    torch.cuda.synchronize()
    assert all(torch.isnan(v).all().item() for v in dout.values())


def test_cloud_overlap_argument_errors():
    from ccpprad import lib
    from test_oracle_overlap import CON_PI, flags, outputs, overlap_case
    ncol, nlev = 10, 4
    inp = _dev(overlap_case(ncol, nlev))
    out = _dev(outputs(ncol, nlev))
    ints = flags(ncol, nlev, 1, 5, 4, 2, True)
    for drop, msg in (("lat", "lat"), ("deltaZc", "deltaZc"), ("cld_frac", "cld_frac"),
                      ("cnv_cloud_overlap_param", "cnv_cloud_overlap_param")):
        a = {**inp, **out}
        del a[drop]
        with pytest.raises(RuntimeError, match=msg):
            lib.cloud_overlap_dev(ints, dict(con_pi=CON_PI), a)
    with pytest.raises(RuntimeError, match="dimensions"):
        lib.cloud_overlap_dev({**ints, "nLev": 0}, dict(con_pi=CON_PI), {**inp, **out})


@pytest.mark.parametrize("top", [False, True])
def test_layer_geometry_matches_oracle(top):
    from ccpprad import lib
    from oracle import oracle as O
    from test_oracle_overlap import geometry_case
    ncol, nlev = 777, 64
    p_lev, p_lay, t_lay, q_lay = geometry_case(ncol, nlev, seed=4, top=top)
    tv, dz, dzc, dp = O.layer_geometry(p_lev, p_lay, t_lay, q_lay)
    d = _dev(dict(p_lev=p_lev, p_lay=p_lay, t_lay=t_lay, q_lay=q_lay))
    out = {k: torch.full((nlev, ncol), -1.0, dtype=torch.float64, device="cuda") for k in ("tv", "dz", "dzc", "dp")}
    lib.layer_geometry_dev(ncol, nlev, d["p_lev"], d["p_lay"], d["t_lay"], d["q_lay"], out["tv"], out["dz"],
                           out["dzc"], out["dp"])
    torch.cuda.synchronize()
    assert np.array_equal(out["dp"].cpu().numpy(), dp)
    assert np.allclose(out["tv"].cpu().numpy(), tv, rtol=1e-15, atol=0)
    lp = np.log(p_lev)
    rtol = 8 * np.finfo(np.float64).eps * np.max(np.abs(lp)) / np.min(np.abs(np.diff(lp, axis=0)))
    below = (lambda a: np.vstack([np.zeros((1, ncol)), a[:-1]])) if not top else \
            (lambda a: np.vstack([a[1:], np.zeros((1, ncol))]))
    for k, r, scale in (("dz", dz, np.abs(dz)), ("dzc", dzc, dz + below(dz))):
        g = out[k].cpu().numpy()
        err = np.max(np.abs(g - r) / scale)
        print(f"{k}: max err {err:.3g} of its scale, tolerance {rtol:.3g}")
        assert err <= rtol, k
    # optional outputs may be NULL
    dzc2 = torch.zeros_like(out["dzc"])
    lib.layer_geometry_dev(ncol, nlev, d["p_lev"], d["p_lay"], d["t_lay"], d["q_lay"], None, None, dzc2, None)
    torch.cuda.synchronize()
    assert torch.equal(dzc2, out["dzc"])
