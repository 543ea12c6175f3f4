"""Pins of the oracle's restatements of GFS_rrtmgp_cloud_overlap_run and of
GFS_rrtmgp_pre's layer geometry.

The overlap scheme is checked bit for bit against the reference's own
Fortran (GFS_rrtmgp_cloud_overlap.F90 and radiation_cloud_overlap.F90,
compiled unmodified into oracle/_ref with the compiler's default real
literals and with default-real-8), over every decorrelation-length method,
every overlap flag, both halves of the year and the convective branch; and by
closed forms that do not depend on either build.  The layer geometry (its
reference routine cannot be built here: GFS_rrtmgp_pre needs modules absent
from the tree) is pinned by the isothermal closed form, by the centre
heights of log-midpoint layers, and by the identity between the two vertical
orders."""
import itertools
import math

import numpy as np
import pytest

# the host's flag values (GFS_typedefs): overlap methods, decorrelation-length
# methods and deep-convection schemes
IOVR = dict(iovr_dcorr=3, iovr_exp=4, iovr_exprand=5)
IDCOR = dict(idcor_con=0, idcor_hogan=1, idcor_oreopoulos=2)
CNV = dict(imfdeepcnv_samf=2, imfdeepcnv_gf=3)
CON_PI = 4.0 * math.atan(1.0)


def overlap_case(ncol=23, nlev=31, seed=5):
    """Latitudes over the globe (poles included), thicknesses between layer
    centres of 20 m - 2 km, cloud fractions with runs of clear layers."""
    rng = np.random.default_rng(seed)
    lat = rng.uniform(-0.5 * CON_PI, 0.5 * CON_PI, ncol)
    lat[0], lat[1], lat[2] = 0.0, 0.5 * CON_PI, -0.5 * CON_PI
    dzc = rng.uniform(20.0, 2000.0, (nlev, ncol))
    cf = rng.uniform(0.0, 1.0, (nlev, ncol))
    cf[rng.uniform(size=cf.shape) < 0.4] = 0.0
    ccf = rng.uniform(0.0, 0.3, (nlev, ncol))
    ccf[rng.uniform(size=ccf.shape) < 0.6] = 0.0
    return dict(lat=lat, deltaZc=dzc, cld_frac=cf, cld_cnv_frac=ccf)


def outputs(ncol, nlev, fill=np.nan):
    return dict(de_lgth=np.full(ncol, fill), cloud_overlap_param=np.full((nlev, ncol), fill),
                cnv_cloud_overlap_param=np.full((nlev, ncol), fill),
                precip_overlap_param=np.full((nlev, ncol), fill))


def flags(ncol, nlev, idcor, iovr, iovr_convcld, imfdeepcnv, r8):
    return dict(nCol=ncol, nLev=nlev, yearlen=365, doSWrad=1, doLWrad=1, literal_r8=int(r8), idcor=idcor,
                iovr=iovr, iovr_convcld=iovr_convcld, imfdeepcnv=imfdeepcnv, **IOVR, **IDCOR, **CNV)


CASES = list(itertools.product([0, 1, 2, 7],        # idcor: con, hogan, oreopoulos, none
                               [1, 3, 4, 5],        # iovr: max-random, dcorr, exp, exprand
                               [(2, 5), (3, 4), (0, 3)],   # (imfdeepcnv, iovr_convcld)
                               [100.0, 250.5]))     # julian day (both halves of the year)


@pytest.mark.parametrize("r8", [False, True])
def test_overlap_matches_reference_fortran(r8):
    from oracle import oracle as O
    R = O.ref_lib_r8() if r8 else O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    ncol, nlev = 23, 31
    inp = overlap_case(ncol, nlev)
    for idcor, iovr, (cnv, iovr_c), jul in CASES:
        ints = flags(ncol, nlev, idcor, iovr, iovr_c, cnv, r8)
        reals = dict(julian=jul, con_pi=CON_PI, dcorr_con=2.5)
        mine, ref = outputs(ncol, nlev), outputs(ncol, nlev)
        O.cloud_overlap(ints, reals, {**inp, **mine})
        assert O.ref_cloud_overlap(ints, reals, {**inp, **ref}, r8=r8)
        keys = ["de_lgth", "cloud_overlap_param", "precip_overlap_param"]
        if cnv in CNV.values():
            keys.append("cnv_cloud_overlap_param")
        for k in keys:
            np.testing.assert_array_equal(mine[k], ref[k], err_msg=f"{k} idcor={idcor} iovr={iovr} cnv={cnv}")


def test_literal_precision_changes_hogan_length():
    """Hogan's fit with single- and double-precision literals: the two builds
    differ in the last digits only (so the flag selects a real difference)."""
    from oracle import oracle as O
    ncol, nlev = 23, 4
    inp = overlap_case(ncol, nlev)
    res = []
    for r8 in (0, 1):
        out = outputs(ncol, nlev)
        O.cloud_overlap(flags(ncol, nlev, 1, 3, 0, 0, r8), dict(con_pi=CON_PI), {**inp, **out})
        res.append(out["de_lgth"])
    assert not np.array_equal(res[0], res[1])
    np.testing.assert_allclose(res[0], res[1], rtol=2e-7)
    # equator and pole: 2.78 km, and the 0.6 km floor (2.78 - 4.6/2 < 0.6)
    assert res[1][0] == 2.78 and res[1][1] == 0.6 and res[1][2] == 0.6


def test_overlap_closed_forms():
    """Oreopoulos: the length peaks at am1 + am2 where the latitude in degrees
    equals am3 (eq. 11b: 28 (julian - 91) / yearlen).  alpha = e^-1 where the
    layer spacing equals the length; the constant method passes dcorr_con
    through; exponential-random zeroes exactly the clear-after-cloudy layers;
    precip = cloud; layer 1 is 0."""
    from oracle import oracle as O
    ncol, nlev, jul = 3, 6, 150.0
    am3 = 28.0 * (jul - 91.0) / 365.0
    inp = dict(lat=np.array([am3 * CON_PI / 180.0, 0.0, -0.3]), deltaZc=np.full((nlev, ncol), 1000.0),
               cld_frac=np.array([[0.5] * 3, [0.0] * 3, [0.0] * 3, [0.2] * 3, [0.3] * 3, [0.0] * 3]),
               cld_cnv_frac=np.zeros((nlev, ncol)))
    out = outputs(ncol, nlev)
    O.cloud_overlap(flags(ncol, nlev, 2, 4, 0, 0, 1), dict(julian=jul, con_pi=CON_PI), {**inp, **out})
    assert abs(out["de_lgth"][0] - (1.4315 + 2.1219)) < 1e-12
    assert np.all(out["de_lgth"][1:] < 1.4315 + 2.1219)
    out = outputs(ncol, nlev)
    O.cloud_overlap(flags(ncol, nlev, 0, 3, 0, 0, 1), dict(con_pi=CON_PI, dcorr_con=1.0), {**inp, **out})
    assert np.all(out["de_lgth"] == 1.0)
    assert np.all(out["cloud_overlap_param"][0] == 0.0)
    np.testing.assert_allclose(out["cloud_overlap_param"][1:], math.exp(-1.0), rtol=1e-15)
    np.testing.assert_array_equal(out["precip_overlap_param"], out["cloud_overlap_param"])
    assert np.all(np.isnan(out["cnv_cloud_overlap_param"]))     # no deep convection scheme: not written
    out = outputs(ncol, nlev)
    O.cloud_overlap(flags(ncol, nlev, 0, 5, 0, 0, 1), dict(con_pi=CON_PI, dcorr_con=1.0), {**inp, **out})
    zero = [k for k in range(nlev) if out["cloud_overlap_param"][k, 0] == 0.0]
    assert zero == [0, 1, 5]      # layer 1, and the clear layers that follow a cloudy one
    out = outputs(ncol, nlev)
    O.cloud_overlap(flags(ncol, nlev, 0, 1, 0, 0, 1), dict(con_pi=CON_PI, dcorr_con=1.0), {**inp, **out})
    assert np.all(out["cloud_overlap_param"] == 0.0)          # not an exponential method
    out = outputs(ncol, nlev)
    ints = flags(ncol, nlev, 0, 3, 0, 0, 1)
    ints["doSWrad"] = ints["doLWrad"] = 0
    O.cloud_overlap(ints, dict(con_pi=CON_PI, dcorr_con=1.0), {**inp, **out})
    assert all(np.all(np.isnan(v)) for v in out.values())      # no radiation this step: nothing written


def geometry_case(ncol=7, nlev=20, seed=2, top=False):
    rng = np.random.default_rng(seed)
    p_sfc = rng.uniform(9.0e4, 1.03e5, ncol)
    sig = np.linspace(1.0, 0.002, nlev + 1)
    p_lev = sig[:, None] * p_sfc[None, :]
    p_lay = np.sqrt(p_lev[:-1] * p_lev[1:]) * rng.uniform(0.97, 1.03, (nlev, ncol))
    p_lay = np.clip(p_lay, p_lev[1:] * 1.0001, p_lev[:-1] * 0.9999)
    t_lay = rng.uniform(190.0, 300.0, (nlev, ncol))
    q_lay = rng.uniform(1e-6, 0.02, (nlev, ncol))
    if top:
        p_lev, p_lay, t_lay, q_lay = p_lev[::-1], p_lay[::-1], t_lay[::-1], q_lay[::-1]
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return c(p_lev), c(p_lay), c(t_lay), c(q_lay)


def test_geometry_orders_agree():
    """The top-first form of a column is the surface-first form read
    backwards: every field equal bit for bit (heights accumulate from the
    surface in both)."""
    from oracle import oracle as O
    sfc = O.layer_geometry(*geometry_case(top=False))
    top = O.layer_geometry(*geometry_case(top=True))
    for a, b in zip(sfc, top):
        np.testing.assert_array_equal(a, b[::-1])


def test_geometry_closed_forms():
    """Isothermal dry layers with centres at the log midpoints: deltaZ =
    (R/g) T ln(p_bot / p_top); the centre sits halfway up each layer, so the
    spacing between centres is the mean of the two thicknesses, and half the
    lowest layer above the surface.  tv = t (1 + fvirt q); deltaP = |dp|."""
    from oracle import oracle as O
    ncol, nlev, T, R, g = 2, 12, 250.0, 287.05, 9.80665
    p_lev = np.outer(np.geomspace(1.0e5, 1.0e3, nlev + 1), [1.0, 0.95])
    p_lay = np.sqrt(p_lev[:-1] * p_lev[1:])
    t = np.full((nlev, ncol), T)
    tv, dz, dzc, dp = O.layer_geometry(p_lev, p_lay, t, np.zeros((nlev, ncol)), con_rd=R, con_g=g)
    ref = R / g * T * np.log(p_lev[:-1] / p_lev[1:])
    np.testing.assert_allclose(dz, ref, rtol=1e-13)
    np.testing.assert_allclose(dzc[0], 0.5 * dz[0], rtol=1e-12)
    np.testing.assert_allclose(dzc[1:], 0.5 * (dz[1:] + dz[:-1]), rtol=1e-12)
    np.testing.assert_allclose(dp, p_lev[:-1] - p_lev[1:], rtol=0)
    q = np.full((nlev, ncol), 0.01)
    tv, *_ = O.layer_geometry(p_lev, p_lay, t, q, con_fvirt=0.61)
    np.testing.assert_allclose(tv, T * (1.0 + 0.61 * 0.01), rtol=1e-15)
