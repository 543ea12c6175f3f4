"""Deterministic synthetic k-distributions, cloud lookup tables and profiles.

The real RRTMGP data files (rrtmgp-data-lw-g128-210809.nc,
rrtmgp-data-sw-g112-210809.nc, rrtmgp-cloud-optics-coeffs-{lw,sw}.nc) are
not available, so every table here is synthetic but carries the same array
roles and index conventions as the ones the CCPP loader hands to
``ty_gas_optics_rrtmgp%load`` (physics/Radiation/RRTMGP/rrtmgp_lw_gas_optics.F90:24-74,
:193-254, :439-448; rrtmgp_sw_gas_optics.F90:488-496) and to
``ty_cloud_optics%load`` (rrtmgp_lw_cloud_optics.F90:184-191, :362-367).

Layout convention: every numpy array is C-contiguous with its shape equal to
the REVERSED Fortran shape, so the raw memory is exactly the Fortran
column-major array.  For example ``kmajor`` is Fortran
``kmajor(ngpt, neta, npres+1, ntemp)`` and numpy shape
``(ntemp, npres+1, neta, ngpt)``; a CCPP field ``p_lay(ncol, nlay)`` has numpy
shape ``(nlay, ncol)``.  Integer index tables are 1-based, as in Fortran.

Synthetic k values are smooth functions of (eta, pressure row, temperature)
whose per-g-point level is drawn log-uniform in [1e-8, 1e2] in units of
1e-24 cm^2/molecule, so layer optical depths span roughly 1e-9 .. 1e2.
"""
from __future__ import annotations

import dataclasses
import numpy as np

# Gas order of the reference namelist active_gases = 'h2o_co2_o3_n2o_ch4_o2'
# (docs/pdftxt/RE210/SCM_GFS_v17_p8_input.nml:105-115).  Index 0 is dry air.
GAS_NAMES = ("h2o", "co2", "o3", "n2o", "ch4", "o2")
IH2O, ICO2, IO3, IN2O, ICH4, IO2 = 1, 2, 3, 4, 5, 6

K_UNIT = 1.0e-24  # cm^2 / molecule

# Band limits in cm-1 of the 16 LW and 14 SW RRTMGP bands.  The SW limits
# matter to the scheme layer: rrtmgp_sw_main.F90:211-213,368-387 maps the
# NIR/UV-VIS albedos by comparing the lower limit with 12850/16000 cm-1 and
# picks the UV-B band by the lower limit 29000 cm-1.
LW_BAND_LIMS = np.array([
    [10, 250], [250, 500], [500, 630], [630, 700], [700, 820], [820, 980],
    [980, 1080], [1080, 1180], [1180, 1390], [1390, 1480], [1480, 1800],
    [1800, 2080], [2080, 2250], [2250, 2390], [2390, 2680], [2680, 3250]],
    dtype=np.float64)
SW_BAND_LIMS = np.array([
    [820, 2680], [2680, 3250], [3250, 4000], [4000, 4650], [4650, 5150],
    [5150, 6150], [6150, 7700], [7700, 8050], [8050, 12850], [12850, 16000],
    [16000, 22650], [22650, 29000], [29000, 38000], [38000, 50000]],
    dtype=np.float64)

# Flavors: pairs of key species (gas indices).  9 flavors.
FLAVORS = np.array([
    [IH2O, ICO2], [IH2O, IO3], [IH2O, IN2O], [IH2O, ICH4], [IO3, ICO2],
    [IN2O, ICO2], [ICH4, IN2O], [IH2O, IO2], [IO2, IO3]], dtype=np.int32)

# (lower-atmosphere flavor, upper-atmosphere flavor) per band, 1-based
LW_BAND_FLAVOR = [(1, 5), (1, 1), (1, 6), (3, 6), (1, 5), (2, 5), (2, 9), (2, 5),
                  (4, 7), (1, 2), (1, 8), (3, 6), (6, 6), (1, 6), (7, 4), (8, 9)]
SW_BAND_FLAVOR = [(1, 1), (4, 7), (1, 5), (1, 6), (1, 8), (4, 7), (1, 5), (8, 9),
                  (1, 8), (8, 9), (2, 9), (2, 5), (9, 9), (9, 5)]

# Minor-absorber intervals: (gas, first g, last g [1-based], scales_with_density,
# scaling gas (0 = none), scale_by_complement).  Every flag combination appears
# in each regime, and some g-points are covered by two intervals.
LW_MINOR_LOWER = [(IN2O, 17, 20, 1, IH2O, 0), (ICO2, 41, 44, 1, IH2O, 1),
                  (ICH4, 65, 72, 0, 0, 0), (IO2, 1, 8, 1, 0, 0),
                  (IO3, 50, 53, 0, 0, 0), (IH2O, 25, 30, 1, IH2O, 0),
                  (IH2O, 25, 32, 1, IH2O, 1), (IN2O, 89, 96, 0, 0, 0)]
LW_MINOR_UPPER = [(IO3, 49, 56, 0, 0, 0), (ICO2, 105, 110, 1, IH2O, 1),
                  (ICH4, 73, 76, 1, 0, 0), (IN2O, 18, 22, 1, IH2O, 0),
                  (IO2, 121, 128, 0, 0, 0), (IH2O, 9, 12, 1, IH2O, 1)]
SW_MINOR_LOWER = [(IN2O, 9, 12, 1, IH2O, 0), (ICO2, 17, 22, 1, IH2O, 1),
                  (ICH4, 41, 48, 0, 0, 0), (IO2, 57, 64, 1, 0, 0),
                  (IO3, 81, 84, 0, 0, 0), (IH2O, 1, 6, 1, IH2O, 0),
                  (IH2O, 1, 8, 1, IH2O, 1)]
SW_MINOR_UPPER = [(IO3, 89, 96, 0, 0, 0), (ICO2, 25, 30, 1, IH2O, 1),
                  (ICH4, 9, 12, 1, 0, 0), (IO2, 73, 80, 1, IH2O, 0),
                  (IH2O, 33, 36, 1, IH2O, 1)]


@dataclasses.dataclass
class KDist:
    """A k-distribution (the arguments of ty_gas_optics_rrtmgp%load)."""
    is_lw: bool
    ngas: int
    nband: int
    ngpt: int
    ntemp: int
    npres: int
    neta: int
    nflav: int
    band_lims_gpt: np.ndarray      # int32 (nband, 2)  <- Fortran (2, nband)
    band_lims_wvn: np.ndarray      # f64   (nband, 2)
    flavor: np.ndarray             # int32 (nflav, 2)
    gpoint_flavor: np.ndarray      # int32 (ngpt, 2)   <- Fortran (2, ngpt)
    press_ref: np.ndarray          # f64 (npres,) Pa, decreasing
    temp_ref: np.ndarray           # f64 (ntemp,) K, increasing, uniform
    press_ref_trop: float
    vmr_ref: np.ndarray            # f64 (ntemp, ngas+1, 2) <- Fortran (2, 0:ngas, ntemp)
    kmajor: np.ndarray             # f64 (ntemp, npres+1, neta, ngpt)
    minor_lower: dict
    minor_upper: dict
    # LW only
    planck_frac: np.ndarray | None = None   # same shape as kmajor
    totplnk: np.ndarray | None = None       # f64 (nband, nplanck) <- (nplanck, nband)
    optimal_angle_fit: np.ndarray | None = None  # f64 (nband, 2) <- (2, nband)
    # SW only
    krayl_lower: np.ndarray | None = None   # f64 (ntemp, neta, ngpt)
    krayl_upper: np.ndarray | None = None
    solar_src: np.ndarray | None = None     # f64 (ngpt,)

    @property
    def temp_ref_min(self):
        return float(self.temp_ref[0])

    @property
    def temp_ref_max(self):
        return float(self.temp_ref[-1])

    @property
    def press_min(self):
        return float(self.press_ref.min())

    @property
    def press_max(self):
        return float(self.press_ref.max())


def _minor_tables(rng, spec, ntemp, neta):
    n = len(spec)
    lims = np.array([[s[1], s[2]] for s in spec], dtype=np.int32)
    width = lims[:, 1] - lims[:, 0] + 1
    start = np.concatenate([[1], 1 + np.cumsum(width)[:-1]]).astype(np.int32)
    ncontrib = int(width.sum())
    # kminor(ncontrib, neta, ntemp): smooth in eta and T, level log-uniform
    lvl = rng.uniform(np.log(1e-8), np.log(1e2), ncontrib)
    bt = rng.uniform(-1.0, 1.0, ncontrib)
    ce = rng.uniform(-1.0, 1.0, ncontrib)
    xt = np.linspace(0.0, 1.0, ntemp)[:, None, None]
    xe = np.linspace(-0.5, 0.5, neta)[None, :, None]
    lnk = lvl[None, None, :] + bt[None, None, :] * xt + ce[None, None, :] * xe
    lnk = np.clip(lnk, np.log(1e-8), np.log(1e2))
    return dict(
        n=n, ncontrib=ncontrib,
        limits_gpt=lims,
        scales_with_density=np.array([s[3] for s in spec], dtype=np.int32),
        idx_scaling=np.array([s[4] for s in spec], dtype=np.int32),
        scale_by_complement=np.array([s[5] for s in spec], dtype=np.int32),
        idx_minor=np.array([s[0] for s in spec], dtype=np.int32),
        kminor_start=start,
        kminor=np.ascontiguousarray(np.exp(lnk) * K_UNIT),  # (ntemp, neta, ncontrib)
    )


def _planck_band(wvn_lo, wvn_hi, temps, nsub=64):
    """Band-integrated Planck radiance [W m-2 sr-1] by midpoint quadrature."""
    h, c, kb = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    edges = np.linspace(wvn_lo, wvn_hi, nsub + 1)
    nu = 0.5 * (edges[1:] + edges[:-1]) * 100.0  # m-1
    dnu = (edges[1] - edges[0]) * 100.0
    t = np.asarray(temps, dtype=np.float64)[:, None]
    x = h * c * nu[None, :] / (kb * t)
    b = 2.0 * h * c * c * nu[None, :] ** 3 / np.expm1(x)
    return (b * dnu).sum(axis=1)


def make_kdist(is_lw: bool, seed: int = 20261203) -> KDist:
    """Build the synthetic LW (16 bands, 128 g) or SW (14 bands, 112 g) table.

    Sizes follow the 14-temperature, 59-pressure, 9-eta, 9-flavor layout named
    for the LW k-distribution; the SW table uses the same grids.
    """
    rng = np.random.default_rng(seed + (0 if is_lw else 1000))
    nband = 16 if is_lw else 14
    ngpt = 8 * nband
    ngas, ntemp, npres, neta = len(GAS_NAMES), 14, 59, 9
    nflav = FLAVORS.shape[0]
    band_lims_gpt = np.array([[8 * b + 1, 8 * b + 8] for b in range(nband)], dtype=np.int32)
    band_lims_wvn = (LW_BAND_LIMS if is_lw else SW_BAND_LIMS).copy()
    bflav = LW_BAND_FLAVOR if is_lw else SW_BAND_FLAVOR
    gpoint_flavor = np.zeros((ngpt, 2), dtype=np.int32)
    for b in range(nband):
        gpoint_flavor[8 * b:8 * b + 8, 0] = bflav[b][0]
        gpoint_flavor[8 * b:8 * b + 8, 1] = bflav[b][1]
    press_ref = np.exp(np.linspace(np.log(109663.2), np.log(1.0), npres))
    temp_ref = np.linspace(160.0, 355.0, ntemp)
    press_ref_trop = 9948.431564193395
    # reference vmr: (2, 0:ngas, ntemp); lower / upper atmosphere values
    base_lo = np.array([1.0, 1.0e-2, 4.0e-4, 3.0e-8, 3.2e-7, 1.8e-6, 0.209])
    base_up = np.array([1.0, 4.0e-6, 4.0e-4, 5.0e-6, 2.0e-7, 1.2e-6, 0.209])
    tfac = 1.0 + 0.1 * np.linspace(-1.0, 1.0, ntemp)
    vmr_ref = np.empty((ntemp, ngas + 1, 2))
    vmr_ref[:, :, 0] = base_lo[None, :] * tfac[:, None]
    vmr_ref[:, :, 1] = base_up[None, :] * tfac[:, None]
    vmr_ref[:, 0, :] = 1.0
    # major absorption: ln k smooth in (eta, pressure row, T) per g-point
    lvl = rng.uniform(np.log(1e-8), np.log(1e2), ngpt)
    for b in range(nband):            # ascending within each band
        lvl[8 * b:8 * b + 8].sort()
    ap = rng.uniform(-2.0, 0.5, ngpt)
    bt = rng.uniform(-1.0, 1.0, ngpt)
    ce = rng.uniform(-1.5, 1.5, ngpt)
    dpe = rng.uniform(-0.5, 0.5, ngpt)
    xt = np.linspace(0.0, 1.0, ntemp)[:, None, None, None]
    xp = np.linspace(0.0, 1.0, npres + 1)[None, :, None, None]
    xe = np.linspace(-0.5, 0.5, neta)[None, None, :, None]
    lnk = (lvl + ap * xp + bt * xt + ce * xe + dpe * xp * xe)
    lnk = np.clip(lnk, np.log(1e-8), np.log(1e2))
    kmajor = np.ascontiguousarray(np.exp(lnk) * K_UNIT)
    ml = _minor_tables(rng, LW_MINOR_LOWER if is_lw else SW_MINOR_LOWER, ntemp, neta)
    mu = _minor_tables(rng, LW_MINOR_UPPER if is_lw else SW_MINOR_UPPER, ntemp, neta)
    kd = KDist(is_lw=is_lw, ngas=ngas, nband=nband, ngpt=ngpt, ntemp=ntemp,
               npres=npres, neta=neta, nflav=nflav, band_lims_gpt=band_lims_gpt,
               band_lims_wvn=band_lims_wvn, flavor=FLAVORS.copy(),
               gpoint_flavor=gpoint_flavor, press_ref=press_ref, temp_ref=temp_ref,
               press_ref_trop=press_ref_trop, vmr_ref=np.ascontiguousarray(vmr_ref),
               kmajor=kmajor, minor_lower=ml, minor_upper=mu)
    if is_lw:
        # Planck fractions: positive, smooth, summing to 1 over each band's
        # g-points at every table node (so sum over a band of lay_src = B_b(T)).
        u = rng.uniform(0.0, 1.0, ngpt)
        v = rng.uniform(-1.0, 1.0, ngpt)
        w = rng.uniform(-1.0, 1.0, ngpt)
        s = u + v * xt + w * xe + 0.2 * xp  # (ntemp, npres+1, neta, ngpt)
        s = np.broadcast_to(s, (ntemp, npres + 1, neta, ngpt))
        e = np.exp(s)
        pf = np.empty_like(e)
        for b in range(nband):
            blk = e[..., 8 * b:8 * b + 8]
            pf[..., 8 * b:8 * b + 8] = blk / blk.sum(axis=-1, keepdims=True)
        kd.planck_frac = np.ascontiguousarray(pf)
        nplanck = 196
        tpl = np.linspace(temp_ref[0], temp_ref[-1], nplanck)
        kd.totplnk = np.ascontiguousarray(np.stack(
            [_planck_band(lo, hi, tpl) for lo, hi in band_lims_wvn]))
        fit = np.empty((nband, 2))
        fit[:, 0] = rng.uniform(-0.15, 0.05, nband)
        fit[:, 1] = 1.66 + rng.uniform(-0.05, 0.05, nband)
        kd.optimal_angle_fit = fit
    else:
        # Rayleigh cross-section ~ nu^4 scaled to 6.7e-27 cm2 at 20000 cm-1,
        # weakly dependent on eta and T.
        nu_c = band_lims_wvn.mean(axis=1)
        sig_b = 6.7e-27 * (nu_c / 2.0e4) ** 4
        sig_g = np.repeat(sig_b, 8) * rng.uniform(0.9, 1.1, ngpt)
        xt3 = np.linspace(0.0, 1.0, ntemp)[:, None, None]
        xe3 = np.linspace(-0.5, 0.5, neta)[None, :, None]
        kd.krayl_lower = np.ascontiguousarray(sig_g * (1 + 0.02 * xe3) * (1 + 0.01 * xt3))
        kd.krayl_upper = np.ascontiguousarray(sig_g * (1 - 0.02 * xe3) * (1 + 0.01 * xt3))
        pb = np.array([_planck_band(lo, hi, [5778.0])[0] for lo, hi in band_lims_wvn])
        frac = rng.uniform(0.5, 1.5, (nband, 8))
        frac /= frac.sum(axis=1, keepdims=True)
        kd.solar_src = np.ascontiguousarray((pb[:, None] * frac).reshape(-1))
    return kd


@dataclasses.dataclass
class CloudLUT:
    """Lookup-table cloud optics (the arguments of ty_cloud_optics%load, LUT form)."""
    nband: int
    nsize_liq: int
    nsize_ice: int
    nrghice: int
    radliq_lwr: float
    radliq_upr: float
    radice_lwr: float
    radice_upr: float
    extliq: np.ndarray   # (nband, nsize_liq)  <- Fortran (nsize_liq, nband)
    ssaliq: np.ndarray
    asyliq: np.ndarray
    extice: np.ndarray   # (nrghice, nband, nsize_ice) <- (nsize_ice, nband, nrghice)
    ssaice: np.ndarray
    asyice: np.ndarray
    icergh: int = 2      # ice roughness category used (1-based)
    # Pade form: None, or a dict of the ccpprad_cloud_lut pade_* arrays
    # (numpy shapes reversed from Fortran: (ncoeff, 3, nband), ice
    # (nrghice, ncoeff, 3, nband); bounds (4,))
    pade: dict | None = None


def make_cloud_pade(is_lw: bool, seed: int = 20261203) -> CloudLUT:
    """Pade-form cloud optics of the same band count: per band and size
    regime, extinction ~ a re / (1 + b re + c re^2 + d re^3) (falling like
    1/re), co-albedo and asymmetry nearly constant, positive denominators.
    Size-regime bounds (b1, b2, b3, b4): liquid (2.5, 6.5, 11, 21.5), ice
    (10, 30, 70, 180); the regime rule min(floor((re - b2) / b3) + 2, 3) puts
    the regime edges at b2 and b2 + b3: liquid 6.5 and 17.5 um, ice 30 and
    100 um."""
    rng = np.random.default_rng(seed + (11 if is_lw else 1011))
    nb, nr = (16 if is_lw else 14), 3
    lut = make_cloud_lut(is_lw, seed)

    def coeffs(nband, ext_scale, coa, asy):
        e = np.zeros((6, nr, nband))
        e[1] = ext_scale * rng.uniform(0.9, 1.1, (nr, nband))        # numerator: a re
        e[3] = rng.uniform(0.0, 0.2, (nr, nband))                     # denominator: 1 + b re + c re^2 + d re^3
        e[4] = rng.uniform(0.8, 1.2, (nr, nband))
        e[5] = rng.uniform(0.0, 1e-3, (nr, nband))
        s = np.zeros((5, nr, nband))
        s[0] = coa * rng.uniform(0.5, 1.5, (nr, nband))
        s[1] = coa * rng.uniform(-0.01, 0.01, (nr, nband))
        s[3] = rng.uniform(0.0, 0.05, (nr, nband))
        a = np.zeros((5, nr, nband))
        a[0] = asy * rng.uniform(0.9, 1.0, (nr, nband))
        a[3] = rng.uniform(0.0, 0.01, (nr, nband))
        return e, s, a

    el, sl, al = coeffs(nb, 1.5, 0.3 if is_lw else 1e-4, 0.85)
    ice = [coeffs(nb, 2.5, 0.4 if is_lw else 1e-3, 0.8) for _ in range(lut.nrghice)]
    pade = dict(pade_extliq=el, pade_ssaliq=sl, pade_asyliq=al,
                pade_extice=np.stack([c[0] for c in ice]), pade_ssaice=np.stack([c[1] for c in ice]),
                pade_asyice=np.stack([c[2] for c in ice]))
    bl, bi = np.array([2.5, 6.5, 11.0, 21.5]), np.array([10.0, 30.0, 70.0, 180.0])
    for k in ("ext", "ssa", "asy"):
        pade[f"pade_sizreg_{k}liq"] = bl.copy()
        pade[f"pade_sizreg_{k}ice"] = bi.copy()
    return dataclasses.replace(lut, radliq_lwr=2.5, radliq_upr=21.5, radice_lwr=10.0, radice_upr=180.0, pade=pade)


def make_cloud_lut(is_lw: bool, seed: int = 20261203) -> CloudLUT:
    rng = np.random.default_rng(seed + (7 if is_lw else 1007))
    nband = 16 if is_lw else 14
    nl, ni, nr = 20, 18, 3
    rl = np.linspace(2.5, 21.5, nl)
    ri = np.linspace(10.0, 180.0, ni)
    bl = rng.uniform(0.8, 1.2, nband)[:, None]
    bi = rng.uniform(0.8, 1.2, nband)[:, None]
    extliq = bl * 1.5 / rl[None, :]                 # m2/g: tau = lwp * ext
    if is_lw:
        ssaliq = np.clip(0.15 + 0.35 * rng.uniform(0, 1, nband)[:, None] + 0.01 * rl[None, :], 0, 0.95)
        ssaice = np.clip(0.30 + 0.30 * rng.uniform(0, 1, nband)[:, None] + 0.001 * ri[None, :], 0, 0.95)
    else:
        ssaliq = np.clip(1.0 - 1e-5 * rl[None, :] * rng.uniform(0.5, 50, nband)[:, None], 0.5, 0.999999)
        ssaice = np.clip(1.0 - 1e-5 * ri[None, :] * rng.uniform(0.5, 50, nband)[:, None], 0.5, 0.999999)
    asyliq = np.clip(0.80 + 0.004 * rl[None, :] + 0.02 * rng.uniform(-1, 1, nband)[:, None], 0, 0.95)
    ext_i = bi * 3.0 / (0.917 * ri[None, :])
    asy_i = np.clip(0.74 + 0.0006 * ri[None, :] + 0.02 * rng.uniform(-1, 1, nband)[:, None], 0, 0.95)
    rfac = np.array([0.97, 1.0, 1.03])[:, None, None]
    extice = ext_i[None] * rfac
    ssai = np.broadcast_to(ssaice[None], (nr, nband, ni)) * np.array([1.0, 0.995, 0.99])[:, None, None]
    asyice = asy_i[None] * np.array([1.02, 1.0, 0.98])[:, None, None]
    return CloudLUT(nband=nband, nsize_liq=nl, nsize_ice=ni, nrghice=nr,
                    radliq_lwr=float(rl[0]), radliq_upr=float(rl[-1]),
                    radice_lwr=float(ri[0]), radice_upr=float(ri[-1]),
                    extliq=np.ascontiguousarray(extliq), ssaliq=np.ascontiguousarray(ssaliq),
                    asyliq=np.ascontiguousarray(asyliq), extice=np.ascontiguousarray(extice),
                    ssaice=np.ascontiguousarray(ssai), asyice=np.ascontiguousarray(np.clip(asyice, 0, 0.95)))


# ---------------------------------------------------------------------------
# Profiles
# ---------------------------------------------------------------------------
AMDW = 28.9644 / 18.0154   # con_amd / con_amw (GFS_rrtmgp_pre.F90:370)
GRAV, RDGAS = 9.80665, 287.05


def _temp_of_p(p):
    """288 K surface, 6.5 K/km lapse to a 217 K isotherm (hydrostatic, closed form)."""
    gam = 6.5e-3
    expo = RDGAS * gam / GRAV
    t = 288.0 * (p / 1.0e5) ** expo
    return np.maximum(t, 217.0)


def _z_of_p(p):
    gam = 6.5e-3
    expo = RDGAS * gam / GRAV
    zt = (288.0 - 217.0) / gam
    pt = 1.0e5 * (217.0 / 288.0) ** (1.0 / expo)
    z_trop = 288.0 / gam * (1.0 - (p / 1.0e5) ** expo)
    z_strat = zt + RDGAS * 217.0 / GRAV * np.log(pt / p)
    return np.where(p > pt, z_trop, z_strat)


def cmp_tlev(p_lay, t_lay, p_lev, tsfc, min_p):
    """Level temperatures for surface-first ordering.

    Restates physics/Radiation/radiation_tools.F90:16-84 (branch
    ``top_at_1 = .false.``).  Arrays use the (nlay, ncol) numpy convention.
    """
    nlay = p_lay.shape[0]
    la = np.log(p_lay)
    lb = np.log(p_lev[:nlay])
    lb = lb.copy()
    lb[nlay - 1] = np.log(np.maximum(min_p, p_lev[nlay - 1]))
    t_lev = np.empty((nlay + 1, p_lay.shape[1]))
    t_lev[0] = tsfc
    t_lev[1:nlay] = t_lay[:nlay - 1] + (t_lay[1:nlay] - t_lay[:nlay - 1]) * \
        (lb[1:nlay] - la[:nlay - 1]) / (la[1:nlay] - la[:nlay - 1])
    t_lev[nlay] = t_lay[nlay - 1]
    return t_lev


@dataclasses.dataclass
class Profiles:
    ncol: int
    nlay: int
    p_lay: np.ndarray   # (nlay, ncol) Pa
    t_lay: np.ndarray
    p_lev: np.ndarray   # (nlay+1, ncol)
    t_lev: np.ndarray
    tsfc: np.ndarray    # (ncol,)
    vmr: dict           # gas name -> (nlay, ncol)
    # all-sky / surface fields (zeros for clear-sky configs)
    semis: np.ndarray
    cld_frac: np.ndarray
    cld_lwp: np.ndarray
    cld_reliq: np.ndarray
    cld_iwp: np.ndarray
    cld_reice: np.ndarray
    coszen: np.ndarray
    alb_nir_dir: np.ndarray
    alb_nir_dif: np.ndarray
    alb_uvvis_dir: np.ndarray
    alb_uvvis_dif: np.ndarray

    def subset(self, cols):
        cols = np.asarray(cols)
        kw = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name == "ncol":
                kw[f.name] = len(cols)
            elif f.name == "nlay":
                kw[f.name] = v
            elif f.name == "vmr":
                kw[f.name] = {k: np.ascontiguousarray(a[:, cols]) for k, a in v.items()}
            elif v.ndim == 1:
                kw[f.name] = np.ascontiguousarray(v[cols])
            else:
                kw[f.name] = np.ascontiguousarray(v[:, cols])
        return Profiles(**kw)


def make_profiles(ncol: int, nlay: int, seed: int = 20261203, clouds: bool = True,
                  kd: KDist | None = None) -> Profiles:
    """Seeded profiles: log-p levels 1000->1 hPa (surface first, top_at_1 false),
    lapse-rate T + N(0, 3 K), exponential H2O (2 km scale height, 0.015 kg/kg),
    Gaussian O3 (8 ppmv at 10 hPa, sigma 1.0 in ln p), CO2/CH4/N2O 400/1.8/0.32
    ppmv x U(0.8, 1.2) per column, O2 0.209.  Clouds: fraction U[0,1] in layers
    with 300 <= p <= 850 hPa, LWP log-normal (median 50 g/m2, sigma_ln 1.0) with
    r_e U[5,20] um, IWP log-normal (median 20 g/m2, sigma_ln 1.0) with r_e
    U[20,80] um.  Emissivity U[0.9,1]; cos(SZA) U[0.05,1]; albedos U[0.05,0.8].
    """
    rng = np.random.default_rng(seed)
    nlev = nlay + 1
    lev1 = np.exp(np.linspace(np.log(1.0e5), np.log(100.0), nlev))
    p_lev = np.repeat(lev1[:, None], ncol, axis=1)
    p_lay = np.exp(0.5 * (np.log(p_lev[:-1]) + np.log(p_lev[1:])))
    t_lay = _temp_of_p(p_lay) + rng.normal(0.0, 3.0, (nlay, ncol))
    tsfc = 288.0 + rng.normal(0.0, 3.0, ncol)
    if kd is not None:  # GFS_rrtmgp_pre.F90:249-264 clamps
        eps = np.finfo(np.float64).eps
        t_lay = np.where(t_lay <= kd.temp_ref_min, kd.temp_ref_min + eps, t_lay)
        t_lay = np.where(t_lay >= kd.temp_ref_max, kd.temp_ref_max - eps, t_lay)
    t_lev = cmp_tlev(p_lay, t_lay, p_lev, tsfc, 1.0 if kd is None else kd.press_min)
    z = _z_of_p(p_lay)
    q = 0.015 * np.exp(-z / 2000.0)
    q = np.maximum(q, 1.0e-6)
    vmr = {
        "h2o": (q / (1.0 - q)) * AMDW,
        "o3": 8.0e-6 * np.exp(-(np.log(p_lay) - np.log(1000.0)) ** 2 / 2.0),
        "co2": np.repeat((400e-6 * rng.uniform(0.8, 1.2, ncol))[None, :], nlay, 0),
        "ch4": np.repeat((1.8e-6 * rng.uniform(0.8, 1.2, ncol))[None, :], nlay, 0),
        "n2o": np.repeat((0.32e-6 * rng.uniform(0.8, 1.2, ncol))[None, :], nlay, 0),
        "o2": np.full((nlay, ncol), 0.209),
    }
    semis = rng.uniform(0.9, 1.0, ncol)
    z2 = np.zeros((nlay, ncol))
    if clouds:
        incl = (p_lay >= 3.0e4) & (p_lay <= 8.5e4)
        cld_frac = np.where(incl, rng.uniform(0.0, 1.0, (nlay, ncol)), 0.0)
        cloudy = cld_frac > 0.0
        cld_lwp = np.where(cloudy, 50.0 * np.exp(rng.normal(0.0, 1.0, (nlay, ncol))), 0.0)
        cld_iwp = np.where(cloudy, 20.0 * np.exp(rng.normal(0.0, 1.0, (nlay, ncol))), 0.0)
        cld_reliq = np.where(cloudy, rng.uniform(5.0, 20.0, (nlay, ncol)), 10.0)
        cld_reice = np.where(cloudy, rng.uniform(20.0, 80.0, (nlay, ncol)), 50.0)
    else:
        cld_frac, cld_lwp, cld_iwp = z2, z2.copy(), z2.copy()
        cld_reliq = np.full((nlay, ncol), 10.0)
        cld_reice = np.full((nlay, ncol), 50.0)
    coszen = rng.uniform(0.05, 1.0, ncol)
    albs = [rng.uniform(0.05, 0.8, ncol) for _ in range(4)]
    c = np.ascontiguousarray
    return Profiles(ncol=ncol, nlay=nlay, p_lay=c(p_lay), t_lay=c(t_lay), p_lev=c(p_lev),
                    t_lev=c(t_lev), tsfc=c(tsfc), vmr={k: c(v) for k, v in vmr.items()},
                    semis=semis, cld_frac=c(cld_frac), cld_lwp=c(cld_lwp),
                    cld_reliq=c(cld_reliq), cld_iwp=c(cld_iwp), cld_reice=c(cld_reice),
                    coszen=coszen, alb_nir_dir=albs[0], alb_nir_dif=albs[1],
                    alb_uvvis_dir=albs[2], alb_uvvis_dif=albs[3])
