"""ctypes mirror of include/ccpprad.h and helpers to marshal tables/args.

The structs carry plain pointers; the Python objects that own the memory
(numpy arrays or torch tensors) are kept alive on the returned wrapper.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

from . import synthetic as syn

P_I = C.POINTER(C.c_int)
P_D = C.POINTER(C.c_double)


class Minor(C.Structure):
    _fields_ = [("n", C.c_int), ("ncontrib", C.c_int), ("limits_gpt", P_I),
                ("scales_with_density", P_I), ("idx_scaling", P_I),
                ("scale_by_complement", P_I), ("idx_minor", P_I),
                ("kminor_start", P_I), ("kminor", P_D)]


class KDistC(C.Structure):
    _fields_ = [("is_lw", C.c_int), ("ngas", C.c_int), ("nband", C.c_int), ("ngpt", C.c_int),
                ("ntemp", C.c_int), ("npres", C.c_int), ("neta", C.c_int), ("nflav", C.c_int),
                ("nplanck", C.c_int), ("band_lims_gpt", P_I), ("band_lims_wvn", P_D),
                ("flavor", P_I), ("gpoint_flavor", P_I), ("press_ref", P_D), ("temp_ref", P_D),
                ("press_ref_trop", C.c_double), ("vmr_ref", P_D), ("kmajor", P_D),
                ("minor_lower", Minor), ("minor_upper", Minor), ("planck_frac", P_D),
                ("totplnk", P_D), ("optimal_angle_fit", P_D), ("krayl_lower", P_D),
                ("krayl_upper", P_D), ("solar_src", P_D)]


PADE_FIELDS = ("pade_extliq", "pade_ssaliq", "pade_asyliq", "pade_extice", "pade_ssaice", "pade_asyice",
               "pade_sizreg_extliq", "pade_sizreg_ssaliq", "pade_sizreg_asyliq", "pade_sizreg_extice",
               "pade_sizreg_ssaice", "pade_sizreg_asyice")


class CloudLUTC(C.Structure):
    _fields_ = [("nband", C.c_int), ("nsize_liq", C.c_int), ("nsize_ice", C.c_int),
                ("nrghice", C.c_int), ("icergh", C.c_int), ("radliq_lwr", C.c_double),
                ("radliq_upr", C.c_double), ("radice_lwr", C.c_double),
                ("radice_upr", C.c_double), ("extliq", P_D), ("ssaliq", P_D), ("asyliq", P_D),
                ("extice", P_D), ("ssaice", P_D), ("asyice", P_D), ("pade", C.c_int)]
    _fields_ += [(n, P_D) for n in PADE_FIELDS]


_LW_FIELDS = (
    [(n, C.c_int) for n in ("doLWrad", "doLWclrsky", "top_at_1", "doGP_lwscat",
                            "use_LW_jacobian", "doGP_sgs_cnv", "doGP_sgs_pbl", "nCol", "nLay",
                            "nGases", "rrtmgp_phys_blksz", "nGauss_angles")]
    + [("icseed_lw", P_I)]
    + [(n, C.c_int) for n in ("iovr", "iovr_convcld", "iovr_max", "iovr_maxrand", "iovr_rand",
                              "iovr_dcorr", "iovr_exp", "iovr_exprand", "isubc_lw")]
    + [(n, P_D) for n in (
        "semis", "tsfg", "p_lay", "p_lev", "t_lay", "t_lev", "vmr_o2", "vmr_h2o", "vmr_o3",
        "vmr_ch4", "vmr_n2o", "vmr_co2", "cld_frac", "cld_lwp", "cld_reliq", "cld_iwp",
        "cld_reice", "cld_swp", "cld_resnow", "cld_rwp", "cld_rerain", "precip_frac",
        "cld_cnv_lwp", "cld_cnv_reliq", "cld_cnv_iwp", "cld_cnv_reice", "cld_pbl_lwp",
        "cld_pbl_reliq", "cld_pbl_iwp", "cld_pbl_reice", "cloud_overlap_param", "aerlw_tau",
        "aerlw_ssa", "aerlw_g", "fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky",
        "fluxlwDOWN_clrsky", "fluxlwUP_jac", "fluxlwUP_radtime", "fluxlwDOWN_radtime",
        "htrlw", "htrlwc")])

_SW_FIELDS = (
    [(n, C.c_int) for n in ("doSWrad", "doSWclrsky", "top_at_1", "doGP_sgs_cnv", "doGP_sgs_pbl",
                            "nCol", "nDay", "nLay", "nGases", "rrtmgp_phys_blksz")]
    + [("idx", P_I), ("icseed_sw", P_I)]
    + [(n, C.c_int) for n in ("iovr", "iovr_convcld", "iovr_max", "iovr_maxrand", "iovr_rand",
                              "iovr_dcorr", "iovr_exp", "iovr_exprand", "isubc_sw", "iSFC")]
    + [(n, P_D) for n in (
        "sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir", "sfc_alb_uvvis_dif", "coszen",
        "p_lay", "p_lev", "t_lay", "t_lev", "vmr_o2", "vmr_h2o", "vmr_o3", "vmr_ch4", "vmr_n2o",
        "vmr_co2", "cld_frac", "cld_lwp", "cld_reliq", "cld_iwp", "cld_reice", "cld_swp",
        "cld_resnow", "cld_rwp", "cld_rerain", "precip_frac", "cld_cnv_lwp", "cld_cnv_reliq",
        "cld_cnv_iwp", "cld_cnv_reice", "cld_pbl_lwp", "cld_pbl_reliq", "cld_pbl_iwp",
        "cld_pbl_reice", "cloud_overlap_param", "aersw_tau", "aersw_ssa", "aersw_g")]
    + [("solcon", C.c_double)]
    + [(n, P_D) for n in ("scmpsw", "fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky",
                          "fluxswDOWN_clrsky", "cldtausw", "htrsw", "htrswc")])


_POST_FIELDS = (
    [(n, C.c_int) for n in ("nCol", "nLev", "nDay", "iSFC", "iTOA", "doLWrad", "doSWrad", "save_diag",
                            "naerodp", "nfxr")]
    + [(n, C.c_double) for n in ("fhlwr", "fhswr", "raddt")]
    + [(n, P_D) for n in (
        "fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
        "fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky", "scmpsw",
        "sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir", "sfc_alb_uvvis_dif",
        "tsfa", "coszen", "coszdg", "htrlw", "p_lev", "cldsa", "cldtaulw", "cldtausw", "aerodp")]
    + [("mtopa", P_I), ("mbota", P_I)]
    + [(n, P_D) for n in (
        "topflw", "sfcflw", "sfcdlw", "sfculw", "tsflw", "htrlwu", "topfsw", "sfcfsw",
        "nirbmdi", "nirdfdi", "visbmdi", "visdfdi", "nirbmui", "nirdfui", "visbmui", "visdfui",
        "sfcnsw", "sfcdsw", "fluxr")])


class PostArgs(C.Structure):
    """ccpprad_post_args (GFS_rrtmgp_post_run's bookkeeping)."""
    _fields_ = _POST_FIELDS


class PostHrArgs(C.Structure):
    """ccpprad_post_hr_args (GFS_rrtmgp_post_run's heating rates)."""
    _fields_ = [("do_lw_clrsky_hr", C.c_int), ("do_sw_clrsky_hr", C.c_int), ("idxday", P_I),
                ("htrlw", P_D), ("htrlwc", P_D), ("htrsw", P_D), ("htrswc", P_D)]


def post_hr_struct(ints: dict, arrays: dict) -> Holder:
    """Fill a PostHrArgs from the two flags and arrays (idxday int32, the
    heating rates FP64; missing names become NULL)."""
    return _fill(PostHrArgs, ints, {}, arrays)


_OVL_FIELDS = (
    [(n, C.c_int) for n in ("nCol", "nLev", "yearlen", "doSWrad", "doLWrad", "literal_r8",
                            "idcor", "idcor_con", "idcor_hogan", "idcor_oreopoulos",
                            "iovr", "iovr_convcld", "iovr_dcorr", "iovr_exp", "iovr_exprand",
                            "imfdeepcnv", "imfdeepcnv_gf", "imfdeepcnv_samf")]
    + [(n, C.c_double) for n in ("julian", "con_pi", "dcorr_con")]
    + [(n, P_D) for n in ("lat", "deltaZc", "cld_frac", "cld_cnv_frac", "de_lgth", "cloud_overlap_param",
                          "cnv_cloud_overlap_param", "precip_overlap_param")])


class OverlapArgs(C.Structure):
    """ccpprad_cloud_overlap_args (GFS_rrtmgp_cloud_overlap_run)."""
    _fields_ = _OVL_FIELDS


def post_struct(ints: dict, reals: dict, arrays: dict) -> Holder:
    """Fill a PostArgs from ints, reals (fhlwr, fhswr, raddt) and arrays
    (numpy or torch; missing names become NULL)."""
    return _fill(PostArgs, ints, reals, arrays)


def overlap_struct(ints: dict, reals: dict, arrays: dict) -> Holder:
    """Fill an OverlapArgs from ints, reals (julian, con_pi, dcorr_con) and
    arrays (numpy or torch; missing names become NULL)."""
    return _fill(OverlapArgs, ints, reals, arrays)


def _fill(cls, ints: dict, reals: dict, arrays: dict) -> Holder:
    s = cls()
    keep = []
    for name, typ in cls._fields_:
        if typ is C.c_int:
            setattr(s, name, int(ints.get(name, 0)))
        elif typ is C.c_double:
            setattr(s, name, float(reals.get(name, 0.0)))
        else:
            a = arrays.get(name)
            if a is not None:
                keep.append(a)
            setattr(s, name, _ptr(a, typ))
    return Holder(s, keep)


class LWArgs(C.Structure):
    _fields_ = _LW_FIELDS


class SWArgs(C.Structure):
    _fields_ = _SW_FIELDS


LW_INPUT_FIELDS = [n for n, t in _LW_FIELDS if t is P_D and not n.startswith(("flux", "htr"))]
LW_OUTPUT_FIELDS = [n for n, t in _LW_FIELDS if t is P_D and n.startswith(("flux", "htr"))]
SW_INPUT_FIELDS = [n for n, t in _SW_FIELDS if t is P_D and n not in (
    "scmpsw", "fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky",
    "cldtausw", "htrsw", "htrswc")]
SW_OUTPUT_FIELDS = ["scmpsw", "fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky",
                    "fluxswDOWN_clrsky", "cldtausw", "htrsw", "htrswc"]

# CCPP overlap-method flags (GFS namelist convention: iovr 1 = random,
# 2 = maximum-random, 3 = maximum, 4 = decorrelation, 5 = exponential,
# 6 = exponential-random).
IOVR = dict(iovr_rand=1, iovr_maxrand=2, iovr_max=3, iovr_dcorr=4, iovr_exp=5, iovr_exprand=6)


def _ptr(a, ctype):
    if a is None:
        return C.cast(None, ctype)
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctype)
    # torch tensor (device or host)
    return C.cast(C.c_void_p(a.data_ptr()), ctype)


class Holder:
    """Keeps referenced buffers alive alongside a ctypes struct."""

    def __init__(self, struct, keep):
        self.struct = struct
        self.keep = keep

    def ref(self):
        return C.byref(self.struct)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def minor_struct(m, keep):
    arrs = {k: _i32(m[k]) for k in ("limits_gpt", "scales_with_density", "idx_scaling",
                                    "scale_by_complement", "idx_minor", "kminor_start")}
    arrs["kminor"] = _f64(m["kminor"])
    keep.extend(arrs.values())
    return Minor(m["n"], m["ncontrib"], _ptr(arrs["limits_gpt"], P_I),
                 _ptr(arrs["scales_with_density"], P_I), _ptr(arrs["idx_scaling"], P_I),
                 _ptr(arrs["scale_by_complement"], P_I), _ptr(arrs["idx_minor"], P_I),
                 _ptr(arrs["kminor_start"], P_I), _ptr(arrs["kminor"], P_D))


def kdist_struct(kd: syn.KDist) -> Holder:
    keep = []

    def d(a):
        if a is None:
            return C.cast(None, P_D)
        a = _f64(a)
        keep.append(a)
        return _ptr(a, P_D)

    def i(a):
        a = _i32(a)
        keep.append(a)
        return _ptr(a, P_I)

    nplanck = 0 if kd.totplnk is None else kd.totplnk.shape[1]
    s = KDistC(int(kd.is_lw), kd.ngas, kd.nband, kd.ngpt, kd.ntemp, kd.npres, kd.neta, kd.nflav,
               nplanck, i(kd.band_lims_gpt), d(kd.band_lims_wvn), i(kd.flavor),
               i(kd.gpoint_flavor), d(kd.press_ref), d(kd.temp_ref), kd.press_ref_trop,
               d(kd.vmr_ref), d(kd.kmajor), minor_struct(kd.minor_lower, keep),
               minor_struct(kd.minor_upper, keep), d(kd.planck_frac), d(kd.totplnk),
               d(kd.optimal_angle_fit), d(kd.krayl_lower), d(kd.krayl_upper), d(kd.solar_src))
    return Holder(s, keep)


def lut_struct(lut: syn.CloudLUT) -> Holder:
    keep = []

    def d(a):
        a = _f64(a)
        keep.append(a)
        return _ptr(a, P_D)

    pade = getattr(lut, "pade", None)
    s = CloudLUTC(lut.nband, lut.nsize_liq, lut.nsize_ice, lut.nrghice, lut.icergh,
                  lut.radliq_lwr, lut.radliq_upr, lut.radice_lwr, lut.radice_upr,
                  *(d(getattr(lut, k)) if getattr(lut, k) is not None else None
                    for k in ("extliq", "ssaliq", "asyliq", "extice", "ssaice", "asyice")),
                  int(pade is not None), *(d(pade[k]) if pade is not None else None for k in PADE_FIELDS))
    return Holder(s, keep)


def lw_default_ints(ncol, nlay, blksz=None, nangle=3, clrsky=False, jacobian=False,
                    iovr="iovr_maxrand", isubc=1, scat=False, top_at_1=False):
    d = dict(doLWrad=1, doLWclrsky=int(clrsky), top_at_1=int(top_at_1), doGP_lwscat=int(scat),
             use_LW_jacobian=int(jacobian), doGP_sgs_cnv=0, doGP_sgs_pbl=0, nCol=ncol,
             nLay=nlay, nGases=6, rrtmgp_phys_blksz=blksz or ncol, nGauss_angles=nangle,
             iovr=IOVR[iovr], iovr_convcld=IOVR["iovr_maxrand"], isubc_lw=isubc)
    d.update(IOVR)
    return d


def sw_default_ints(ncol, nday, nlay, blksz=None, clrsky=False, iovr="iovr_maxrand", isubc=1, top_at_1=False):
    # iSFC: the surface level, nLay + 1 when the vertical index runs from the
    # top (as GFS_rrtmgp_pre sets it)
    d = dict(doSWrad=1, doSWclrsky=int(clrsky), top_at_1=int(top_at_1), doGP_sgs_cnv=0, doGP_sgs_pbl=0,
             nCol=ncol, nDay=nday, nLay=nlay, nGases=6, rrtmgp_phys_blksz=blksz or nday,
             iovr=IOVR[iovr], iovr_convcld=IOVR["iovr_maxrand"], isubc_sw=isubc,
             iSFC=nlay + 1 if top_at_1 else 1)
    d.update(IOVR)
    return d


def profile_fields(pr: syn.Profiles) -> dict:
    """Map a Profiles object onto the scheme argument names (numpy, host)."""
    nl, nc = pr.nlay, pr.ncol
    z = np.zeros((nl, nc))
    f = dict(
        p_lay=pr.p_lay, p_lev=pr.p_lev, t_lay=pr.t_lay, t_lev=pr.t_lev,
        vmr_o2=pr.vmr["o2"], vmr_h2o=pr.vmr["h2o"], vmr_o3=pr.vmr["o3"], vmr_ch4=pr.vmr["ch4"],
        vmr_n2o=pr.vmr["n2o"], vmr_co2=pr.vmr["co2"], cld_frac=pr.cld_frac, cld_lwp=pr.cld_lwp,
        cld_reliq=pr.cld_reliq, cld_iwp=pr.cld_iwp, cld_reice=pr.cld_reice, cld_swp=z,
        cld_resnow=np.full((nl, nc), 20.0), cld_rwp=z, cld_rerain=np.full((nl, nc), 100.0),
        precip_frac=z, cld_cnv_lwp=z, cld_cnv_reliq=np.full((nl, nc), 10.0), cld_cnv_iwp=z,
        cld_cnv_reice=np.full((nl, nc), 50.0), cld_pbl_lwp=z,
        cld_pbl_reliq=np.full((nl, nc), 10.0), cld_pbl_iwp=z,
        cld_pbl_reice=np.full((nl, nc), 50.0), cloud_overlap_param=np.full((nl, nc), 0.5),
        semis=pr.semis, tsfg=pr.tsfc, coszen=pr.coszen, sfc_alb_nir_dir=pr.alb_nir_dir,
        sfc_alb_nir_dif=pr.alb_nir_dif, sfc_alb_uvvis_dir=pr.alb_uvvis_dir,
        sfc_alb_uvvis_dif=pr.alb_uvvis_dif)
    # every field its own buffer (several start as the same zeros), so that
    # setting one field in place leaves the others alone
    return {k: np.array(v, dtype=np.float64, order="C", copy=True) for k, v in f.items()}


def build_args(cls, ints: dict, arrays: dict, solcon=None) -> Holder:
    """Fill an LWArgs / SWArgs from ints and a name->buffer mapping (numpy or
    torch, host or device; missing names become NULL)."""
    s = cls()
    keep = []
    for name, typ in cls._fields_:
        if typ is C.c_int:
            setattr(s, name, int(ints.get(name, 0)))
        elif typ is C.c_double:
            setattr(s, name, float(solcon if solcon is not None else 1361.0))
        elif typ is P_I:
            a = arrays.get(name)
            if a is not None:
                if isinstance(a, np.ndarray):
                    a = _i32(a)
                keep.append(a)
            setattr(s, name, _ptr(a, P_I))
        else:
            a = arrays.get(name)
            if a is not None:
                keep.append(a)
            setattr(s, name, _ptr(a, P_D))
    return Holder(s, keep)


_ERRLEN = 512


def errbuf():
    return C.create_string_buffer(_ERRLEN), _ERRLEN


class CcppradError(RuntimeError):
    pass


def check(rc, buf):
    if rc != 0:
        raise CcppradError(buf.value.decode(errors="replace") or f"status {rc}")
