"""ctypes wrapper of the FP64 oracle (oracle/rrtmgp_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  Never imported by the product
package.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
_PKG = os.path.join(_REPO, "ccpp-physics_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from ccpprad import capi, synthetic as syn  # noqa: E402

P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int)
P_U8 = C.POINTER(C.c_ubyte)

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_gas_optics.argtypes = [C.c_void_p, C.c_int, C.c_int, P_D, P_D, P_D, P_D, P_D, P_D, P_D]
        L.orc_lw_sources.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P_D, P_D, P_D, P_D,
                                     P_D, P_D, P_D, P_D, P_D]
        L.orc_lw_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, P_D, P_D, P_D,
                                   P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D]
        L.orc_lw2s_gpt.argtypes = [C.c_int, C.c_int, P_D, P_D, P_D, P_D, C.c_double, C.c_double, P_D, P_D]
        L.orc_lw_solve_2str.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [P_D] * 9
        L.orc_optimal_angles.argtypes = [C.c_void_p, C.c_int, C.c_int, P_D, P_D]
        L.orc_prepare.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4 + [P_D] * 13
        L.orc_prepare.restype = C.c_int
        L.orc_post.argtypes = [C.c_void_p]
        L.orc_cloud_overlap.argtypes = [C.c_void_p]
        L.orc_layer_geometry.argtypes = [C.c_int, C.c_int] + [C.c_double] * 3 + [P_D] * 8
        L.orc_gauss.argtypes = [C.c_int, P_D, P_D]
        L.orc_heating_rate.argtypes = [C.c_int, C.c_int, P_D, P_D, P_D, P_D]
        L.orc_mt_draws.argtypes = [C.c_int, C.c_int, P_D]
        L.orc_mcica_mask.argtypes = [C.c_int, C.c_int, C.c_int, P_I, P_D, P_D] + [C.c_int] * 7 + [P_U8]
        L.orc_cloud_optics.argtypes = [C.c_void_p, C.c_int, C.c_int, P_D, P_D, P_D, P_D, P_D, P_D, P_D]
        L.orc_cloud_optics.restype = C.c_int
        L.orc_lw_main_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_sw_main_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_sw_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P_D, P_D, P_D, P_D, P_D,
                                   P_D, P_D, P_D, P_D, P_D]
        L.orc_planck_band.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_planck_band.restype = C.c_double
        L.orc_interp_layer.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, P_D, P_I, P_I,
                                       P_I, P_D, P_D, P_D, P_I, P_D, P_D]
        L.orc_nthreads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


_lib32 = None


def lib32():
    """The FP32 build of the same source (liboracle32.so: per-g-point optics
    and transport in FP32, reductions in FP64), for the FP32 CPU baseline.
    Only its scheme entries are bound; it is timed, never used as a checker."""
    global _lib32
    if _lib32 is None:
        path = os.path.join(_HERE, "liboracle32.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle32.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_lw_main_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_sw_main_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.orc_nthreads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _lib32 = L
    return _lib32


def ref_lib():
    """The reference's own Fortran (mersenne_twister.f, radiation_tools.F90)
    built by `make -C oracle ref`; None when not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libccpp_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_mt_draws.argtypes = [C.c_int, C.c_int, P_D]
        R.ref_mt_draws_chunked.argtypes = [C.c_int, C.c_int, C.c_int, P_D]
        R.ref_cmp_tlev.argtypes = [C.c_int, C.c_int, C.c_double, P_D, P_D, P_D, P_D, P_D]
        R.ref_cloud_overlap.argtypes = _REF_OVL_ARGTYPES
        _ref = R
    return _ref


# ref_cloud_overlap(ncol, nlev, yearlen, julian, con_pi, dcorr_con, idcor,
# idcor_con, idcor_hogan, idcor_oreopoulos, iovr, iovr_convcld, iovr_dcorr,
# iovr_exp, iovr_exprand, imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf, lat,
# deltaZc, cld_frac, cld_cnv_frac, de_lgth, cloud, cnv, precip)
_REF_OVL_ARGTYPES = [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int] * 12 + [P_D] * 8
_ref_r8 = None


def ref_lib_r8():
    """The reference's cloud-overlap scheme built with default-real-8
    (`make -C oracle ref`, _ref/r8/); None when not built."""
    global _ref_r8
    if _ref_r8 is None:
        path = os.path.join(_HERE, "_ref", "r8", "libccpp_ref_r8.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_cloud_overlap.argtypes = _REF_OVL_ARGTYPES
        _ref_r8 = R
    return _ref_r8


def _p(a, t=P_D):
    return None if a is None else a.ctypes.data_as(t)


def set_threads(n):
    lib().orc_set_threads(int(n))
    if _lib32 is not None:
        _lib32.orc_set_threads(int(n))


def nthreads():
    return lib().orc_nthreads()


def vmr_stack(pr: syn.Profiles):
    """(ngas, nlay, ncol) numpy == Fortran (ncol, nlay, ngas), k-dist gas order."""
    return np.ascontiguousarray(np.stack([pr.vmr[g] for g in syn.GAS_NAMES]))


def gas_optics(kd, pr):
    h = capi.kdist_struct(kd)
    nc, nl, ng = pr.ncol, pr.nlay, kd.ngpt
    tau = np.zeros((nc, nl, ng))
    pf = np.zeros((nc, nl, ng)) if kd.is_lw else None
    ray = None if kd.is_lw else np.zeros((nc, nl, ng))
    v = vmr_stack(pr)
    lib().orc_gas_optics(C.byref(h.struct), nc, nl, _p(pr.p_lay), _p(pr.p_lev), _p(pr.t_lay),
                         _p(v), _p(tau), _p(pf), _p(ray))
    return (tau, pf) if kd.is_lw else (tau, ray)


def lw_sources(kd, pr, pfrac, top_at_1=False):
    h = capi.kdist_struct(kd)
    nc, nl, ng = pr.ncol, pr.nlay, kd.ngpt
    out = [np.zeros((nc, nl, ng)) for _ in range(3)] + [np.zeros((nc, ng)) for _ in range(2)]
    lib().orc_lw_sources(C.byref(h.struct), nc, nl, int(top_at_1), _p(pr.t_lay), _p(pr.t_lev),
                         _p(np.ascontiguousarray(pr.tsfc)), _p(np.ascontiguousarray(pfrac)),
                         *[_p(o) for o in out])
    return dict(lay_src=out[0], lev_inc=out[1], lev_dec=out[2], sfc_src=out[3], sfc_jac=out[4])


def gauss(n):
    D = np.zeros(max(n, 1))
    w = np.zeros(max(n, 1))
    lib().orc_gauss(n, _p(D), _p(w))
    return D, w


def lw_solve(kd, ncol, nlay, tau, src, emis_b, nang=3, Ds=None, top_at_1=False, jac=False):
    h = capi.kdist_struct(kd)
    up = np.zeros((nlay + 1, ncol))
    dn = np.zeros((nlay + 1, ncol))
    fj = np.zeros((nlay + 1, ncol)) if jac else None
    c = np.ascontiguousarray
    lib().orc_lw_solve(C.byref(h.struct), ncol, nlay, int(top_at_1), nang,
                       _p(None if Ds is None else c(Ds)), _p(c(tau)), _p(c(src["lay_src"])),
                       _p(c(src["lev_inc"])), _p(c(src["lev_dec"])), _p(c(emis_b)),
                       _p(c(src["sfc_src"])), _p(c(src["sfc_jac"])), _p(up), _p(dn), _p(fj))
    return up, dn, fj


def lw2s_gpt(tau, ssa, g, levs, emis, sfc, top_at_1=True):
    """LW two-stream fluxes of one g-point column with given level sources
    (array order); returns up, dn at the nlay+1 levels."""
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    nl = len(tau)
    up, dn = np.zeros(nl + 1), np.zeros(nl + 1)
    lib().orc_lw2s_gpt(nl, int(top_at_1), _p(c(tau)), _p(c(ssa)), _p(c(g)), _p(c(levs)), float(emis),
                       float(sfc), _p(up), _p(dn))
    return up, dn


def lw_solve_2str(kd, ncol, nlay, tau, ssa, g, src, emis_b, top_at_1=False):
    """lw_solver_2stream + band/broadband sums; tau/ssa/g (ncol, nlay, ngpt)."""
    h = capi.kdist_struct(kd)
    up = np.zeros((nlay + 1, ncol))
    dn = np.zeros((nlay + 1, ncol))
    c = np.ascontiguousarray
    lib().orc_lw_solve_2str(C.byref(h.struct), ncol, nlay, int(top_at_1), _p(c(tau)), _p(c(ssa)), _p(c(g)),
                            _p(c(src["lev_inc"])), _p(c(src["lev_dec"])), _p(c(emis_b)),
                            _p(c(src["sfc_src"])), _p(up), _p(dn))
    return up, dn


def prepare(kd, prsi, prsl, tgrs, q_h2o, q_o3, tsfc):
    """GFS_rrtmgp_pre's input preparation; arrays (nlev[+1], ncol) numpy =
    Fortran (ncol, nlev[+1]).  Returns (dict of outputs, top_at_1)."""
    c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    nl, nc = prsl.shape
    out = {k: np.zeros((nl, nc)) for k in ("p_lay", "t_lay", "q_lay", "vmr_h2o", "vmr_o3")}
    out["p_lev"] = np.zeros((nl + 1, nc))
    out["t_lev"] = np.zeros((nl + 1, nc))
    pr = np.asarray(kd.press_ref)
    top = lib().orc_prepare(nc, nl, float(pr.min()), float(pr.max()), float(kd.temp_ref[0]),
                            float(kd.temp_ref[-1]), _p(c(prsi)), _p(c(prsl)), _p(c(tgrs)), _p(c(q_h2o)),
                            _p(c(q_o3)), _p(c(tsfc)), _p(out["p_lay"]), _p(out["p_lev"]), _p(out["t_lay"]),
                            _p(out["t_lev"]), _p(out["q_lay"]), _p(out["vmr_h2o"]), _p(out["vmr_o3"]))
    return out, bool(top)


def post(ints, reals, arrays):
    """GFS_rrtmgp_post's bookkeeping on host arrays (updated in place)."""
    h = capi.post_struct(ints, reals, arrays)
    lib().orc_post(h.ref())


def layer_geometry(p_lev, p_lay, t_lay, q_lay, con_rd=287.05, con_g=9.80665, con_fvirt=0.60779):
    """GFS_rrtmgp_pre's tv_lay, deltaZ, deltaZc, deltaP from (nlev[+1], ncol)
    C-order arrays (= Fortran (nCol, nLev[+1])); returns them the same way."""
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    nl, nc = np.shape(t_lay)
    tv, dz, dzc, dp = (np.zeros((nl, nc)) for _ in range(4))
    lib().orc_layer_geometry(nc, nl, con_rd, con_g, con_fvirt, _p(c(p_lev)), _p(c(p_lay)), _p(c(t_lay)),
                             _p(c(q_lay)), _p(tv), _p(dz), _p(dzc), _p(dp))
    return tv, dz, dzc, dp


def cloud_overlap(ints, reals, arrays):
    """GFS_rrtmgp_cloud_overlap_run on host arrays (outputs updated in place)."""
    h = capi.overlap_struct(ints, reals, arrays)
    lib().orc_cloud_overlap(h.ref())


def ref_cloud_overlap(ints, reals, arrays, r8=False):
    """The reference's own GFS_rrtmgp_cloud_overlap_run (built from its
    sources into _ref/); the outputs in arrays are updated in place.  None
    when that build is absent."""
    R = ref_lib_r8() if r8 else ref_lib()
    if R is None:
        return None
    g = lambda k: int(ints.get(k, 0))
    nc, nl = g("nCol"), g("nLev")
    zero = np.zeros((nl, nc))
    a = {k: arrays.get(k, zero) for k in ("lat", "deltaZc", "cld_frac", "cld_cnv_frac")}
    if "lat" not in arrays:
        a["lat"] = np.zeros(nc)
    R.ref_cloud_overlap(nc, nl, g("yearlen"), float(reals.get("julian", 0.0)), float(reals["con_pi"]),
                        float(reals.get("dcorr_con", 0.0)), g("idcor"), g("idcor_con"), g("idcor_hogan"),
                        g("idcor_oreopoulos"), g("iovr"), g("iovr_convcld"), g("iovr_dcorr"), g("iovr_exp"),
                        g("iovr_exprand"), g("imfdeepcnv"), g("imfdeepcnv_gf"), g("imfdeepcnv_samf"),
                        *(_p(np.ascontiguousarray(a[k])) for k in ("lat", "deltaZc", "cld_frac", "cld_cnv_frac")),
                        *(_p(arrays[k]) for k in ("de_lgth", "cloud_overlap_param", "cnv_cloud_overlap_param",
                                                  "precip_overlap_param")))
    return True


def optimal_angles(kd, ncol, nlay, tau):
    h = capi.kdist_struct(kd)
    Ds = np.zeros((ncol, kd.ngpt))
    lib().orc_optimal_angles(C.byref(h.struct), ncol, nlay, _p(np.ascontiguousarray(tau)), _p(Ds))
    return Ds


def sw_solve(kd, ncol, nlay, mu0, tau, ssa, g, toa, alb_dir, alb_dif, top_at_1=False):
    h = capi.kdist_struct(kd)
    nb, nv = kd.nband, nlay + 1
    bup, bdn, bdir = (np.zeros((ncol, nv, nb)) for _ in range(3))
    c = np.ascontiguousarray
    lib().orc_sw_solve(C.byref(h.struct), ncol, nlay, int(top_at_1), _p(c(mu0)), _p(c(tau)),
                       _p(c(ssa)), _p(c(g)), _p(c(toa)), _p(c(alb_dir)), _p(c(alb_dif)),
                       _p(bup), _p(bdn), _p(bdir))
    return bup, bdn, bdir


def heating_rate(up, dn, p_lev):
    nv, nc = up.shape
    h = np.zeros((nv - 1, nc))
    c = np.ascontiguousarray
    lib().orc_heating_rate(nc, nv - 1, _p(c(up)), _p(c(dn)), _p(c(p_lev)), _p(h))
    return h


def mt_draws(seed, n):
    out = np.zeros(n)
    lib().orc_mt_draws(int(seed), int(n), _p(out))
    return out


def mcica_mask(seeds, cf, ngpt, iovr="iovr_maxrand", overlap=None):
    nl, nc = cf.shape
    m = np.zeros((nc, ngpt, nl), dtype=np.uint8)
    io = capi.IOVR
    lib().orc_mcica_mask(nc, nl, ngpt, _p(np.ascontiguousarray(seeds, dtype=np.int32), P_I),
                         _p(np.ascontiguousarray(cf)),
                         _p(None if overlap is None else np.ascontiguousarray(overlap)),
                         io[iovr], io["iovr_max"], io["iovr_maxrand"], io["iovr_rand"],
                         io["iovr_dcorr"], io["iovr_exp"], io["iovr_exprand"], _p(m, P_U8))
    return m  # (ncol, ngpt, nlay) == Fortran (nlay, ngpt, ncol)


def cloud_optics(lut, lwp, iwp, reliq, reice):
    hl = capi.lut_struct(lut)
    nl, nc = lwp.shape
    t, s, g = (np.zeros((lut.nband, nl, nc)) for _ in range(3))
    c = np.ascontiguousarray
    rc = lib().orc_cloud_optics(C.byref(hl.struct), nc, nl, _p(c(lwp)), _p(c(iwp)), _p(c(reliq)),
                                _p(c(reice)), _p(t), _p(s), _p(g))
    if rc:
        raise ValueError("effective radius outside LUT range")
    return t, s, g


def _run(fn, kd, lut, cls, ints, arrays, out_names, solcon=None):
    hk = capi.kdist_struct(kd)
    hl = capi.lut_struct(lut)
    ha = capi.build_args(cls, ints, arrays, solcon=solcon)
    buf, n = capi.errbuf()
    rc = fn(C.byref(hk.struct), C.byref(hl.struct), C.byref(ha.struct), buf, n)
    capi.check(rc, buf)
    return {k: arrays[k] for k in out_names if k in arrays}


def lw_main_run(kd, lut, ints, arrays, fp32=False):
    """rrtmgp_lw_main_run on host numpy arrays; outputs are written in place
    into the arrays dict (which must hold the output buffers).  fp32: the
    FP32 build (baseline timing only)."""
    L = lib32() if fp32 else lib()
    return _run(L.orc_lw_main_run, kd, lut, capi.LWArgs, ints, arrays, capi.LW_OUTPUT_FIELDS)


def sw_main_run(kd, lut, ints, arrays, solcon=1361.0, fp32=False):
    L = lib32() if fp32 else lib()
    return _run(L.orc_sw_main_run, kd, lut, capi.SWArgs, ints, arrays, capi.SW_OUTPUT_FIELDS,
                solcon=solcon)


def lw_outputs(ncol, nlay, hr=True):
    nv = nlay + 1
    o = {n: np.zeros((nv, ncol)) for n in capi.LW_OUTPUT_FIELDS if n.startswith("flux")}
    if hr:
        o["htrlw"] = np.zeros((nlay, ncol))
        o["htrlwc"] = np.zeros((nlay, ncol))
    return o


def sw_outputs(ncol, nlay, hr=True):
    nv = nlay + 1
    o = {n: np.zeros((nv, ncol)) for n in ("fluxswUP_allsky", "fluxswDOWN_allsky",
                                           "fluxswUP_clrsky", "fluxswDOWN_clrsky")}
    o["scmpsw"] = np.zeros((ncol, 6))
    o["cldtausw"] = np.zeros((nlay, ncol))
    if hr:
        o["htrsw"] = np.zeros((nlay, ncol))
        o["htrswc"] = np.zeros((nlay, ncol))
    return o
