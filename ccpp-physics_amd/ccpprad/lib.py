"""ctypes binding of libccpprad_hip.so and the Python mirror of the CCPP
scheme interface (rrtmgp_lw_main / rrtmgp_sw_main init + run).

The product path is the HIP library: if it is missing, importing this module
raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

from . import capi, synthetic as syn
from .build import HIP_LIB

P_D = C.POINTER(C.c_double)
P_F = C.POINTER(C.c_float)

_lib = None
_lib_path = None

# Every symbol the C-ABI header declares (checked by the CPU test suite).
EXPORTS = ["ccpprad_lw_init", "ccpprad_sw_init", "ccpprad_lw_init_files", "ccpprad_sw_init_files",
           "ccpprad_lw_run", "ccpprad_sw_run", "ccpprad_lw_run_dev", "ccpprad_sw_run_dev",
           "ccpprad_finalize", "ccpprad_heating_rate_dev", "ccpprad_lw_gas_optics_dev",
           "ccpprad_sw_gas_optics_dev", "ccpprad_ngpt", "ccpprad_nband", "ccpprad_version",
           "ccpprad_prepare_inputs_dev", "ccpprad_post_dev", "ccpprad_layer_geometry_dev",
           "ccpprad_cloud_overlap_dev", "ccpprad_lw_init_files_pade", "ccpprad_sw_init_files_pade",
           "ccpprad_heating_rate", "ccpprad_set_launch_columns", "ccpprad_launch_columns",
           "ccpprad_lw_mcica_mask_dev", "ccpprad_sw_mcica_mask_dev", "ccpprad_lw_mcica_mask_noalloc_dev",
           "ccpprad_post", "ccpprad_cloud_overlap"]


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib, _lib_path
    if _lib is not None:
        return _lib
    path = os.environ.get("CCPPRAD_LIB", HIP_LIB)   # an alternative build of the same library
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.ccpprad_lw_init.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp), C.c_char_p, C.c_int]
    L.ccpprad_sw_init.argtypes = L.ccpprad_lw_init.argtypes
    L.ccpprad_lw_init_files.argtypes = [C.c_char_p, C.c_char_p, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(vp), C.c_char_p, C.c_int]
    L.ccpprad_sw_init_files.argtypes = L.ccpprad_lw_init_files.argtypes
    L.ccpprad_lw_init_files_pade.argtypes = L.ccpprad_lw_init_files.argtypes
    L.ccpprad_sw_init_files_pade.argtypes = L.ccpprad_lw_init_files.argtypes
    L.ccpprad_post.argtypes = [vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_cloud_overlap.argtypes = [vp, C.c_char_p, C.c_int]
    L.ccpprad_lw_run.argtypes = [vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_sw_run.argtypes = [vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_lw_run_dev.argtypes = [vp, vp, vp, C.c_int, C.c_char_p, C.c_int]
    L.ccpprad_sw_run_dev.argtypes = [vp, vp, vp, C.c_int, C.c_char_p, C.c_int]
    L.ccpprad_finalize.argtypes = [vp]
    L.ccpprad_finalize.restype = None
    L.ccpprad_heating_rate_dev.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_heating_rate.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_set_launch_columns.argtypes = [vp, C.c_int]
    L.ccpprad_launch_columns.argtypes = [vp]
    L.ccpprad_lw_mcica_mask_dev.argtypes = [vp, vp, vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_sw_mcica_mask_dev.argtypes = [vp, vp, vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_lw_mcica_mask_noalloc_dev.argtypes = [vp, vp, vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_lw_gas_optics_dev.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_sw_gas_optics_dev.argtypes = L.ccpprad_lw_gas_optics_dev.argtypes
    L.ccpprad_prepare_inputs_dev.argtypes = [vp, C.c_int, C.c_int] + [vp] * 14 + [vp, C.c_char_p, C.c_int]
    L.ccpprad_post_dev.argtypes = [vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_layer_geometry_dev.argtypes = ([C.c_int, C.c_int] + [C.c_double] * 3 + [vp] * 8 +
                                             [vp, C.c_char_p, C.c_int])
    L.ccpprad_cloud_overlap_dev.argtypes = [vp, vp, C.c_char_p, C.c_int]
    L.ccpprad_ngpt.argtypes = [vp]
    L.ccpprad_nband.argtypes = [vp]
    L.ccpprad_version.restype = C.c_char_p
    _lib = L
    _lib_path = os.path.abspath(path)
    return L


def loaded_path():
    """Path of the loaded HIP library."""
    load()
    return _lib_path


def _gas_names(names):
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    return arr


def _dp(t):
    """Raw pointer of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


class _Scheme:
    is_lw = True

    def __init__(self, kd: syn.KDist, lut: syn.CloudLUT | None, nlay: int, blksz: int = 1,
                 gas_names=syn.GAS_NAMES):
        self.L = load()
        self.kd, self.lut, self.nlay = kd, lut, nlay
        hk = capi.kdist_struct(kd)
        hl = capi.lut_struct(lut) if lut is not None else None
        names = _gas_names(list(gas_names))
        ctx = C.c_void_p()
        buf, n = capi.errbuf()
        fn = self.L.ccpprad_lw_init if self.is_lw else self.L.ccpprad_sw_init
        rc = fn(C.addressof(hk.struct), C.addressof(hl.struct) if hl else None, names, len(gas_names),
                nlay, blksz, C.byref(ctx), buf, n)
        capi.check(rc, buf)
        self.ctx = ctx

    @classmethod
    def from_files(cls, gas_file, cloud_file, nlay: int, blksz: int = 1, gas_names=syn.GAS_NAMES, pade=False,
                   nrghice=0):
        """rrtmgp_*_main_init from table files: the reference's netCDF files
        (classic formats) or this package's binary table files; pade: the
        cloud file's Pade coefficients (doGP_cldoptics_PADE); nrghice: the
        namelist's ice-roughness category (0 = the file's)."""
        self = cls.__new__(cls)
        self.L = load()
        self.kd, self.lut, self.nlay = None, None, nlay
        names = _gas_names(list(gas_names))
        ctx = C.c_void_p()
        buf, n = capi.errbuf()
        if pade:
            fn = self.L.ccpprad_lw_init_files_pade if cls.is_lw else self.L.ccpprad_sw_init_files_pade
        else:
            fn = self.L.ccpprad_lw_init_files if cls.is_lw else self.L.ccpprad_sw_init_files
        rc = fn(str(gas_file).encode(), (str(cloud_file) if cloud_file else "").encode(), int(nrghice), names,
                len(gas_names), nlay, blksz, C.byref(ctx), buf, n)
        capi.check(rc, buf)
        self.ctx = ctx
        return self

    def set_launch_columns(self, ncol):
        """Columns per launch (bounds the per-stream work buffers)."""
        if self.L.ccpprad_set_launch_columns(self.ctx, int(ncol)) != 0:
            raise capi.CcppradError("ccpprad_set_launch_columns failed")
        return self.L.ccpprad_launch_columns(self.ctx)

    @property
    def ngpt(self):
        return self.L.ccpprad_ngpt(self.ctx)

    @property
    def nband(self):
        return self.L.ccpprad_nband(self.ctx)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.ccpprad_finalize(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LWScheme(_Scheme):
    """rrtmgp_lw_main: init builds device tables; run/run_dev execute _run."""
    is_lw = True

    def run_dev(self, ints: dict, arrays: dict, stream=None, validate=True):
        """arrays: name -> torch CUDA tensor (FP64, Fortran layout).  Outputs
        are written in place."""
        h = capi.build_args(capi.LWArgs, ints, arrays)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_lw_run_dev(self.ctx, C.addressof(h.struct), stream, int(validate), buf, n)
        capi.check(rc, buf)

    def run(self, ints: dict, arrays: dict):
        """Host numpy arrays (the Fortran drop-in path)."""
        h = capi.build_args(capi.LWArgs, ints, arrays)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_lw_run(self.ctx, C.addressof(h.struct), buf, n)
        capi.check(rc, buf)

    def prepare_inputs_dev(self, ncol, nlev, prsi, prsl, tgrs, q_h2o, q_o3, tsfc, out, top_at_1=None, stream=None):
        """GFS_rrtmgp_pre's input preparation on the device; out holds the
        device arrays p_lay, p_lev, t_lay, t_lev, vmr_h2o and optionally q_lay,
        vmr_o3 (nlev = layers)."""
        L = load()
        buf, n = capi.errbuf()
        rc = L.ccpprad_prepare_inputs_dev(self.ctx, ncol, nlev, _dp(prsi), _dp(prsl), _dp(tgrs), _dp(q_h2o),
                                          _dp(q_o3), _dp(tsfc), _dp(out["p_lay"]), _dp(out["p_lev"]),
                                          _dp(out["t_lay"]), _dp(out["t_lev"]), _dp(out.get("q_lay")),
                                          _dp(out["vmr_h2o"]), _dp(out.get("vmr_o3")), _dp(top_at_1), stream,
                                          buf, n)
        capi.check(rc, buf)

    def mcica_mask_dev(self, ints: dict, arrays: dict, mask, stream=None):
        """The McICA mask of a _run call (mask: device uint32/int32 tensor of
        nCol * nLay * 4 words, [col][lay][word])."""
        h = capi.build_args(capi.LWArgs, ints, arrays)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_lw_mcica_mask_dev(self.ctx, C.addressof(h.struct), _dp(mask), stream, buf, n)
        capi.check(rc, buf)

    def mcica_mask_noalloc_dev(self, ints: dict, arrays: dict, mask, stream=None):
        """mcica_mask_dev through the allocation-free entry (its block flags
        in the library's fixed device array)."""
        h = capi.build_args(capi.LWArgs, ints, arrays)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_lw_mcica_mask_noalloc_dev(self.ctx, C.addressof(h.struct), _dp(mask), stream, buf, n)
        capi.check(rc, buf)

    def gas_optics_dev(self, ncol, nlay, p_lay, p_lev, t_lay, vmr, tau, pfrac, stream=None):
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_lw_gas_optics_dev(self.ctx, ncol, nlay, _dp(p_lay), _dp(p_lev), _dp(t_lay),
                                              _dp(vmr), _dp(tau), _dp(pfrac), stream, buf, n)
        capi.check(rc, buf)


class SWScheme(_Scheme):
    is_lw = False

    def run_dev(self, ints: dict, arrays: dict, solcon=1361.0, stream=None, validate=True):
        h = capi.build_args(capi.SWArgs, ints, arrays, solcon=solcon)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_sw_run_dev(self.ctx, C.addressof(h.struct), stream, int(validate), buf, n)
        capi.check(rc, buf)

    def run(self, ints: dict, arrays: dict, solcon=1361.0):
        h = capi.build_args(capi.SWArgs, ints, arrays, solcon=solcon)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_sw_run(self.ctx, C.addressof(h.struct), buf, n)
        capi.check(rc, buf)

    def mcica_mask_dev(self, ints: dict, arrays: dict, mask, stream=None):
        """The McICA mask of a _run call over the daylit columns (mask:
        device tensor of nDay * nLay * 4 32-bit words)."""
        h = capi.build_args(capi.SWArgs, ints, arrays)
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_sw_mcica_mask_dev(self.ctx, C.addressof(h.struct), _dp(mask), stream, buf, n)
        capi.check(rc, buf)

    def gas_optics_dev(self, ncol, nlay, p_lay, p_lev, t_lay, vmr, tau_abs, tau_ray, stream=None):
        buf, n = capi.errbuf()
        rc = self.L.ccpprad_sw_gas_optics_dev(self.ctx, ncol, nlay, _dp(p_lay), _dp(p_lev), _dp(t_lay),
                                              _dp(vmr), _dp(tau_abs), _dp(tau_ray), stream, buf, n)
        capi.check(rc, buf)


def heating_rate_dev(ncol, nlay, up, dn, p_lev, htr, stream=None):
    L = load()
    buf, n = capi.errbuf()
    rc = L.ccpprad_heating_rate_dev(ncol, nlay, _dp(up), _dp(dn), _dp(p_lev), _dp(htr), stream, buf, n)
    capi.check(rc, buf)


def heating_rate(ncol, nlay, up, dn, p_lev):
    """compute_heating_rate with host arrays (numpy FP64, Fortran (ncol,
    nlay+1) order); returns the (nlay, ncol) numpy heating rate in K/s."""
    L = load()
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    up, dn, p_lev = c(up), c(dn), c(p_lev)
    htr = np.zeros((nlay, ncol))
    buf, n = capi.errbuf()
    rc = L.ccpprad_heating_rate(ncol, nlay, _dp(up), _dp(dn), _dp(p_lev), _dp(htr), buf, n)
    capi.check(rc, buf)
    return htr


def layer_geometry_dev(ncol, nlev, p_lev, p_lay, t_lay, q_lay, tv_lay, dz, dzc, dp, con_rd=287.05,
                       con_g=9.80665, con_fvirt=0.60779, stream=None):
    """GFS_rrtmgp_pre's virtual temperature and layer thicknesses on the
    device (torch device tensors, Fortran (nCol, nLev[+1]) order; tv_lay, dz
    and dp may be None)."""
    L = load()
    buf, n = capi.errbuf()
    rc = L.ccpprad_layer_geometry_dev(ncol, nlev, con_rd, con_g, con_fvirt, _dp(p_lev), _dp(p_lay), _dp(t_lay),
                                      _dp(q_lay), _dp(tv_lay), _dp(dz), _dp(dzc), _dp(dp), stream, buf, n)
    capi.check(rc, buf)


def cloud_overlap_dev(ints, reals, arrays, stream=None):
    """GFS_rrtmgp_cloud_overlap_run on the device (arrays: torch device
    tensors by ccpprad_cloud_overlap_args field name)."""
    L = load()
    h = capi.overlap_struct(ints, reals, arrays)
    buf, n = capi.errbuf()
    rc = L.ccpprad_cloud_overlap_dev(C.addressof(h.struct), stream, buf, n)
    capi.check(rc, buf)


def post_dev(ints, reals, arrays, stream=None):
    """GFS_rrtmgp_post's flux bookkeeping on the device (arrays: torch device
    tensors by ccpprad_post_args field name)."""
    L = load()
    h = capi.post_struct(ints, reals, arrays)
    buf, n = capi.errbuf()
    rc = L.ccpprad_post_dev(C.addressof(h.struct), stream, buf, n)
    capi.check(rc, buf)
