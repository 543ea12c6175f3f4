"""Writer of the package's binary table files, which ccpprad_{lw,sw}_init_files
read besides the reference's netCDF files.

Format: b"CCPPRADT", int32 version 1, int32 record count, then records of
char name[32], int32 type (1 int32, 2 float64), int32 ndim, int64 dims[4] and
the data in Fortran order (the raw memory of the numpy arrays, whose shapes
are the reversed Fortran shapes)."""
from __future__ import annotations

import struct
import numpy as np

from . import synthetic as syn


def _rec(f, name, arr):
    a = np.asarray(arr)
    if a.dtype.kind in "iub":
        a = np.ascontiguousarray(a, dtype=np.int32)
        t = 1
    else:
        a = np.ascontiguousarray(a, dtype=np.float64)
        t = 2
    dims = list(a.shape[::-1]) if a.ndim else [1]
    nd = len(dims)
    assert nd <= 4, name
    f.write(name.encode().ljust(32, b"\0"))
    f.write(struct.pack("<ii", t, nd))
    f.write(struct.pack("<4q", *(dims + [0] * (4 - nd))))
    f.write(a.tobytes(order="C"))


def _write(path, recs):
    with open(path, "wb") as f:
        f.write(b"CCPPRADT")
        f.write(struct.pack("<ii", 1, len(recs)))
        for n, a in recs:
            _rec(f, n, a)


def write_kdist(path, kd: syn.KDist):
    i = lambda v: np.array([v], dtype=np.int32)
    d = lambda v: np.array([v], dtype=np.float64)
    nplanck = 0 if kd.totplnk is None else kd.totplnk.shape[1]
    recs = [("is_lw", i(int(kd.is_lw))), ("ngas", i(kd.ngas)), ("nband", i(kd.nband)),
            ("ngpt", i(kd.ngpt)), ("ntemp", i(kd.ntemp)), ("npres", i(kd.npres)),
            ("neta", i(kd.neta)), ("nflav", i(kd.nflav)), ("nplanck", i(nplanck)),
            ("band_lims_gpt", kd.band_lims_gpt), ("band_lims_wvn", kd.band_lims_wvn),
            ("flavor", kd.flavor), ("gpoint_flavor", kd.gpoint_flavor), ("press_ref", kd.press_ref),
            ("temp_ref", kd.temp_ref), ("press_ref_trop", d(kd.press_ref_trop)),
            ("vmr_ref", kd.vmr_ref), ("kmajor", kd.kmajor)]
    for pre, m in (("minor_lower_", kd.minor_lower), ("minor_upper_", kd.minor_upper)):
        recs += [(pre + "n", i(m["n"])), (pre + "ncontrib", i(m["ncontrib"]))]
        for k in ("limits_gpt", "scales_with_density", "idx_scaling", "scale_by_complement",
                  "idx_minor", "kminor_start", "kminor"):
            recs.append((pre + k, m[k]))
    if kd.is_lw:
        recs += [("planck_frac", kd.planck_frac), ("totplnk", kd.totplnk),
                 ("optimal_angle_fit", kd.optimal_angle_fit)]
    else:
        recs += [("krayl_lower", kd.krayl_lower), ("krayl_upper", kd.krayl_upper),
                 ("solar_src", kd.solar_src)]
    _write(path, recs)


def write_lut(path, lut: syn.CloudLUT):
    i = lambda v: np.array([v], dtype=np.int32)
    d = lambda v: np.array([v], dtype=np.float64)
    recs = [("nband", i(lut.nband)), ("nsize_liq", i(lut.nsize_liq)),
            ("nsize_ice", i(lut.nsize_ice)), ("nrghice", i(lut.nrghice)),
            ("icergh", i(lut.icergh)), ("radliq_lwr", d(lut.radliq_lwr)),
            ("radliq_upr", d(lut.radliq_upr)), ("radice_lwr", d(lut.radice_lwr)),
            ("radice_upr", d(lut.radice_upr)), ("extliq", lut.extliq), ("ssaliq", lut.ssaliq),
            ("asyliq", lut.asyliq), ("extice", lut.extice), ("ssaice", lut.ssaice),
            ("asyice", lut.asyice)]
    _write(path, recs)
