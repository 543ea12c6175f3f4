"""Writes a synthetic k-distribution and cloud LUT in the layout of the
reference's netCDF table files (the variable names and dimension orders read
at rrtmgp_lw_gas_optics.F90:130-327, rrtmgp_sw_gas_optics.F90:288-376 and
rrtmgp_lw_cloud_optics.F90:130-290), as netCDF classic files.

The file form carries more than the reduced tables: its gas list has two gases
the run does not use ('co' and 'n2', at positions that shift the others), each
with minor-absorber intervals interleaved among the real ones, so reading the
file back must drop them, repack the kminor contributors and renumber the gas
indices to give the original tables again."""
import numpy as np
from scipy.io import netcdf_file

FILE_GASES = ["h2o", "co2", "o3", "n2o", "co", "ch4", "o2", "n2"]
SLOT_GAS = ["h2o", "co2", "o3", "n2o", "ch4", "o2"]      # the library's gas slots 1..6
EXTRA = ["co", "n2"]


def _chars(names, width=32):
    a = np.full((len(names), width), b" ", dtype="S1")
    for i, n in enumerate(names):
        for j, ch in enumerate(n.encode()):
            a[i, j] = bytes([ch])
    return a


def _minor_file_form(m, is_lower, rng, neta, ntemp, scaled_by_inactive=False):
    """The synthetic intervals with an extra-gas interval after the first one
    and another at the end, each with its own contributors; with
    scaled_by_inactive, also an o3 interval scaled by n2 (a gas the run does
    not use), which must stay, without its scaling."""
    n = m["n"]
    lims, kmin, start = m["limits_gpt"], m["kminor"], m["kminor_start"]
    rows = []
    for i in range(n):
        w = lims[i, 1] - lims[i, 0] + 1
        s0 = start[i] - 1
        rows.append(dict(gas=SLOT_GAS[m["idx_minor"][i] - 1], lims=lims[i].copy(),
                         dens=int(m["scales_with_density"][i]), comp=int(m["scale_by_complement"][i]),
                         scal=SLOT_GAS[m["idx_scaling"][i] - 1] if m["idx_scaling"][i] > 0 else "",
                         k=kmin[:, :, s0:s0 + w]))
        if i == 0 or i == n - 1:
            gas = EXTRA[0] if (i == 0) == is_lower else EXTRA[1]
            lo = int(lims[i, 0])
            rows.append(dict(gas=gas, lims=np.array([lo, lo + 2], dtype=np.int32), dens=1, comp=0, scal="h2o",
                             k=rng.uniform(1e-22, 1e-20, (ntemp, neta, 3))))
    if scaled_by_inactive:
        rows.append(dict(gas="o3", lims=np.array([1, 4], dtype=np.int32), dens=1, comp=0, scal="n2",
                         k=rng.uniform(1e-21, 1e-19, (ntemp, neta, 4))))
    ks, starts = 1, []
    for r in rows:
        starts.append(ks)
        ks += r["k"].shape[2]
    return rows, np.array(starts, dtype=np.int32), np.concatenate([r["k"] for r in rows], axis=2)


def write_kdist_nc(path, kd, version=2, seed=3, scaled_by_inactive=False):
    """kd: a synthetic KDist.  version: 1 (CDF-1) or 2 (64-bit offset)."""
    rng = np.random.default_rng(seed)
    f = netcdf_file(path, "w", version=version)
    ntemp, npres, neta, ngpt, nbnd = kd.ntemp, kd.npres, kd.neta, kd.ngpt, kd.nband
    nabs = len(FILE_GASES)
    lo_rows, lo_start, lo_k = _minor_file_form(kd.minor_lower, True, rng, neta, ntemp, scaled_by_inactive)
    up_rows, up_start, up_k = _minor_file_form(kd.minor_upper, False, rng, neta, ntemp)
    ident = sorted({r["gas"] for r in lo_rows + up_rows})
    dims = dict(temperature=ntemp, pressure=npres, pressure_interp=npres + 1, absorber=nabs,
                absorber_ext=nabs + 1, minor_absorber=len(ident), mixing_fraction=neta, atmos_layer=2,
                bnd=nbnd, gpt=ngpt, pair=2, string_len=32, contributors_lower=lo_k.shape[2],
                contributors_upper=up_k.shape[2], minor_absorber_intervals_lower=len(lo_rows),
                minor_absorber_intervals_upper=len(up_rows), fit_coeffs=2)
    if kd.is_lw:
        dims["temperature_Planck"] = kd.totplnk.shape[1]
    for k, v in dims.items():
        f.createDimension(k, v)

    def var(name, dtype, dn, data):
        v = f.createVariable(name, dtype, dn)
        v[...] = data

    var("gas_names", "c", ("absorber", "string_len"), _chars(FILE_GASES))
    var("gas_minor", "c", ("minor_absorber", "string_len"), _chars(ident))
    var("identifier_minor", "c", ("minor_absorber", "string_len"), _chars(ident))
    fidx = {g: FILE_GASES.index(g) + 1 for g in FILE_GASES}
    key = np.zeros((nbnd, 2, 2), dtype=np.int32)
    for b in range(nbnd):
        g0 = kd.band_lims_gpt[b, 0] - 1
        for a in range(2):
            pair = kd.flavor[kd.gpoint_flavor[g0, a] - 1]
            key[b, a] = [fidx[SLOT_GAS[pair[0] - 1]], fidx[SLOT_GAS[pair[1] - 1]]]
    var("key_species", "i", ("bnd", "atmos_layer", "pair"), key)
    var("bnd_limits_gpt", "i", ("bnd", "pair"), kd.band_lims_gpt)
    var("bnd_limits_wavenumber", "d", ("bnd", "pair"), kd.band_lims_wvn)
    var("press_ref", "d", ("pressure",), kd.press_ref)
    var("temp_ref", "d", ("temperature",), kd.temp_ref)
    var("press_ref_trop", "d", (), kd.press_ref_trop)
    var("absorption_coefficient_ref_P", "d", (), 101325.0)
    var("absorption_coefficient_ref_T", "d", (), 273.0)
    vref = rng.uniform(1e-8, 1e-6, (ntemp, nabs + 1, 2))
    vref[:, 0, :] = kd.vmr_ref[:, 0, :]
    for s, g in enumerate(SLOT_GAS):
        vref[:, fidx[g], :] = kd.vmr_ref[:, s + 1, :]
    var("vmr_ref", "d", ("temperature", "absorber_ext", "atmos_layer"), vref)
    var("kmajor", "d", ("temperature", "pressure_interp", "mixing_fraction", "gpt"), kd.kmajor)
    for sfx, rows, start, kk in (("lower", lo_rows, lo_start, lo_k), ("upper", up_rows, up_start, up_k)):
        ni = f"minor_absorber_intervals_{sfx}"
        var(f"minor_gases_{sfx}", "c", (ni, "string_len"), _chars([r["gas"] for r in rows]))
        var(f"scaling_gas_{sfx}", "c", (ni, "string_len"), _chars([r["scal"] for r in rows]))
        var(f"minor_limits_gpt_{sfx}", "i", (ni, "pair"), np.stack([r["lims"] for r in rows]))
        var(f"minor_scales_with_density_{sfx}", "i", (ni,), np.array([r["dens"] for r in rows], dtype=np.int32))
        var(f"scale_by_complement_{sfx}", "i", (ni,), np.array([r["comp"] for r in rows], dtype=np.int32))
        var(f"kminor_start_{sfx}", "i", (ni,), start)
        var(f"kminor_{sfx}", "d", ("temperature", "mixing_fraction", f"contributors_{sfx}"), kk)
    if kd.is_lw:
        var("plank_fraction", "d", ("temperature", "pressure_interp", "mixing_fraction", "gpt"), kd.planck_frac)
        var("totplnk", "d", ("bnd", "temperature_Planck"), kd.totplnk)
        var("optimal_angle_fit", "d", ("bnd", "fit_coeffs"), kd.optimal_angle_fit)
    else:
        var("rayl_lower", "d", ("temperature", "mixing_fraction", "gpt"), kd.krayl_lower)
        var("rayl_upper", "d", ("temperature", "mixing_fraction", "gpt"), kd.krayl_upper)
        var("solar_source_quiet", "d", ("gpt",), kd.solar_src)
        var("solar_source_facular", "d", ("gpt",), np.zeros(ngpt))
        var("solar_source_sunspot", "d", ("gpt",), np.zeros(ngpt))
        var("mg_default", "d", (), 0.1567652)
        var("sb_default", "d", (), 902.7126)
        var("tsi_default", "d", (), 1360.85767381726)
    f.close()


def write_lut_nc(path, lut, version=2):
    f = netcdf_file(path, "w", version=version)
    for k, v in dict(nband=lut.nband, nrghice=lut.nrghice, nsize_liq=lut.nsize_liq, nsize_ice=lut.nsize_ice,
                     pair=2).items():
        f.createDimension(k, v)

    def var(name, dn, data):
        v = f.createVariable(name, "d", dn)
        v[...] = data

    for k in ("radliq_lwr", "radliq_upr", "radice_lwr", "radice_upr"):
        var(k, (), getattr(lut, k))
    var("radliq_fac", (), (lut.radliq_upr - lut.radliq_lwr) / (lut.nsize_liq - 1))
    var("radice_fac", (), (lut.radice_upr - lut.radice_lwr) / (lut.nsize_ice - 1))
    for k in ("extliq", "ssaliq", "asyliq"):
        var("lut_" + k, ("nband", "nsize_liq"), getattr(lut, k))
    for k in ("extice", "ssaice", "asyice"):
        var("lut_" + k, ("nrghice", "nband", "nsize_ice"), getattr(lut, k))
    var("bnd_limits_wavenumber", ("nband", "pair"), np.zeros((lut.nband, 2)))
    f.close()


def to_cdf5(src, dst):
    """Re-encode a netCDF classic file (read with scipy) as CDF-5, the format
    `nccopy -k cdf5` writes: 8-byte counts, lengths and offsets, no record
    dimension, attributes dropped (the tables carry none the loader reads)."""
    import struct
    f = netcdf_file(src, "r", mmap=False)
    types = {"b": (1, 1), "c": (2, 1), "h": (3, 2), "i": (4, 4), "f": (5, 4), "d": (6, 8)}
    dims = list(f.dimensions.items())
    dim_id = {n: i for i, (n, _) in enumerate(dims)}

    def name(s):
        b = s.encode()
        return struct.pack(">q", len(b)) + b + b"\0" * (-len(b) % 4)

    vars_ = []
    for n, v in f.variables.items():
        code, size = types[v.typecode()]
        data = np.ascontiguousarray(v[...])
        if v.typecode() == "c":
            raw = data.astype("S1").tobytes()
        else:
            raw = data.astype(data.dtype.newbyteorder(">")).tobytes()
        vars_.append((n, [dim_id[d] for d in v.dimensions], code, raw))
    f.close()

    def header(begins):
        h = b"CDF\x05" + struct.pack(">q", 0)
        h += struct.pack(">iq", 0x0A, len(dims)) + b"".join(name(n) + struct.pack(">q", L) for n, L in dims)
        h += struct.pack(">iq", 0, 0)
        h += struct.pack(">iq", 0x0B, len(vars_))
        for (n, ids, code, raw), b in zip(vars_, begins):
            h += name(n) + struct.pack(">q", len(ids)) + b"".join(struct.pack(">q", i) for i in ids)
            h += struct.pack(">iq", 0, 0) + struct.pack(">iqq", code, len(raw) + (-len(raw) % 4), b)
        return h

    hlen = len(header([0] * len(vars_)))
    begins, off = [], hlen
    for _, _, _, raw in vars_:
        begins.append(off)
        off += len(raw) + (-len(raw) % 4)
    with open(dst, "wb") as out:
        out.write(header(begins))
        for _, _, _, raw in vars_:
            out.write(raw + b"\0" * (-len(raw) % 4))


def write_pade_nc(path, lut, version=2, with_lut=False):
    """Pade-form cloud optics (lut.pade) in the reference's variable names and
    dimension orders (rrtmgp_lw_cloud_optics.F90:130-150,192-205,261-288);
    with_lut: the lookup tables too, as the reference's files hold both."""
    f = netcdf_file(path, "w", version=version)
    dims = dict(nband=lut.nband, nrghice=lut.nrghice, nsizereg=3, ncoeff_ext=6, ncoeff_ssa_g=5, nbound=4, pair=2)
    if with_lut:
        dims.update(nsize_liq=lut.nsize_liq, nsize_ice=lut.nsize_ice)
    for k, v in dims.items():
        f.createDimension(k, v)

    def var(name, dn, data):
        v = f.createVariable(name, "d", dn)
        v[...] = data

    for k in ("radliq_lwr", "radliq_upr", "radice_lwr", "radice_upr"):
        var(k, (), getattr(lut, k))
    P = lut.pade
    for ph in ("liq", "ice"):
        lead = ("nrghice",) if ph == "ice" else ()
        var(f"pade_ext{ph}", lead + ("ncoeff_ext", "nsizereg", "nband"), P[f"pade_ext{ph}"])
        var(f"pade_ssa{ph}", lead + ("ncoeff_ssa_g", "nsizereg", "nband"), P[f"pade_ssa{ph}"])
        var(f"pade_asy{ph}", lead + ("ncoeff_ssa_g", "nsizereg", "nband"), P[f"pade_asy{ph}"])
        for k in ("ext", "ssa", "asy"):
            var(f"pade_sizreg_{k}{ph}", ("nbound",), P[f"pade_sizreg_{k}{ph}"])
    if with_lut:
        for k in ("extliq", "ssaliq", "asyliq"):
            var("lut_" + k, ("nband", "nsize_liq"), getattr(lut, k))
        for k in ("extice", "ssaice", "asyice"):
            var("lut_" + k, ("nrghice", "nband", "nsize_ice"), getattr(lut, k))
    var("bnd_limits_wavenumber", ("nband", "pair"), np.zeros((lut.nband, 2)))
    f.close()
