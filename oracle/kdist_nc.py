"""Oracle: the k-distribution and cloud-LUT netCDF files reduced to the active
gases, restated from rte-rrtmgp's ty_gas_optics_rrtmgp%load /
init_abs_coeffs (the call sites are rrtmgp_lw_gas_optics.F90:439-448 and
rrtmgp_sw_gas_optics.F90:487-496; the rte-rrtmgp submodule is absent from the
reference tree, so the steps follow the published BSD-3 code: gases kept in
file order when active; vmr_ref columns of the kept gases; reduce_minor_arrays
dropping intervals whose gas is not active and packing their contributors;
create_key_species_reduce; create_flavor = the distinct key-species pairs over
(band, atmosphere layer), (0, 0) standing for (2, 2); create_gpoint_flavor).
Gas indices are renumbered to the library's fixed order h2o co2 o3 n2o ch4
o2.  The netCDF files are read with scipy's classic-format reader.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the library's native
loader; never imported by the product."""
import numpy as np
from scipy.io import netcdf_file

SLOTS = ["h2o", "co2", "o3", "n2o", "ch4", "o2"]


def _names(v):
    a = np.asarray(v[:])
    return ["".join(ch.decode() for ch in row).rstrip(" \x00").lower() for row in a]


def load_kdist(path, is_lw, active=tuple(SLOTS)):
    """A synthetic.KDist (the library's reduced form) from a netCDF file."""
    from ccpprad.synthetic import KDist
    f = netcdf_file(path, "r", mmap=False)
    V = f.variables
    fgas = _names(V["gas_names"])
    active = [a.lower() for a in active]
    kept = [g for g in fgas if g in active]
    slot = {g: SLOTS.index(g) + 1 for g in kept}
    key = np.asarray(V["key_species"][:])                 # (bnd, atmos_layer, pair), file indices
    nbnd = key.shape[0]
    flav, bflav = [], np.zeros((nbnd, 2), dtype=np.int32)
    for b in range(nbnd):
        for a in range(2):
            p = []
            for fi in key[b, a]:
                if fi == 0:
                    p.append(0)
                    continue
                g = fgas[fi - 1]
                if g not in active:
                    raise ValueError(f"key species {g} is not active")
                p.append(slot[g])
            if p == [0, 0]:
                p = [slot[kept[1]]] * 2
            if tuple(p) not in flav:
                flav.append(tuple(p))
            bflav[b, a] = flav.index(tuple(p)) + 1
    b2g = np.asarray(V["bnd_limits_gpt"][:])
    ngpt = int(b2g.max())
    gpf = np.zeros((ngpt, 2), dtype=np.int32)
    for b in range(nbnd):
        gpf[b2g[b, 0] - 1:b2g[b, 1], :] = bflav[b]
    vref = np.asarray(V["vmr_ref"][:])                    # (temperature, absorber_ext, atmos_layer)
    ntemp = vref.shape[0]
    vr = np.ones((ntemp, len(SLOTS) + 1, 2))
    vr[:, 0, :] = vref[:, 0, :]
    for g in kept:
        vr[:, slot[g], :] = vref[:, fgas.index(g) + 1, :]
    gas_minor, ident = _names(V["gas_minor"]), _names(V["identifier_minor"])

    def minor(sfx):
        gases = _names(V[f"minor_gases_{sfx}"])
        scal = _names(V[f"scaling_gas_{sfx}"])
        lims = np.asarray(V[f"minor_limits_gpt_{sfx}"][:])
        dens = np.asarray(V[f"minor_scales_with_density_{sfx}"][:])
        comp = np.asarray(V[f"scale_by_complement_{sfx}"][:])
        start = np.asarray(V[f"kminor_start_{sfx}"][:])
        km = np.asarray(V[f"kminor_{sfx}"][:])            # (temperature, mixing_fraction, contributors)
        keep = [i for i in range(len(gases)) if gas_minor[ident.index(gases[i])] in active]
        ks, starts, blocks = 1, [], []
        for i in keep:
            w = lims[i, 1] - lims[i, 0] + 1
            starts.append(ks)
            blocks.append(km[:, :, start[i] - 1:start[i] - 1 + w])
            ks += w
        k = np.concatenate(blocks, axis=2) if blocks else np.zeros((ntemp, km.shape[1], 1))
        return dict(n=len(keep), ncontrib=int(ks - 1), limits_gpt=lims[keep].astype(np.int32),
                    scales_with_density=(dens[keep] != 0).astype(np.int32),
                    idx_scaling=np.array([slot.get(scal[i], 0) if scal[i] in kept else 0 for i in keep],
                                         dtype=np.int32),
                    scale_by_complement=(comp[keep] != 0).astype(np.int32),
                    idx_minor=np.array([SLOTS.index(gas_minor[ident.index(gases[i])]) + 1 for i in keep],
                                       dtype=np.int32),
                    kminor_start=np.array(starts, dtype=np.int32), kminor=np.ascontiguousarray(k))

    km = np.ascontiguousarray(V["kmajor"][:])
    kd = KDist(is_lw=is_lw, ngas=len(SLOTS), nband=nbnd, ngpt=ngpt, ntemp=ntemp, npres=len(V["press_ref"][:]),
               neta=km.shape[2], nflav=len(flav), band_lims_gpt=b2g.astype(np.int32),
               band_lims_wvn=np.asarray(V["bnd_limits_wavenumber"][:], dtype=np.float64),
               flavor=np.array(flav, dtype=np.int32), gpoint_flavor=gpf,
               press_ref=np.asarray(V["press_ref"][:], dtype=np.float64),
               temp_ref=np.asarray(V["temp_ref"][:], dtype=np.float64),
               press_ref_trop=float(V["press_ref_trop"].getValue()), vmr_ref=np.ascontiguousarray(vr),
               kmajor=km, minor_lower=minor("lower"), minor_upper=minor("upper"))
    if is_lw:
        kd.planck_frac = np.ascontiguousarray(V["plank_fraction"][:])
        kd.totplnk = np.ascontiguousarray(V["totplnk"][:])
        kd.optimal_angle_fit = np.ascontiguousarray(V["optimal_angle_fit"][:])
    else:
        kd.krayl_lower = np.ascontiguousarray(V["rayl_lower"][:])
        kd.krayl_upper = np.ascontiguousarray(V["rayl_upper"][:])
        mg = float(V["mg_default"].getValue()) if "mg_default" in V else 0.1567652
        sb = float(V["sb_default"].getValue()) if "sb_default" in V else 902.7126
        kd.solar_src = np.ascontiguousarray(V["solar_source_quiet"][:] + (mg - 0.1495954) * V["solar_source_facular"][:]
                                            + (sb - 0.00066696) * V["solar_source_sunspot"][:])
    f.close()
    return kd
