// Host side: k-distribution and cloud-optics tables from the netCDF files
// the reference's _init reads (nc_tables.hip).  The holders own the arrays
// the ccpprad_kdist / ccpprad_cloud_lut views point into.
#pragma once
#include "../../include/ccpprad.h"

#include <string>
#include <vector>

namespace ccpprad {

struct NcMinorHold {
    ccpprad_minor m{};
    std::vector<int> limits, dens, scaling, comp, idx, start;
    std::vector<double> kminor;
};

struct NcKdistHold {
    ccpprad_kdist kd{};
    std::vector<int> band_gpt, flavor, gpt_flav;
    std::vector<double> band_wvn, press_ref, temp_ref, vmr_ref, kmajor, planck, totplnk, fit;
    std::vector<double> rayl_lower, rayl_upper, solar;
    NcMinorHold lower, upper;
};

struct NcLutHold {
    ccpprad_cloud_lut lut{};
    std::vector<double> extliq, ssaliq, asyliq, extice, ssaice, asyice;
    std::vector<double> pade[6], sizreg[6];   // Pade form: ext, ssa, asy (liquid, then ice)
};

// true when the file starts with a netCDF classic or netCDF-4 (HDF5) signature
bool is_netcdf_file(const char *path);
// the k-distribution, reduced to the active gases (gas_names) as rte-rrtmgp's
// load does; false with a message in err on any missing or malformed field
bool load_kdist_nc(const char *path, int want_lw, const char *const *gas_names, int ngases, NcKdistHold &h,
                   std::string &err);
// the cloud optics: the lookup tables when the file has them, else the Pade
// coefficients (the last ice-roughness category selected)
// (force_pade: the Pade coefficients even when the file also has tables)
bool load_lut_nc(const char *path, NcLutHold &h, std::string &err, bool force_pade = false);

}  // namespace ccpprad
