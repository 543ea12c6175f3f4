// Host-pointer scheme entries (the drop-in path the Fortran scheme modules
// call) and the table-file loader used by *_init_files.
//
// ccpprad_lw_run / ccpprad_sw_run copy every non-NULL array of the argument
// list to device staging buffers (outputs too: the CCPP outputs are
// intent(inout), rrtmgp_lw_main.F90:183-190 / rrtmgp_sw_main.F90:175-182, and
// columns the scheme does not touch keep their values), run the device entry
// on a private stream, copy the outputs back and synchronise.  Outputs the
// scheme overwrites in every column are not copied in: the LW level fluxes
// (the reference zeroes or writes them for every block,
// rrtmgp_lw_main.F90:281-292,529-533,604-609) and the LW heating rate.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <map>

#include "../../include/ccpprad.h"
#include "nc_tables.h"

extern "C" int ccpprad_ctx_stage(ccpprad_ctx *c, void *stream, int slot, size_t bytes, void **p);
extern "C" void *ccpprad_ctx_host_stream(ccpprad_ctx *c);

namespace {

using ccpprad::NcKdistHold;
using ccpprad::NcLutHold;
using ccpprad::is_netcdf_file;
using ccpprad::load_kdist_nc;
using ccpprad::load_lut_nc;

enum Shape { S_COL, S_LAY, S_LEV, S_BAND, S_SCMP, S_DAY };
struct Field {
    size_t off;
    Shape shape;
    bool is_int;
    bool out;
    bool whole = false;   // an output written in every column: not copied in
};

#define LWF(name, sh, isint, out) {offsetof(ccpprad_lw_args, name), sh, isint, out}
#define LWW(name) {offsetof(ccpprad_lw_args, name), S_LEV, false, true, true}
#define SWF(name, sh, isint, out) {offsetof(ccpprad_sw_args, name), sh, isint, out}

const Field LW_FIELDS[] = {
    LWF(icseed_lw, S_COL, true, false), LWF(semis, S_COL, false, false), LWF(tsfg, S_COL, false, false),
    LWF(p_lay, S_LAY, false, false), LWF(p_lev, S_LEV, false, false), LWF(t_lay, S_LAY, false, false),
    LWF(t_lev, S_LEV, false, false), LWF(vmr_o2, S_LAY, false, false), LWF(vmr_h2o, S_LAY, false, false),
    LWF(vmr_o3, S_LAY, false, false), LWF(vmr_ch4, S_LAY, false, false), LWF(vmr_n2o, S_LAY, false, false),
    LWF(vmr_co2, S_LAY, false, false), LWF(cld_frac, S_LAY, false, false), LWF(cld_lwp, S_LAY, false, false),
    LWF(cld_reliq, S_LAY, false, false), LWF(cld_iwp, S_LAY, false, false), LWF(cld_reice, S_LAY, false, false),
    LWF(cld_swp, S_LAY, false, false), LWF(cld_resnow, S_LAY, false, false), LWF(cld_rwp, S_LAY, false, false),
    LWF(cld_rerain, S_LAY, false, false), LWF(precip_frac, S_LAY, false, false),
    LWF(cld_cnv_lwp, S_LAY, false, false), LWF(cld_cnv_reliq, S_LAY, false, false),
    LWF(cld_cnv_iwp, S_LAY, false, false), LWF(cld_cnv_reice, S_LAY, false, false),
    LWF(cld_pbl_lwp, S_LAY, false, false), LWF(cld_pbl_reliq, S_LAY, false, false),
    LWF(cld_pbl_iwp, S_LAY, false, false), LWF(cld_pbl_reice, S_LAY, false, false),
    LWF(cloud_overlap_param, S_LAY, false, false), LWF(aerlw_tau, S_BAND, false, false),
    LWF(aerlw_ssa, S_BAND, false, false), LWF(aerlw_g, S_BAND, false, false),
    LWW(fluxlwUP_allsky), LWW(fluxlwDOWN_allsky), LWW(fluxlwUP_clrsky), LWW(fluxlwDOWN_clrsky),
    LWF(fluxlwUP_jac, S_LEV, false, true), LWW(fluxlwUP_radtime), LWW(fluxlwDOWN_radtime),
    {offsetof(ccpprad_lw_args, htrlw), S_LAY, false, true, true}, LWF(htrlwc, S_LAY, false, true),
};

const Field SW_FIELDS[] = {
    SWF(idx, S_DAY, true, false), SWF(icseed_sw, S_COL, true, false),
    SWF(sfc_alb_nir_dir, S_COL, false, false), SWF(sfc_alb_nir_dif, S_COL, false, false),
    SWF(sfc_alb_uvvis_dir, S_COL, false, false), SWF(sfc_alb_uvvis_dif, S_COL, false, false),
    SWF(coszen, S_COL, false, false), SWF(p_lay, S_LAY, false, false), SWF(p_lev, S_LEV, false, false),
    SWF(t_lay, S_LAY, false, false), SWF(t_lev, S_LEV, false, false), SWF(vmr_o2, S_LAY, false, false),
    SWF(vmr_h2o, S_LAY, false, false), SWF(vmr_o3, S_LAY, false, false), SWF(vmr_ch4, S_LAY, false, false),
    SWF(vmr_n2o, S_LAY, false, false), SWF(vmr_co2, S_LAY, false, false), SWF(cld_frac, S_LAY, false, false),
    SWF(cld_lwp, S_LAY, false, false), SWF(cld_reliq, S_LAY, false, false), SWF(cld_iwp, S_LAY, false, false),
    SWF(cld_reice, S_LAY, false, false), SWF(cld_swp, S_LAY, false, false), SWF(cld_resnow, S_LAY, false, false),
    SWF(cld_rwp, S_LAY, false, false), SWF(cld_rerain, S_LAY, false, false), SWF(precip_frac, S_LAY, false, false),
    SWF(cld_cnv_lwp, S_LAY, false, false), SWF(cld_cnv_reliq, S_LAY, false, false),
    SWF(cld_cnv_iwp, S_LAY, false, false), SWF(cld_cnv_reice, S_LAY, false, false),
    SWF(cld_pbl_lwp, S_LAY, false, false), SWF(cld_pbl_reliq, S_LAY, false, false),
    SWF(cld_pbl_iwp, S_LAY, false, false), SWF(cld_pbl_reice, S_LAY, false, false),
    SWF(cloud_overlap_param, S_LAY, false, false), SWF(aersw_tau, S_BAND, false, false),
    SWF(aersw_ssa, S_BAND, false, false), SWF(aersw_g, S_BAND, false, false), SWF(scmpsw, S_SCMP, false, true),
    SWF(fluxswUP_allsky, S_LEV, false, true), SWF(fluxswDOWN_allsky, S_LEV, false, true),
    SWF(fluxswUP_clrsky, S_LEV, false, true), SWF(fluxswDOWN_clrsky, S_LEV, false, true),
    SWF(cldtausw, S_LAY, false, true), SWF(htrsw, S_LAY, false, true), SWF(htrswc, S_LAY, false, true),
};

size_t field_elems(Shape s, size_t ncol, size_t nlay, size_t nband, size_t nday)
{
    switch (s) {
    case S_COL: return ncol;
    case S_LAY: return ncol * nlay;
    case S_LEV: return ncol * (nlay + 1);
    case S_BAND: return ncol * nlay * nband;
    case S_SCMP: return 6 * ncol;
    case S_DAY: return nday;
    }
    return 0;
}

template <typename Args, size_t N>
int run_host(ccpprad_ctx *c, const Args *a, const Field (&F)[N], int ncol, int nlay, int nday,
             int (*dev)(ccpprad_ctx *, const Args *, void *, int, char *, int), char *errmsg, int errlen)
{
    if (!c || !a) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "null context or arguments");
        return 1;
    }
    const size_t nband = (size_t)ccpprad_nband(c);
    Args d = *a;
    hipStream_t st = (hipStream_t)ccpprad_ctx_host_stream(c);
    if (!st) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "stream creation failed");
        return 2;
    }
    for (size_t i = 0; i < N; ++i) {
        void **hp = (void **)((char *)&d + F[i].off);
        if (!*hp) continue;
        const size_t n = field_elems(F[i].shape, (size_t)ncol, (size_t)nlay, nband, (size_t)nday);
        const size_t bytes = n * (F[i].is_int ? sizeof(int) : sizeof(double));
        void *dp = nullptr;
        int rc = ccpprad_ctx_stage(c, (void *)st, (int)i, bytes, &dp);
        if (rc) {
            if (errmsg && errlen > 0) snprintf(errmsg, errlen, "device staging allocation failed");
            return 2;
        }
        if (bytes && !F[i].whole && hipMemcpyAsync(dp, *hp, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
            if (errmsg && errlen > 0) snprintf(errmsg, errlen, "host to device copy failed");
            return 2;
        }
        *hp = dp;
    }
    int rc = dev(c, &d, (void *)st, 1, errmsg, errlen);
    if (rc) return rc;
    for (size_t i = 0; i < N; ++i) {
        if (!F[i].out) continue;
        void *const *hp = (void *const *)((const char *)a + F[i].off);
        void *const *dp = (void *const *)((const char *)&d + F[i].off);
        if (!*hp) continue;
        const size_t n = field_elems(F[i].shape, (size_t)ncol, (size_t)nlay, nband, (size_t)nday);
        if (n && hipMemcpyAsync(*hp, *dp, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) {
            if (errmsg && errlen > 0) snprintf(errmsg, errlen, "device to host copy failed");
            return 2;
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "stream synchronisation failed");
        return 2;
    }
    return 0;
}

}  // namespace

extern "C" int ccpprad_lw_run(ccpprad_ctx *c, const ccpprad_lw_args *a, char *errmsg, int errlen)
{
    if (a && !a->doLWrad) return 0;
    return run_host(c, a, LW_FIELDS, a ? a->nCol : 0, a ? a->nLay : 0, 0, ccpprad_lw_run_dev, errmsg, errlen);
}

extern "C" int ccpprad_sw_run(ccpprad_ctx *c, const ccpprad_sw_args *a, char *errmsg, int errlen)
{
    if (a && !a->doSWrad) return 0;
    return run_host(c, a, SW_FIELDS, a ? a->nCol : 0, a ? a->nLay : 0, a ? (a->nDay > 0 ? a->nDay : 0) : 0,
                    ccpprad_sw_run_dev, errmsg, errlen);
}

// ---------------------------------------------------------------------------
// Table files: "CCPPRADT", int32 version (1), int32 record count, then per
// record: char name[32], int32 type (1 int32, 2 float64), int32 ndim,
// int64 dims[4], data in Fortran order.  Written by ccpprad.tables.
// ---------------------------------------------------------------------------
namespace {
struct Rec {
    int type = 0;
    std::vector<long long> dims;
    std::vector<int> i;
    std::vector<double> d;
};
using RecMap = std::map<std::string, Rec>;

bool read_file(const char *path, RecMap &m, std::string &err)
{
    FILE *f = fopen(path, "rb");
    if (!f) { err = std::string("cannot open ") + path; return false; }
    char magic[8];
    int ver = 0, nrec = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "CCPPRADT", 8) == 0 && fread(&ver, 4, 1, f) == 1 &&
              ver == 1 && fread(&nrec, 4, 1, f) == 1 && nrec >= 0 && nrec < 100000;
    for (int r = 0; ok && r < nrec; ++r) {
        char name[33] = {0};
        int type = 0, ndim = 0;
        long long dims[4];
        ok = fread(name, 1, 32, f) == 32 && fread(&type, 4, 1, f) == 1 && fread(&ndim, 4, 1, f) == 1 &&
             fread(dims, 8, 4, f) == 4 && ndim >= 0 && ndim <= 4 && (type == 1 || type == 2);
        if (!ok) break;
        long long n = 1;
        Rec rec;
        rec.type = type;
        for (int k = 0; k < ndim; ++k) { if (dims[k] < 0 || dims[k] > (1LL << 32)) ok = false; n *= dims[k]; rec.dims.push_back(dims[k]); }
        if (!ok || n > (1LL << 31)) { ok = false; break; }
        if (type == 1) { rec.i.resize((size_t)n); ok = fread(rec.i.data(), 4, (size_t)n, f) == (size_t)n; }
        else { rec.d.resize((size_t)n); ok = fread(rec.d.data(), 8, (size_t)n, f) == (size_t)n; }
        m[name] = std::move(rec);
    }
    fclose(f);
    if (!ok) err = std::string("malformed table file ") + path;
    return ok;
}

struct TableHold {
    RecMap k, l;
    ccpprad_kdist kd{};
    ccpprad_cloud_lut lut{};
};

int geti(RecMap &m, const char *n, bool *ok) {
    auto it = m.find(n);
    if (it == m.end() || it->second.type != 1 || it->second.i.empty()) { *ok = false; return 0; }
    return it->second.i[0];
}
double getd(RecMap &m, const char *n, bool *ok) {
    auto it = m.find(n);
    if (it == m.end() || it->second.type != 2 || it->second.d.empty()) { *ok = false; return 0; }
    return it->second.d[0];
}
const int *pi(RecMap &m, const char *n, bool *ok, size_t need) {
    auto it = m.find(n);
    if (it == m.end() || it->second.type != 1 || it->second.i.size() < need) { *ok = false; return nullptr; }
    return it->second.i.data();
}
const double *pd(RecMap &m, const char *n, bool *ok, size_t need, bool optional = false) {
    auto it = m.find(n);
    if (it == m.end()) { if (!optional) *ok = false; return nullptr; }
    if (it->second.type != 2 || it->second.d.size() < need) { *ok = false; return nullptr; }
    return it->second.d.data();
}

bool fill_minor(RecMap &m, const char *pre, int neta, int ntemp, ccpprad_minor &mm)
{
    bool ok = true;
    std::string p(pre);
    mm.n = geti(m, (p + "n").c_str(), &ok);
    mm.ncontrib = geti(m, (p + "ncontrib").c_str(), &ok);
    if (!ok || mm.n < 0 || mm.ncontrib < 0) return false;
    const size_t n = (size_t)mm.n;
    mm.limits_gpt = pi(m, (p + "limits_gpt").c_str(), &ok, 2 * n);
    mm.scales_with_density = pi(m, (p + "scales_with_density").c_str(), &ok, n);
    mm.idx_scaling = pi(m, (p + "idx_scaling").c_str(), &ok, n);
    mm.scale_by_complement = pi(m, (p + "scale_by_complement").c_str(), &ok, n);
    mm.idx_minor = pi(m, (p + "idx_minor").c_str(), &ok, n);
    mm.kminor_start = pi(m, (p + "kminor_start").c_str(), &ok, n);
    mm.kminor = pd(m, (p + "kminor").c_str(), &ok, (size_t)mm.ncontrib * neta * ntemp);
    return ok;
}

// The binary table file's k-distribution records into h.kd.
bool kdist_from_records(TableHold &h)
{
    bool ok = true;
    ccpprad_kdist &k = h.kd;
    k.is_lw = geti(h.k, "is_lw", &ok);
    k.ngas = geti(h.k, "ngas", &ok); k.nband = geti(h.k, "nband", &ok); k.ngpt = geti(h.k, "ngpt", &ok);
    k.ntemp = geti(h.k, "ntemp", &ok); k.npres = geti(h.k, "npres", &ok); k.neta = geti(h.k, "neta", &ok);
    k.nflav = geti(h.k, "nflav", &ok); k.nplanck = geti(h.k, "nplanck", &ok);
    if (ok && (k.ngpt <= 0 || k.nband <= 0 || k.ntemp <= 1 || k.npres <= 1 || k.neta <= 1 || k.nflav <= 0 || k.ngas <= 0))
        ok = false;
    if (ok) {
        const size_t nk = (size_t)k.ngpt * k.neta * (k.npres + 1) * k.ntemp;
        k.band_lims_gpt = pi(h.k, "band_lims_gpt", &ok, 2 * (size_t)k.nband);
        k.band_lims_wvn = pd(h.k, "band_lims_wvn", &ok, 2 * (size_t)k.nband);
        k.flavor = pi(h.k, "flavor", &ok, 2 * (size_t)k.nflav);
        k.gpoint_flavor = pi(h.k, "gpoint_flavor", &ok, 2 * (size_t)k.ngpt);
        k.press_ref = pd(h.k, "press_ref", &ok, k.npres);
        k.temp_ref = pd(h.k, "temp_ref", &ok, k.ntemp);
        k.press_ref_trop = getd(h.k, "press_ref_trop", &ok);
        k.vmr_ref = pd(h.k, "vmr_ref", &ok, 2 * (size_t)(k.ngas + 1) * k.ntemp);
        k.kmajor = pd(h.k, "kmajor", &ok, nk);
        ok = ok && fill_minor(h.k, "minor_lower_", k.neta, k.ntemp, k.minor_lower) &&
             fill_minor(h.k, "minor_upper_", k.neta, k.ntemp, k.minor_upper);
        if (k.is_lw) {
            k.planck_frac = pd(h.k, "planck_frac", &ok, nk);
            k.totplnk = pd(h.k, "totplnk", &ok, (size_t)k.nplanck * k.nband);
            k.optimal_angle_fit = pd(h.k, "optimal_angle_fit", &ok, 2 * (size_t)k.nband, true);
        } else {
            const size_t nr = (size_t)k.ngpt * k.neta * k.ntemp;
            k.krayl_lower = pd(h.k, "krayl_lower", &ok, nr);
            k.krayl_upper = pd(h.k, "krayl_upper", &ok, nr);
            k.solar_src = pd(h.k, "solar_src", &ok, k.ngpt);
        }
    }
    return ok;
}

// The binary table file's cloud records into h.lut.
bool lut_from_records(TableHold &h)
{
    bool ok = true;
    ccpprad_cloud_lut &L = h.lut;
    L.nband = geti(h.l, "nband", &ok); L.nsize_liq = geti(h.l, "nsize_liq", &ok);
    L.nsize_ice = geti(h.l, "nsize_ice", &ok); L.nrghice = geti(h.l, "nrghice", &ok);
    L.icergh = geti(h.l, "icergh", &ok);
    L.radliq_lwr = getd(h.l, "radliq_lwr", &ok); L.radliq_upr = getd(h.l, "radliq_upr", &ok);
    L.radice_lwr = getd(h.l, "radice_lwr", &ok); L.radice_upr = getd(h.l, "radice_upr", &ok);
    if (ok && (L.nband <= 0 || L.nsize_liq < 2 || L.nsize_ice < 2 || L.nrghice < 1)) ok = false;
    if (ok) {
        const size_t nl = (size_t)L.nsize_liq * L.nband, ni = (size_t)L.nsize_ice * L.nband * L.nrghice;
        L.extliq = pd(h.l, "extliq", &ok, nl); L.ssaliq = pd(h.l, "ssaliq", &ok, nl); L.asyliq = pd(h.l, "asyliq", &ok, nl);
        L.extice = pd(h.l, "extice", &ok, ni); L.ssaice = pd(h.l, "ssaice", &ok, ni); L.asyice = pd(h.l, "asyice", &ok, ni);
    }
    return ok;
}

// Tables from the k-distribution and cloud files named at _init: the
// reference's netCDF files (read natively, classic formats) or this
// package's binary table files, each file's format taken from its signature.
int init_files(int lw, const char *gas_file, const char *cloud_file, int nrghice, const char *const *gas_names,
               int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg, int errlen, bool pade = false)
{
    TableHold h;
    NcKdistHold kh;
    NcLutHold lh;
    std::string err;
    auto fail = [&](const char *file, const std::string &e) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "%s: %s", file ? file : "(null)", e.c_str());
        return 1;
    };
    if (!gas_file) return fail(gas_file, "null gas file");
    const ccpprad_kdist *kd = nullptr;
    const ccpprad_cloud_lut *lut = nullptr;
    if (is_netcdf_file(gas_file)) {
        if (!load_kdist_nc(gas_file, lw, gas_names, ngases, kh, err)) return fail(gas_file, err);
        kd = &kh.kd;
    } else {
        if (!read_file(gas_file, h.k, err)) return fail(gas_file, err);
        if (!kdist_from_records(h)) return fail(gas_file, "table file lacks a required record or has a bad size");
        kd = &h.kd;
    }
    if (cloud_file && cloud_file[0]) {
        if (is_netcdf_file(cloud_file)) {
            if (!load_lut_nc(cloud_file, lh, err, pade)) return fail(cloud_file, err);
            lut = &lh.lut;
        } else {
            if (pade) return fail(cloud_file, "Pade cloud optics come from a netCDF cloud file");
            if (!read_file(cloud_file, h.l, err)) return fail(cloud_file, err);
            if (!lut_from_records(h)) return fail(cloud_file, "table file lacks a required record or has a bad size");
            lut = &h.lut;
        }
        // the namelist's ice-roughness category overrides the file's when
        // nonzero (rrtmgp_lw_cloud_optics.F90:171-173,377;
        // rrtmgp_sw_cloud_optics.F90:160-162,366)
        if (nrghice < 0 || nrghice > lut->nrghice) {
            char m[96];
            snprintf(m, sizeof m, "nrghice=%d outside 0..%d (the file's ice-roughness categories)", nrghice,
                     lut->nrghice);
            return fail(cloud_file, m);
        }
        if (nrghice > 0) (lut == &lh.lut ? lh.lut : h.lut).icergh = nrghice;
    }
    return lw ? ccpprad_lw_init(kd, lut, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen)
              : ccpprad_sw_init(kd, lut, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen);
}
}  // namespace

extern "C" int ccpprad_lw_init_files(const char *gas_file, const char *cloud_file, int nrghice,
                                     const char *const *gas_names,
                                     int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg, int errlen)
{
    return init_files(1, gas_file, cloud_file, nrghice, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen);
}

extern "C" int ccpprad_sw_init_files(const char *gas_file, const char *cloud_file, int nrghice,
                                     const char *const *gas_names,
                                     int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg, int errlen)
{
    return init_files(0, gas_file, cloud_file, nrghice, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen);
}

extern "C" int ccpprad_lw_init_files_pade(const char *gas_file, const char *cloud_file, int nrghice,
                                     const char *const *gas_names,
                                          int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg,
                                          int errlen)
{
    return init_files(1, gas_file, cloud_file, nrghice, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen, true);
}

extern "C" int ccpprad_sw_init_files_pade(const char *gas_file, const char *cloud_file, int nrghice,
                                     const char *const *gas_names,
                                          int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg,
                                          int errlen)
{
    return init_files(0, gas_file, cloud_file, nrghice, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen, true);
}

// compute_heating_rate with host arrays (GFS_rrtmgp_post.F90:159-171,249-266):
// the fluxes and pressures are copied to the device, the FP64 kernel runs
// there and the heating rate is copied back.  Synchronous; the device
// staging is allocated per call.
namespace ccpprad {
void launch_heating_rate(int ncol, int nlay, const double *up, const double *dn, const double *p_lev,
                         double *htr, hipStream_t st);
void launch_mark_days(int nday, const int *idx, unsigned char *day, hipStream_t st);           // lw_solver.hip
void launch_heating_rate_day(int ncol, int nlay, const unsigned char *day, const double *up, const double *dn,
                             const double *p_lev, double *htr, hipStream_t st);
}

namespace {

// Staging of a context-free host entry: one device block and one private
// stream per call.  Each Stage names a pointer field of the caller's
// argument copy (a host pointer on entry, replaced by its device copy);
// IN arrays are copied in, OUT arrays back, SCRATCH ones are device-only.
enum { ST_IN = 1, ST_OUT = 2, ST_SCRATCH = 4 };
struct Stage {
    void **slot;
    size_t bytes;
    int mode;
};

template <typename Dev>
int staged_call(const char *who, std::vector<Stage> &S, Dev &&dev, char *errmsg, int errlen)
{
    std::vector<void *> host(S.size());
    std::vector<size_t> off(S.size());
    size_t total = 0;
    for (size_t i = 0; i < S.size(); ++i) {
        host[i] = *S[i].slot;
        off[i] = total;
        if (host[i] || (S[i].mode & ST_SCRATCH)) total += (S[i].bytes + 255) & ~(size_t)255;
    }
    int rc = 0;
    auto fail = [&](const char *what, hipError_t e) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "%s: %s: %s", who, what, hipGetErrorString(e));
        rc = 2;
    };
    hipStream_t st = nullptr;
    char *d = nullptr;
    hipError_t e = hipStreamCreate(&st);
    if (e != hipSuccess) { fail("stream creation", e); return rc; }
    if (total && (e = hipMalloc(&d, total)) != hipSuccess) {
        fail("device allocation", e);
        (void)hipStreamDestroy(st);
        return rc;
    }
    for (size_t i = 0; i < S.size() && !rc; ++i) {
        if (!host[i] && !(S[i].mode & ST_SCRATCH)) continue;
        *S[i].slot = d + off[i];
        if ((S[i].mode & ST_IN) && S[i].bytes &&
            (e = hipMemcpyAsync(d + off[i], host[i], S[i].bytes, hipMemcpyHostToDevice, st)) != hipSuccess)
            fail("host to device copy", e);
    }
    if (!rc) rc = dev(st);
    for (size_t i = 0; i < S.size() && !rc; ++i) {
        if (!host[i] || !(S[i].mode & ST_OUT) || !S[i].bytes) continue;
        if ((e = hipMemcpyAsync(host[i], d + off[i], S[i].bytes, hipMemcpyDeviceToHost, st)) != hipSuccess)
            fail("device to host copy", e);
    }
    if (!rc && (e = hipStreamSynchronize(st)) != hipSuccess) fail("stream synchronisation", e);
    (void)hipStreamSynchronize(st);
    if (d) (void)hipFree(d);
    (void)hipStreamDestroy(st);
    return rc;
}

template <typename T>
Stage stage(T *&p, size_t n, int mode) { return Stage{(void **)&p, n * sizeof(T), mode}; }

}  // namespace

extern "C" int ccpprad_cloud_overlap(const ccpprad_cloud_overlap_args *a, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!a) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "cloud_overlap: null arguments");
        return 1;
    }
    ccpprad_cloud_overlap_args d = *a;
    if (d.nCol < 0 || d.nLev < 1 || d.nCol == 0 || !(d.doLWrad || d.doSWrad))   // checked or a no-op there
        return ccpprad_cloud_overlap_dev(&d, nullptr, errmsg, errlen);
    const size_t nc = (size_t)d.nCol, nl = (size_t)d.nLev;
    std::vector<Stage> S = {
        stage(d.lat, nc, ST_IN), stage(d.deltaZc, nc * nl, ST_IN), stage(d.cld_frac, nc * nl, ST_IN),
        stage(d.cld_cnv_frac, nc * nl, ST_IN), stage(d.de_lgth, nc, ST_IN | ST_OUT),
        stage(d.cloud_overlap_param, nc * nl, ST_IN | ST_OUT),
        stage(d.cnv_cloud_overlap_param, nc * nl, ST_IN | ST_OUT),
        stage(d.precip_overlap_param, nc * nl, ST_IN | ST_OUT)};
    return staged_call("cloud_overlap", S,
                       [&](hipStream_t st) { return ccpprad_cloud_overlap_dev(&d, (void *)st, errmsg, errlen); },
                       errmsg, errlen);
}

extern "C" int ccpprad_post(const ccpprad_post_args *a, const ccpprad_post_hr_args *h, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    auto err = [&](const char *m) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "post: %s", m);
        return 1;
    };
    if (!a || !h) return err("null arguments");
    if (a->nCol < 0 || a->nLev < 1) return err("bad dimensions");
    if (a->nCol == 0 || !(a->doLWrad || a->doSWrad)) return 0;
    ccpprad_post_args d = *a;
    ccpprad_post_hr_args r = *h;
    const int nc = a->nCol, nday = a->doSWrad ? a->nDay : 0;
    if (!a->doLWrad) r.htrlw = r.htrlwc = nullptr;
    if (!a->doSWrad) r.htrsw = r.htrswc = nullptr;
    if (!r.do_lw_clrsky_hr) r.htrlwc = nullptr;
    if (!r.do_sw_clrsky_hr) r.htrswc = nullptr;
    if (a->doLWrad && !r.htrlw) return err("LW: htrlw is NULL");
    if (a->doLWrad && r.do_lw_clrsky_hr && !r.htrlwc) return err("do_lw_clrsky_hr: htrlwc is NULL");
    if (a->doSWrad && !r.htrsw) return err("SW: htrsw is NULL");
    if (a->doSWrad && r.do_sw_clrsky_hr && !r.htrswc) return err("do_sw_clrsky_hr: htrswc is NULL");
    if (a->doSWrad && (a->nDay < 0 || a->nDay > nc)) return err("nDay outside 0..nCol");
    if (nday > 0) {
        if (!r.idxday) return err("idxday is NULL");
        std::vector<unsigned char> seen((size_t)nc, 0);
        for (int i = 0; i < nday; ++i) {
            const int c = r.idxday[i];
            if (c < 1 || c > nc || seen[(size_t)c - 1]) return err("idxday(1:nDay) must be distinct columns in 1..nCol");
            seen[(size_t)c - 1] = 1;
        }
    } else {
        r.idxday = nullptr;
    }
    const size_t n = (size_t)nc, nl = (size_t)a->nLev, nv = n * (nl + 1);
    unsigned char *day = nullptr;
    std::vector<Stage> S = {
        stage(d.fluxlwUP_allsky, nv, ST_IN), stage(d.fluxlwDOWN_allsky, nv, ST_IN),
        stage(d.fluxlwUP_clrsky, nv, ST_IN), stage(d.fluxlwDOWN_clrsky, nv, ST_IN),
        stage(d.fluxswUP_allsky, nv, ST_IN), stage(d.fluxswDOWN_allsky, nv, ST_IN),
        stage(d.fluxswUP_clrsky, nv, ST_IN), stage(d.fluxswDOWN_clrsky, nv, ST_IN),
        stage(d.scmpsw, 6 * n, ST_IN), stage(d.sfc_alb_nir_dir, n, ST_IN), stage(d.sfc_alb_nir_dif, n, ST_IN),
        stage(d.sfc_alb_uvvis_dir, n, ST_IN), stage(d.sfc_alb_uvvis_dif, n, ST_IN), stage(d.tsfa, n, ST_IN),
        stage(d.coszen, n, ST_IN), stage(d.coszdg, n, ST_IN), stage(d.p_lev, nv, ST_IN),
        stage(d.cldsa, 5 * n, ST_IN), stage(d.cldtaulw, n * nl, ST_IN), stage(d.cldtausw, n * nl, ST_IN),
        stage(d.aerodp, n * (size_t)(a->naerodp > 0 ? a->naerodp : 0), ST_IN), stage(d.mtopa, 3 * n, ST_IN),
        stage(d.mbota, 3 * n, ST_IN), stage(d.topflw, 2 * n, ST_IN | ST_OUT), stage(d.sfcflw, 4 * n, ST_IN | ST_OUT),
        stage(d.sfcdlw, n, ST_IN | ST_OUT), stage(d.sfculw, n, ST_IN | ST_OUT), stage(d.tsflw, n, ST_IN | ST_OUT),
        stage(d.htrlwu, n * nl, ST_IN | ST_OUT), stage(d.topfsw, 3 * n, ST_IN | ST_OUT),
        stage(d.sfcfsw, 4 * n, ST_IN | ST_OUT), stage(d.nirbmdi, n, ST_IN | ST_OUT),
        stage(d.nirdfdi, n, ST_IN | ST_OUT), stage(d.visbmdi, n, ST_IN | ST_OUT), stage(d.visdfdi, n, ST_IN | ST_OUT),
        stage(d.nirbmui, n, ST_IN | ST_OUT), stage(d.nirdfui, n, ST_IN | ST_OUT), stage(d.visbmui, n, ST_IN | ST_OUT),
        stage(d.visdfui, n, ST_IN | ST_OUT), stage(d.sfcnsw, n, ST_IN | ST_OUT), stage(d.sfcdsw, n, ST_IN | ST_OUT),
        stage(d.fluxr, n * (size_t)(a->nfxr > 0 ? a->nfxr : 0), ST_IN | ST_OUT),
        stage(r.idxday, (size_t)(nday > 0 ? nday : 0), ST_IN), stage(r.htrlw, n * nl, ST_OUT),
        stage(r.htrlwc, n * nl, ST_OUT), stage(r.htrsw, n * nl, ST_OUT), stage(r.htrswc, n * nl, ST_OUT),
        stage(day, a->doSWrad ? n : 0, ST_SCRATCH)};
    d.htrlw = nullptr;   // the LW heating rate this call computes (r.htrlw) feeds htrlwu
    auto dev = [&](hipStream_t st) -> int {
        const int nlay = a->nLev;
        if (a->doLWrad) {
            if (!d.fluxlwUP_allsky || !d.fluxlwDOWN_allsky || !d.p_lev ||
                (r.htrlwc && (!d.fluxlwUP_clrsky || !d.fluxlwDOWN_clrsky)))
                return err("LW heating rate: NULL flux or p_lev");
            if (r.htrlwc) ccpprad::launch_heating_rate(nc, nlay, d.fluxlwUP_clrsky, d.fluxlwDOWN_clrsky, d.p_lev,
                                                       r.htrlwc, st);
            ccpprad::launch_heating_rate(nc, nlay, d.fluxlwUP_allsky, d.fluxlwDOWN_allsky, d.p_lev, r.htrlw, st);
            d.htrlw = r.htrlw;
        }
        if (a->doSWrad) {
            if (!d.fluxswUP_allsky || !d.fluxswDOWN_allsky || !d.p_lev ||
                (r.htrswc && (!d.fluxswUP_clrsky || !d.fluxswDOWN_clrsky)))
                return err("SW heating rate: NULL flux or p_lev");
            hipError_t e = hipMemsetAsync(day, 0, n, st);
            if (e != hipSuccess) return err("flag clear failed");
            ccpprad::launch_mark_days(nday, r.idxday, day, st);
            if (r.htrswc) ccpprad::launch_heating_rate_day(nc, nlay, day, d.fluxswUP_clrsky, d.fluxswDOWN_clrsky,
                                                           d.p_lev, r.htrswc, st);
            ccpprad::launch_heating_rate_day(nc, nlay, day, d.fluxswUP_allsky, d.fluxswDOWN_allsky, d.p_lev,
                                             r.htrsw, st);
        }
        if (hipGetLastError() != hipSuccess) return err("heating-rate launch failed");
        return ccpprad_post_dev(&d, (void *)st, errmsg, errlen);
    };
    return staged_call("post", S, dev, errmsg, errlen);
}

extern "C" int ccpprad_heating_rate(int nCol, int nLay, const double *flux_up, const double *flux_dn,
                                    const double *p_lev, double *htr, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (nCol < 0 || nLay < 1) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "heating_rate: bad dimensions nCol=%d nLay=%d", nCol, nLay);
        return 1;
    }
    if (!flux_up || !flux_dn || !p_lev || !htr) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "heating_rate: NULL array");
        return 1;
    }
    if (nCol == 0) return 0;
    const size_t nv = (size_t)nCol * (nLay + 1), nl = (size_t)nCol * nLay;
    double *d = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    auto fail = [&](const char *what, hipError_t e) {
        if (errmsg && errlen > 0) snprintf(errmsg, errlen, "heating_rate: %s: %s", what, hipGetErrorString(e));
        rc = 2;
    };
    hipError_t e = hipStreamCreate(&st);
    if (e != hipSuccess) { fail("stream creation", e); return rc; }
    e = hipMalloc(&d, sizeof(double) * (3 * nv + nl));
    if (e != hipSuccess) { fail("device allocation", e); (void)hipStreamDestroy(st); return rc; }
    double *du = d, *dd = d + nv, *dp = d + 2 * nv, *dh = d + 3 * nv;
    if ((e = hipMemcpyAsync(du, flux_up, sizeof(double) * nv, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(dd, flux_dn, sizeof(double) * nv, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(dp, p_lev, sizeof(double) * nv, hipMemcpyHostToDevice, st)) != hipSuccess) {
        fail("host to device copy", e);
    } else {
        rc = ccpprad_heating_rate_dev(nCol, nLay, du, dd, dp, dh, (void *)st, errmsg, errlen);
        if (!rc && (e = hipMemcpyAsync(htr, dh, sizeof(double) * nl, hipMemcpyDeviceToHost, st)) != hipSuccess)
            fail("device to host copy", e);
        if (!rc && (e = hipStreamSynchronize(st)) != hipSuccess) fail("stream synchronisation", e);
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(d);
    (void)hipStreamDestroy(st);
    return rc;
}
