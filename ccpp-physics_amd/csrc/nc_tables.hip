// Host side: the RRTMGP k-distribution and cloud-optics netCDF files read
// natively, and reduced to the active gases the way ty_gas_optics_rrtmgp%load
// does, so *_init_files accept the files the reference's _init reads.
//
//  - Reader: the netCDF classic formats (CDF-1, CDF-2 64-bit offset, CDF-5
//    64-bit data), big-endian, fixed-size variables.  netCDF-4 files are HDF5
//    containers; they are recognised and refused with a message (this build
//    has no HDF5; `nccopy -k cdf5` converts them).  The file is untrusted
//    input: every count, offset and size is checked against the file length.
//  - Variables: those the reference reads at rrtmgp_lw_gas_optics.F90:130-327,
//    rrtmgp_sw_gas_optics.F90:288-376 and rrtmgp_lw_cloud_optics.F90:130-290
//    (LUT form), with the same names and dimension orders.
//  - Reduction (rte-rrtmgp's init_abs_coeffs, called at
//    rrtmgp_lw_gas_optics.F90:439-448 and rrtmgp_sw_gas_optics.F90:487-496):
//    the file's gases cut to the active ones; vmr_ref columns of the kept
//    gases; minor-absorber intervals whose gas is inactive dropped, their
//    kminor contributors packed; key species mapped to the kept gases (a key
//    species that is not active is an error); flavors = the distinct key
//    species pairs over (band, atmosphere layer) in that order, a (0, 0)
//    pair standing for (2, 2); each g-point's flavor per layer from its
//    band's pair.  Gas indices are then renumbered to the library's fixed
//    order h2o co2 o3 n2o ch4 o2.
//  - SW solar source: quiet + (mg - 0.1495954) facular + (sb - 0.00066696)
//    sunspot, with mg_default / sb_default from the file when present, else
//    the scheme's defaults (rrtmgp_sw_gas_optics.F90:21-23,320-337); its
//    scale does not matter, the run normalises the TOA source to solcon.
#include "kernels.h"
#include "../../include/ccpprad.h"
#include "nc_tables.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace ccpprad {
namespace {

// ---------------------------------------------------------------------------
// netCDF classic reader
// ---------------------------------------------------------------------------
enum { NC_BYTE = 1, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE, NC_UBYTE, NC_USHORT, NC_UINT, NC_INT64, NC_UINT64 };
enum { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

int type_size(int t)
{
    switch (t) {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
    }
}

struct NcVar {
    int type = 0;
    std::vector<size_t> shape;     // C order (slowest first)
    std::vector<std::string> dims;
    uint64_t begin = 0;
    bool record = false;
};

struct NcFile {
    std::vector<unsigned char> buf;
    std::map<std::string, size_t> dims;
    std::map<std::string, NcVar> vars;
};

struct Cursor {
    const unsigned char *p, *end;
    bool ok = true;
    bool need(size_t n) { if (ok && (size_t)(end - p) < n) ok = false; return ok; }
    uint64_t be(int n)
    {
        if (!need((size_t)n)) return 0;
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
        p += n;
        return v;
    }
    void skip(uint64_t n) { if (need(n)) p += n; }
};

bool nc_open(const char *path, NcFile &f, std::string &err)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) { err = std::string("cannot open ") + path; return false; }
    fseek(fp, 0, SEEK_END);
    const long len = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    if (len < 8) { fclose(fp); err = std::string("not a netCDF file: ") + path; return false; }
    f.buf.resize((size_t)len);
    const bool rd = fread(f.buf.data(), 1, (size_t)len, fp) == (size_t)len;
    fclose(fp);
    if (!rd) { err = std::string("cannot read ") + path; return false; }
    const unsigned char *b = f.buf.data();
    if (b[0] == 0x89 && b[1] == 'H' && b[2] == 'D' && b[3] == 'F') {
        err = std::string(path) + " is a netCDF-4 (HDF5) file; this build reads the netCDF classic formats: "
                                  "convert it with `nccopy -k cdf5`";
        return false;
    }
    if (!(b[0] == 'C' && b[1] == 'D' && b[2] == 'F' && (b[3] == 1 || b[3] == 2 || b[3] == 5))) {
        err = std::string("not a netCDF classic file: ") + path;
        return false;
    }
    const int ver = b[3];
    const int nsz = ver == 5 ? 8 : 4;            // counts and lengths
    const int osz = ver == 1 ? 4 : 8;            // data offsets
    Cursor c{b + 4, b + len};
    c.be(nsz);                                   // numrecs
    auto name = [&]() {
        const uint64_t n = c.be(nsz);
        std::string s;
        if (n > 4096 || !c.need(n)) { c.ok = false; return s; }
        s.assign((const char *)c.p, (size_t)n);
        c.skip((n + 3) & ~(uint64_t)3);
        return s;
    };
    auto skip_atts = [&]() {
        const uint64_t tag = c.be(4), n = c.be(nsz);
        if (tag == 0 && n == 0) return;
        if (tag != TAG_ATT || n > 100000) { c.ok = false; return; }
        for (uint64_t a = 0; a < n && c.ok; ++a) {
            name();
            const int t = (int)c.be(4);
            const uint64_t cnt = c.be(nsz);
            const int ts = type_size(t);
            if (!ts || cnt > (uint64_t)len) { c.ok = false; return; }
            c.skip((cnt * ts + 3) & ~(uint64_t)3);
        }
    };
    std::vector<std::pair<std::string, size_t>> dimv;
    {
        const uint64_t tag = c.be(4), n = c.be(nsz);
        if (!(tag == 0 && n == 0)) {
            if (tag != TAG_DIM || n > 100000) c.ok = false;
            for (uint64_t i = 0; i < n && c.ok; ++i) {
                std::string nm = name();
                const uint64_t dl = c.be(nsz);
                dimv.push_back({nm, (size_t)dl});
                f.dims[nm] = (size_t)dl;
            }
        }
    }
    skip_atts();                                 // global attributes
    {
        const uint64_t tag = c.be(4), n = c.be(nsz);
        if (!(tag == 0 && n == 0)) {
            if (tag != TAG_VAR || n > 100000) c.ok = false;
            for (uint64_t i = 0; i < n && c.ok; ++i) {
                NcVar v;
                std::string nm = name();
                const uint64_t nd = c.be(nsz);
                if (nd > 16) { c.ok = false; break; }
                for (uint64_t k = 0; k < nd && c.ok; ++k) {
                    const uint64_t id = c.be(nsz);
                    if (id >= dimv.size()) { c.ok = false; break; }
                    if (k == 0 && dimv[id].second == 0) v.record = true;   // the unlimited dimension
                    v.shape.push_back(dimv[id].second);
                    v.dims.push_back(dimv[id].first);
                }
                skip_atts();
                v.type = (int)c.be(4);
                c.be(nsz);                       // vsize (recomputed below)
                v.begin = c.be(osz);
                if (!c.ok) break;
                if (!type_size(v.type)) { c.ok = false; break; }
                if (!v.record) {
                    uint64_t nel = 1;
                    for (size_t d : v.shape) {
                        if (d != 0 && nel > ((uint64_t)1 << 40) / d) { c.ok = false; break; }
                        nel *= d;
                    }
                    const uint64_t bytes = nel * (uint64_t)type_size(v.type);
                    if (!c.ok || v.begin > (uint64_t)len || bytes > (uint64_t)len - v.begin) { c.ok = false; break; }
                }
                f.vars[nm] = v;
            }
        }
    }
    if (!c.ok) { err = std::string("malformed netCDF header in ") + path; return false; }
    return true;
}

size_t nel(const NcVar &v)
{
    size_t n = 1;
    for (size_t d : v.shape) n *= d;
    return n;
}

double elem_d(const unsigned char *p, int t)
{
    uint64_t u = 0;
    const int n = type_size(t);
    for (int i = 0; i < n; ++i) u = (u << 8) | p[i];
    switch (t) {
    case NC_BYTE: return (double)(int8_t)u;
    case NC_UBYTE: return (double)(uint8_t)u;
    case NC_SHORT: return (double)(int16_t)u;
    case NC_USHORT: return (double)(uint16_t)u;
    case NC_INT: return (double)(int32_t)u;
    case NC_UINT: return (double)(uint32_t)u;
    case NC_INT64: return (double)(int64_t)u;
    case NC_UINT64: return (double)u;
    case NC_FLOAT: { uint32_t w = (uint32_t)u; float x; memcpy(&x, &w, 4); return x; }
    case NC_DOUBLE: { double x; memcpy(&x, &u, 8); return x; }
    default: return 0.0;
    }
}

struct Reader {
    NcFile f;
    std::string missing;
    const NcVar *var(const char *n, size_t need)
    {
        auto it = f.vars.find(n);
        if (it == f.vars.end() || it->second.record || it->second.type == NC_CHAR || nel(it->second) < need) {
            if (missing.empty()) missing = n;
            return nullptr;
        }
        return &it->second;
    }
    bool has(const char *n) const { return f.vars.count(n) != 0; }
    size_t dim(const char *n)
    {
        auto it = f.dims.find(n);
        if (it == f.dims.end()) { if (missing.empty()) missing = std::string("dimension ") + n; return 0; }
        return it->second;
    }
    std::vector<double> dbl(const char *n, size_t need)
    {
        std::vector<double> out;
        const NcVar *v = var(n, need);
        if (!v) return out;
        const int ts = type_size(v->type);
        out.resize(nel(*v));
        const unsigned char *p = f.buf.data() + v->begin;
        for (size_t i = 0; i < out.size(); ++i) out[i] = elem_d(p + i * ts, v->type);
        return out;
    }
    std::vector<int> ints(const char *n, size_t need)
    {
        std::vector<double> d = dbl(n, need);
        std::vector<int> out(d.size());
        for (size_t i = 0; i < d.size(); ++i) out[i] = (int)d[i];
        return out;
    }
    double scalar(const char *n) { std::vector<double> d = dbl(n, 1); return d.empty() ? 0.0 : d[0]; }
    // NC_CHAR (n, strlen): n names, trailing blanks and NULs removed, lower case
    std::vector<std::string> names(const char *n, size_t need)
    {
        std::vector<std::string> out;
        auto it = f.vars.find(n);
        if (it == f.vars.end() || it->second.type != NC_CHAR || it->second.shape.size() != 2 ||
            it->second.shape[0] < need) {
            if (missing.empty()) missing = n;
            return out;
        }
        const NcVar &v = it->second;
        const size_t cnt = v.shape[0], sl = v.shape[1];
        const char *p = (const char *)f.buf.data() + v.begin;
        for (size_t i = 0; i < cnt; ++i) {
            std::string s(p + i * sl, sl);
            while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
            for (char &ch : s) ch = (char)tolower((unsigned char)ch);
            out.push_back(s);
        }
        return out;
    }
};

const char *const KNOWN[6] = {"h2o", "co2", "o3", "n2o", "ch4", "o2"};
int known_slot(const std::string &g)
{
    for (int i = 0; i < 6; ++i) if (g == KNOWN[i]) return i + 1;
    return 0;
}

int find(const std::vector<std::string> &v, const std::string &s)
{
    for (size_t i = 0; i < v.size(); ++i) if (v[i] == s) return (int)i;
    return -1;
}

// One regime's minor-absorber arrays after the reduction (rte's
// reduce_minor_arrays and create_idx_minor / create_idx_minor_scaling).
bool reduce_minor(Reader &R, const char *sfx, int neta, int ntemp, const std::vector<std::string> &active,
                  const std::vector<std::string> &kept, const std::vector<std::string> &gas_minor,
                  const std::vector<std::string> &identifier, NcMinorHold &m, std::string &err)
{
    const std::string s(sfx);
    const size_t nint = R.dim(("minor_absorber_intervals_" + s).c_str());
    const size_t ncon = R.dim(("contributors_" + s).c_str());
    const std::vector<std::string> gases = R.names(("minor_gases_" + s).c_str(), nint);
    const std::vector<std::string> scaling = R.names(("scaling_gas_" + s).c_str(), nint);
    const std::vector<int> lims = R.ints(("minor_limits_gpt_" + s).c_str(), 2 * nint);
    const std::vector<int> dens = R.ints(("minor_scales_with_density_" + s).c_str(), nint);
    const std::vector<int> comp = R.ints(("scale_by_complement_" + s).c_str(), nint);
    const std::vector<int> start = R.ints(("kminor_start_" + s).c_str(), nint);
    const std::vector<double> kmin = R.dbl(("kminor_" + s).c_str(), ncon * neta * ntemp);
    if (!R.missing.empty()) return false;
    m.limits.clear(); m.dens.clear(); m.scaling.clear(); m.comp.clear(); m.idx.clear(); m.start.clear();
    std::vector<std::pair<int, int>> keep;   // (file interval, width)
    for (size_t i = 0; i < nint; ++i) {
        const int id = find(identifier, gases[i]);
        if (id < 0) { err = "minor absorber '" + gases[i] + "' has no identifier_minor entry"; return false; }
        const std::string &g = gas_minor[id];
        if (find(active, g) < 0) continue;   // its gas is not active: interval dropped
        const int lo = lims[2 * i], hi = lims[2 * i + 1];
        if (hi < lo || start[i] < 1 || (size_t)(start[i] - 1 + hi - lo) >= ncon) {
            err = "minor interval " + std::to_string(i + 1) + " (" + s + ") has bad limits or start";
            return false;
        }
        keep.push_back({(int)i, hi - lo + 1});
        m.limits.push_back(lo);
        m.limits.push_back(hi);
        m.dens.push_back(dens[i] != 0);
        m.comp.push_back(comp[i] != 0);
        m.idx.push_back(known_slot(g));
        // the scaling gas is looked up among the kept gases: none (0) if absent
        m.scaling.push_back(find(kept, scaling[i]) >= 0 ? known_slot(scaling[i]) : 0);
    }
    int ncontrib = 0;
    for (auto &k : keep) ncontrib += k.second;
    m.kminor.assign((size_t)std::max(ncontrib, 1) * neta * ntemp, 0.0);
    // kminor (ncontrib, neta, ntemp) in Fortran order = C order (ntemp, neta, ncontrib)
    int ks = 1;
    for (auto &k : keep) {
        const int i = k.first, w = k.second, s0 = start[i] - 1;
        m.start.push_back(ks);
        for (int t = 0; t < ntemp; ++t)
            for (int e = 0; e < neta; ++e)
                for (int j = 0; j < w; ++j)
                    m.kminor[((size_t)t * neta + e) * ncontrib + (ks - 1 + j)] =
                        kmin[((size_t)t * neta + e) * ncon + (s0 + j)];
        ks += w;
    }
    m.m.n = (int)keep.size();
    m.m.ncontrib = ncontrib;
    m.m.limits_gpt = m.limits.data();
    m.m.scales_with_density = m.dens.data();
    m.m.idx_scaling = m.scaling.data();
    m.m.scale_by_complement = m.comp.data();
    m.m.idx_minor = m.idx.data();
    m.m.kminor_start = m.start.data();
    m.m.kminor = m.kminor.data();
    return true;
}

}  // namespace

bool is_netcdf_file(const char *path)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return false;
    unsigned char b[4] = {0, 0, 0, 0};
    const bool rd = fread(b, 1, 4, fp) == 4;
    fclose(fp);
    return rd && ((b[0] == 'C' && b[1] == 'D' && b[2] == 'F') || (b[0] == 0x89 && b[1] == 'H' && b[2] == 'D' && b[3] == 'F'));
}

bool load_kdist_nc(const char *path, int want_lw, const char *const *gas_names, int ngases, NcKdistHold &h,
                   std::string &err)
{
    Reader R;
    if (!nc_open(path, R.f, err)) return false;
    const size_t ntemp = R.dim("temperature"), npres = R.dim("pressure"), nabs = R.dim("absorber");
    const size_t neta = R.dim("mixing_fraction"), nbnd = R.dim("bnd"), ngpt = R.dim("gpt");
    const size_t nminor = R.dim("minor_absorber");
    if (!R.missing.empty()) { err = "netCDF k-distribution lacks " + R.missing; return false; }
    if (ntemp < 2 || npres < 2 || neta < 2 || nbnd < 1 || ngpt < 1 || nabs < 1 || nabs > 64) {
        err = "netCDF k-distribution has bad dimensions";
        return false;
    }
    const bool lw = R.has("plank_fraction") || R.has("totplnk");
    if (lw != !!want_lw) { err = "k-distribution file is for the other spectral range"; return false; }
    // active gases: the named ones the file holds (rte keeps the file's order)
    std::vector<std::string> active;
    for (int i = 0; i < ngases && gas_names; ++i) {
        if (!gas_names[i]) continue;
        std::string g(gas_names[i]);
        for (char &ch : g) ch = (char)tolower((unsigned char)ch);
        active.push_back(g);
    }
    const std::vector<std::string> fgas = R.names("gas_names", nabs);
    const std::vector<std::string> gas_minor = R.names("gas_minor", nminor);
    const std::vector<std::string> identifier = R.names("identifier_minor", nminor);
    const std::vector<int> key = R.ints("key_species", 2 * 2 * nbnd);          // (bnd, atmos_layer, pair)
    const std::vector<int> b2g = R.ints("bnd_limits_gpt", 2 * nbnd);
    h.band_wvn = R.dbl("bnd_limits_wavenumber", 2 * nbnd);
    h.press_ref = R.dbl("press_ref", npres);
    h.temp_ref = R.dbl("temp_ref", ntemp);
    const double ptrop = R.scalar("press_ref_trop");
    const std::vector<double> vref = R.dbl("vmr_ref", ntemp * (nabs + 1) * 2);  // (temperature, absorber_ext, atmos_layer)
    h.kmajor = R.dbl("kmajor", ntemp * (npres + 1) * neta * ngpt);
    if (!R.missing.empty()) { err = "netCDF k-distribution lacks " + R.missing; return false; }
    std::vector<std::string> kept;   // this%gas_names: file order, active only
    for (const std::string &g : fgas) if (find(active, g) >= 0) kept.push_back(g);
    for (const std::string &g : kept)
        if (!known_slot(g)) { err = "active gas '" + g + "' has no concentration input in the run arguments"; return false; }
    // key species (file gas indices, 1-based, 0 = none) -> library slots
    auto slot_of_file = [&](int fi, int &slot) {
        if (fi == 0) { slot = 0; return true; }
        if (fi < 1 || (size_t)fi > fgas.size()) { err = "key_species index out of range"; return false; }
        const std::string &g = fgas[fi - 1];
        if (find(active, g) < 0) { err = "key species '" + g + "' is not an active gas"; return false; }
        slot = known_slot(g);
        return true;
    };
    // the (0, 0) pair stands for (2, 2) in the reduced gas list
    int two = 0;
    if (kept.size() >= 2) two = known_slot(kept[1]);
    std::vector<std::pair<int, int>> flav;
    std::vector<int> bflav(2 * nbnd);
    for (size_t b = 0; b < nbnd; ++b)
        for (int a = 0; a < 2; ++a) {
            int s0 = 0, s1 = 0;
            if (!slot_of_file(key[(b * 2 + a) * 2 + 0], s0) || !slot_of_file(key[(b * 2 + a) * 2 + 1], s1)) return false;
            if (s0 == 0 && s1 == 0) {
                if (!two) { err = "a band without key species needs at least two active gases"; return false; }
                s0 = s1 = two;
            }
            if (s0 == 0 || s1 == 0) { err = "key species pair with one gas missing"; return false; }
            int f = -1;
            for (size_t i = 0; i < flav.size(); ++i) if (flav[i].first == s0 && flav[i].second == s1) f = (int)i;
            if (f < 0) { flav.push_back({s0, s1}); f = (int)flav.size() - 1; }
            bflav[2 * b + a] = f + 1;
        }
    h.flavor.clear();
    for (auto &p : flav) { h.flavor.push_back(p.first); h.flavor.push_back(p.second); }
    h.band_gpt.assign(b2g.begin(), b2g.begin() + 2 * nbnd);
    h.gpt_flav.assign(2 * ngpt, 0);
    for (size_t b = 0; b < nbnd; ++b) {
        const int g0 = b2g[2 * b], g1 = b2g[2 * b + 1];
        if (g0 < 1 || g1 < g0 || (size_t)g1 > ngpt) { err = "bnd_limits_gpt out of range"; return false; }
        for (int g = g0; g <= g1; ++g) {
            h.gpt_flav[2 * (g - 1)] = bflav[2 * b];
            h.gpt_flav[2 * (g - 1) + 1] = bflav[2 * b + 1];
        }
    }
    // vmr_ref (2, 0:ngas, ntemp) in the library's gas order; gases that are
    // not kept get 1 (no flavor or minor interval refers to them)
    const int NG = 6;
    h.vmr_ref.assign(2 * (size_t)(NG + 1) * ntemp, 1.0);
    for (size_t t = 0; t < ntemp; ++t)
        for (int a = 0; a < 2; ++a) {
            h.vmr_ref[(t * (NG + 1) + 0) * 2 + a] = vref[(t * (nabs + 1) + 0) * 2 + a];
            for (const std::string &g : kept) {
                const int fi = find(fgas, g);
                h.vmr_ref[(t * (NG + 1) + known_slot(g)) * 2 + a] = vref[(t * (nabs + 1) + fi + 1) * 2 + a];
            }
        }
    if (!reduce_minor(R, "lower", (int)neta, (int)ntemp, active, kept, gas_minor, identifier, h.lower, err) ||
        !reduce_minor(R, "upper", (int)neta, (int)ntemp, active, kept, gas_minor, identifier, h.upper, err)) {
        if (err.empty()) err = "netCDF k-distribution lacks " + R.missing;
        return false;
    }
    ccpprad_kdist &k = h.kd;
    k = ccpprad_kdist{};
    k.is_lw = want_lw;
    k.ngas = NG; k.nband = (int)nbnd; k.ngpt = (int)ngpt; k.ntemp = (int)ntemp; k.npres = (int)npres;
    k.neta = (int)neta; k.nflav = (int)flav.size();
    k.band_lims_gpt = h.band_gpt.data();
    k.band_lims_wvn = h.band_wvn.data();
    k.flavor = h.flavor.data();
    k.gpoint_flavor = h.gpt_flav.data();
    k.press_ref = h.press_ref.data();
    k.temp_ref = h.temp_ref.data();
    k.press_ref_trop = ptrop;
    k.vmr_ref = h.vmr_ref.data();
    k.kmajor = h.kmajor.data();
    k.minor_lower = h.lower.m;
    k.minor_upper = h.upper.m;
    if (want_lw) {
        const size_t npl = R.dim("temperature_Planck");
        h.planck = R.dbl("plank_fraction", ntemp * (npres + 1) * neta * ngpt);
        h.totplnk = R.dbl("totplnk", npl * nbnd);
        if (R.has("optimal_angle_fit")) h.fit = R.dbl("optimal_angle_fit", 2 * nbnd);
        if (!R.missing.empty()) { err = "netCDF k-distribution lacks " + R.missing; return false; }
        k.nplanck = (int)npl;
        k.planck_frac = h.planck.data();
        k.totplnk = h.totplnk.data();
        k.optimal_angle_fit = h.fit.empty() ? nullptr : h.fit.data();
    } else {
        h.rayl_lower = R.dbl("rayl_lower", ntemp * neta * ngpt);
        h.rayl_upper = R.dbl("rayl_upper", ntemp * neta * ngpt);
        const std::vector<double> q = R.dbl("solar_source_quiet", ngpt);
        const std::vector<double> fa = R.dbl("solar_source_facular", ngpt);
        const std::vector<double> ss = R.dbl("solar_source_sunspot", ngpt);
        if (!R.missing.empty()) { err = "netCDF k-distribution lacks " + R.missing; return false; }
        const double mg = R.has("mg_default") ? R.scalar("mg_default") : 0.1567652;
        const double sb = R.has("sb_default") ? R.scalar("sb_default") : 902.7126;
        h.solar.resize(ngpt);
        for (size_t g = 0; g < ngpt; ++g) h.solar[g] = q[g] + (mg - 0.1495954) * fa[g] + (sb - 0.00066696) * ss[g];
        k.nplanck = 0;
        k.krayl_lower = h.rayl_lower.data();
        k.krayl_upper = h.rayl_upper.data();
        k.solar_src = h.solar.data();
    }
    return true;
}

bool load_lut_nc(const char *path, NcLutHold &h, std::string &err, bool force_pade)
{
    Reader R;
    if (!nc_open(path, R.f, err)) return false;
    const size_t nb = R.dim("nband"), nr = R.dim("nrghice");
    const bool lut = R.has("lut_extliq") && !force_pade;
    const size_t nl = lut ? R.dim("nsize_liq") : 0, ni = lut ? R.dim("nsize_ice") : 0;
    if (!R.missing.empty()) { err = "netCDF cloud-optics file lacks " + R.missing; return false; }
    if (!lut) {
        // Pade form (rrtmgp_lw_cloud_optics.F90:192-205,261-288): coefficients
        // (nband, nsizereg, ncoeff[, nrghice]) and the size-regime bounds
        const size_t ns = R.dim("nsizereg"), ne = R.dim("ncoeff_ext"), nsg = R.dim("ncoeff_ssa_g");
        const size_t nbd = R.dim("nbound");
        if (!R.missing.empty()) { err = "netCDF cloud-optics file has neither lookup tables nor Pade " + R.missing; return false; }
        if (nb < 1 || nr < 1 || ns != 3 || ne != 6 || nsg != 5 || nbd != 4) {
            err = "Pade cloud optics need 3 size regimes, 6 extinction and 5 ssa/g coefficients, 4 bounds";
            return false;
        }
        const char *cn[6] = {"pade_extliq", "pade_ssaliq", "pade_asyliq", "pade_extice", "pade_ssaice", "pade_asyice"};
        const char *bn[6] = {"pade_sizreg_extliq", "pade_sizreg_ssaliq", "pade_sizreg_asyliq",
                             "pade_sizreg_extice", "pade_sizreg_ssaice", "pade_sizreg_asyice"};
        for (int k = 0; k < 6; ++k) {
            h.pade[k] = R.dbl(cn[k], nb * ns * (k % 3 == 0 ? ne : nsg) * (k >= 3 ? nr : 1));
            h.sizreg[k] = R.dbl(bn[k], nbd);
        }
        if (!R.missing.empty()) { err = "netCDF cloud-optics file lacks " + R.missing; return false; }
        ccpprad_cloud_lut &L = h.lut;
        L = ccpprad_cloud_lut{};
        L.nband = (int)nb; L.nrghice = (int)nr; L.icergh = (int)nr;
        L.radliq_lwr = h.sizreg[0][0]; L.radliq_upr = h.sizreg[0][3];
        L.radice_lwr = h.sizreg[3][0]; L.radice_upr = h.sizreg[3][3];
        L.pade = 1;
        L.pade_extliq = h.pade[0].data(); L.pade_ssaliq = h.pade[1].data(); L.pade_asyliq = h.pade[2].data();
        L.pade_extice = h.pade[3].data(); L.pade_ssaice = h.pade[4].data(); L.pade_asyice = h.pade[5].data();
        L.pade_sizreg_extliq = h.sizreg[0].data(); L.pade_sizreg_ssaliq = h.sizreg[1].data();
        L.pade_sizreg_asyliq = h.sizreg[2].data(); L.pade_sizreg_extice = h.sizreg[3].data();
        L.pade_sizreg_ssaice = h.sizreg[4].data(); L.pade_sizreg_asyice = h.sizreg[5].data();
        return true;
    }
    if (nb < 1 || nl < 2 || ni < 2 || nr < 1) { err = "netCDF cloud-optics file has bad dimensions"; return false; }
    h.extliq = R.dbl("lut_extliq", nl * nb); h.ssaliq = R.dbl("lut_ssaliq", nl * nb); h.asyliq = R.dbl("lut_asyliq", nl * nb);
    h.extice = R.dbl("lut_extice", ni * nb * nr); h.ssaice = R.dbl("lut_ssaice", ni * nb * nr);
    h.asyice = R.dbl("lut_asyice", ni * nb * nr);
    ccpprad_cloud_lut &L = h.lut;
    L = ccpprad_cloud_lut{};
    L.radliq_lwr = R.scalar("radliq_lwr"); L.radliq_upr = R.scalar("radliq_upr");
    L.radice_lwr = R.scalar("radice_lwr"); L.radice_upr = R.scalar("radice_upr");
    if (!R.missing.empty()) { err = "netCDF cloud-optics file lacks " + R.missing; return false; }
    L.nband = (int)nb; L.nsize_liq = (int)nl; L.nsize_ice = (int)ni; L.nrghice = (int)nr;
    L.icergh = (int)nr;   // the last roughness category (namelist nrghice = the file's count)
    L.extliq = h.extliq.data(); L.ssaliq = h.ssaliq.data(); L.asyliq = h.asyliq.data();
    L.extice = h.extice.data(); L.ssaice = h.ssaice.data(); L.asyice = h.asyice.data();
    return true;
}

}  // namespace ccpprad
