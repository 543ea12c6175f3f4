// C-ABI of the MI355X RRTMGP hot path: context (device tables), scheme
// run entries (device and host pointers), component entries.
// See include/ccpprad.h for the reference interface each entry replaces.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>
#include <initializer_list>
#include <thread>

#include "../../include/ccpprad.h"
#include "ccpprad_dev.h"

#include "kernels.h"

using namespace ccpprad;

// --------------------------------------------------------------------------
// helpers
// --------------------------------------------------------------------------
static void seterr(char *errmsg, int errlen, const char *fmt, ...)
{
    if (!errmsg || errlen <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(errmsg, (size_t)errlen, fmt, ap);
    va_end(ap);
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            seterr(errmsg, errlen, "%s failed: %s", #expr, hipGetErrorString(e_));          \
            return 2;                                                                       \
        }                                                                                   \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Per-call mutable state.  One workspace serves one stream: calls on that
// stream reuse it in stream order, calls on other streams get their own, so
// concurrent scheme calls from several host threads do not share buffers.
struct Workspace {
    hipStream_t stream = nullptr;  // the stream it serves
    bool own_stream = false;       // created for a host thread's staging path
    std::thread::id owner;         // that thread
    DevBuf w_tau, w_aux, w_mask, w_ctau, w_sb, w_blk, w_colf, w_day, w_idx;
    DevBuf w_oob, w_colm, w_aer;
    std::vector<DevBuf> stage;     // host-path staging
    int *err = nullptr;            // device error word, then the SW last-wipe block
    double *toa64 = nullptr;       // SW TOA source scaled to solcon
    float *toa32 = nullptr;
    void release()
    {
        DevBuf *bufs[] = {&w_tau, &w_aux, &w_mask, &w_ctau, &w_sb, &w_blk, &w_colf, &w_day, &w_idx,
                          &w_oob, &w_colm, &w_aer};
        for (DevBuf *b : bufs) b->release();
        for (auto &b : stage) b.release();
        if (err) (void)hipFree(err);
        if (toa64) (void)hipFree(toa64);
        if (toa32) (void)hipFree(toa32);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

struct ccpprad_ctx {
    int is_lw = 0;
    int nlay = 0, blksz = 0, device = 0;
    TablesView T{};
    LutView L{};
    int have_lut = 0;
    std::vector<void *> allocs;         // table allocations
    uint32_t *mtpow = nullptr;
    int *alb_cat = nullptr;
    int ibd = 0, ibd_uv = 0;
    const double *solar_dev = nullptr;
    int chunk = 65536;                  // columns per launch
    int chunk_max = 65536;              // its bound (32-bit cloud-record offsets)
    // learned gas optics (LW): MLP with table fallback, or null = table
    ccpprad_mlp *mlp = nullptr;
    // workspaces and the lock for them and for the timing records; held while
    // a call enqueues its work, not while the GPU runs it
    std::mutex mu;
    std::vector<Workspace *> ws;
    // optional per-kernel event timing
    int timing = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_rec;
};

// The workspace of a stream (caller holds c->mu).
static Workspace *ws_for_stream(ccpprad_ctx *c, hipStream_t st)
{
    for (Workspace *w : c->ws)
        if (w->stream == st) return w;
    Workspace *w = new Workspace();
    w->stream = st;
    const int G = c->T.ngpt > 0 ? c->T.ngpt : 1;
    if (hipMalloc(&w->err, 2 * sizeof(int)) != hipSuccess ||
        hipMalloc(&w->toa64, sizeof(double) * G) != hipSuccess ||
        hipMalloc(&w->toa32, sizeof(float) * G) != hipSuccess) {
        w->release();
        delete w;
        return nullptr;
    }
    c->ws.push_back(w);
    return w;
}

enum { K_GAS = 0, K_LWSOLVE = 1, K_SWSOLVE = 2, K_CLOUD = 3, K_MCICA = 4, K_MLP = 5, K_GASFB = 6, K_AER = 7, K_N = 8 };

static hipEvent_t ev_take(ccpprad_ctx *c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Brackets one launch with events on its stream when timing is on.
struct KScope {
    ccpprad_ctx *c;
    int id;
    hipStream_t st;
    hipEvent_t a = nullptr;
    KScope(ccpprad_ctx *c_, int id_, hipStream_t st_) : c(c_), id(id_), st(st_)
    {
        if (c->timing && (a = ev_take(c))) (void)hipEventRecord(a, st);
    }
    ~KScope()
    {
        if (!a) return;
        hipEvent_t b = ev_take(c);
        if (!b) return;
        (void)hipEventRecord(b, st);
        c->ev_rec.push_back({id, {a, b}});
    }
};

extern "C" int ccpprad_timing_enable(ccpprad_ctx *c, int enable)
{
    if (!c) return 1;
    c->timing = enable;
    c->ev_rec.clear();
    c->ev_used = 0;
    return 0;
}

extern "C" int ccpprad_timing_query(ccpprad_ctx *c, int kernel_id, double *total_ms, int *count)
{
    if (!c || !total_ms || !count) return 1;
    double t = 0.0;
    int n = 0;
    for (auto &r : c->ev_rec) {
        if (r.first != kernel_id) continue;
        if (hipEventSynchronize(r.second.second) != hipSuccess) return 2;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.second.first, r.second.second) != hipSuccess) return 2;
        t += ms;
        ++n;
    }
    *total_ms = t;
    *count = n;
    return 0;
}

// Internal to the library (hidden): the calling thread's workspace for the
// host-pointer entries (a private stream per thread), and its staging buffers.
extern "C" __attribute__((visibility("hidden"))) void *ccpprad_ctx_host_stream(ccpprad_ctx *c)
{
    std::lock_guard<std::mutex> lk(c->mu);
    const std::thread::id me = std::this_thread::get_id();
    for (Workspace *w : c->ws)
        if (w->own_stream && w->owner == me) return (void *)w->stream;
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess) return nullptr;
    Workspace *w = ws_for_stream(c, st);
    if (!w) { (void)hipStreamDestroy(st); return nullptr; }
    w->own_stream = true;
    w->owner = me;
    return (void *)st;
}

extern "C" __attribute__((visibility("hidden"))) int ccpprad_ctx_stage(ccpprad_ctx *c, void *stream, int slot,
                                                                       size_t bytes, void **p)
{
    std::lock_guard<std::mutex> lk(c->mu);
    Workspace *w = ws_for_stream(c, (hipStream_t)stream);
    if (!w) return 1;
    if ((int)w->stage.size() <= slot) w->stage.resize(slot + 1);
    if (w->stage[slot].ensure(bytes ? bytes : 8) != hipSuccess) return 1;
    *p = w->stage[slot].p;
    return 0;
}

template <typename T>
static T *upload(ccpprad_ctx *c, const T *h, size_t n, hipError_t *e)
{
    T *d = nullptr;
    if (*e != hipSuccess) return nullptr;   // keep the first error
    *e = hipMalloc(&d, sizeof(T) * (n ? n : 1));
    if (*e != hipSuccess) return nullptr;
    c->allocs.push_back(d);
    if (n && h) *e = hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice);
    return d;
}

static std::vector<float> to_f32(const double *p, size_t n)
{
    std::vector<float> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (float)p[i];
    return v;
}

static const char *KGAS[MAX_GAS] = {"h2o", "co2", "o3", "n2o", "ch4", "o2"};
static int scratch_setup(int dev, char *errmsg, int errlen);

static int ctx_init(int want_lw, const ccpprad_kdist *kd, const ccpprad_cloud_lut *lut,
                    const char *const *gas_names, int ngases, int nLay, int blksz, ccpprad_ctx **out,
                    char *errmsg, int errlen)
{
    if (!kd || !out) { seterr(errmsg, errlen, "null k-distribution or context pointer"); return 1; }
    if (!!kd->is_lw != !!want_lw) { seterr(errmsg, errlen, "k-distribution is for the other spectral range"); return 1; }
    if (kd->ngpt > MAX_GPT || kd->ngpt < 1) { seterr(errmsg, errlen, "ngpt=%d outside 1..%d", kd->ngpt, MAX_GPT); return 1; }
    if (kd->nband > MAX_BAND || kd->nflav > MAX_FLAV || kd->ngas > MAX_GAS || kd->ngas < 1) {
        seterr(errmsg, errlen, "table dimensions exceed the device limits"); return 1;
    }
    if (kd->minor_lower.n > MAX_MINOR || kd->minor_upper.n > MAX_MINOR) { seterr(errmsg, errlen, "too many minor intervals"); return 1; }
    if (nLay < 2 || nLay > MAX_LAY) { seterr(errmsg, errlen, "nLay=%d outside 2..%d", nLay, MAX_LAY); return 1; }
    if (blksz < 1) { seterr(errmsg, errlen, "rrtmgp_phys_blksz must be positive"); return 1; }
    if ((want_lw ? std::max(lw_solve_smem(nLay, kd->nband), lw_solve_2str_smem(nLay, kd->nband))
                 : sw_solve_smem(nLay, kd->nband)) > 64 * 1024) {
        seterr(errmsg, errlen, "nLay=%d needs more than the solver's 64 KiB of LDS per column", nLay); return 1;
    }
    if (gas_gather_smem(want_lw, kd->nflav, kd->minor_lower.n, kd->minor_upper.n, kd->ngpt, kd->neta) > 64 * 1024) {
        seterr(errmsg, errlen, "table dimensions need more than the gas optics' 64 KiB of LDS per block"); return 1;
    }
    if (want_lw && (!kd->planck_frac || !kd->totplnk || kd->nplanck < 2)) { seterr(errmsg, errlen, "LW tables lack Planck data"); return 1; }
    if (!want_lw && (!kd->krayl_lower || !kd->krayl_upper || !kd->solar_src)) { seterr(errmsg, errlen, "SW tables lack Rayleigh/solar data"); return 1; }
    // active gas names (CCPP active_gases_array) must be known to the k-distribution
    for (int i = 0; i < ngases && gas_names; ++i) {
        bool ok = false;
        for (int g = 0; g < kd->ngas && g < MAX_GAS; ++g)
            if (gas_names[i] && strcmp(gas_names[i], KGAS[g]) == 0) ok = true;
        if (!ok) { seterr(errmsg, errlen, "gas '%s' is not in the k-distribution", gas_names[i] ? gas_names[i] : "(null)"); return 1; }
    }
    ccpprad_ctx *c = new ccpprad_ctx();
    c->is_lw = want_lw;
    c->nlay = nLay;
    c->blksz = blksz;
    (void)hipGetDevice(&c->device);
    // the solvers form byte offsets into a launch's cloud records
    // ([layer][column][band], 12 bytes each) in 32 bits: bound the columns
    // per launch accordingly
    {
        const size_t per_col = (size_t)std::max(nLay, 1) * std::max(kd->nband, 1) * 12;
        const size_t cmax = (((size_t)1 << 32) - 1) / per_col;
        c->chunk_max = (int)std::min<size_t>(std::max<size_t>(cmax, 1), (size_t)1 << 30);
        c->chunk = std::min(c->chunk, c->chunk_max);
    }
    hipError_t e = hipSuccess;
    TablesView &T = c->T;
    T.maxminor_g = 0;
    T.is_lw = want_lw; T.ngas = kd->ngas; T.nband = kd->nband; T.ngpt = kd->ngpt; T.ntemp = kd->ntemp;
    T.npres = kd->npres; T.neta = kd->neta; T.nflav = kd->nflav; T.nplanck = kd->nplanck;
    T.press_ref_log1 = log(kd->press_ref[0]);
    T.press_ref_log_delta = (log(kd->press_ref[0]) - log(kd->press_ref[kd->npres - 1])) / (kd->npres - 1);
    T.press_ref_trop_log = log(kd->press_ref_trop);
    T.temp_ref_min = kd->temp_ref[0];
    T.temp_ref_delta = (kd->temp_ref[kd->ntemp - 1] - kd->temp_ref[0]) / (kd->ntemp - 1);
    T.totplnk_delta = want_lw ? (kd->temp_ref[kd->ntemp - 1] - kd->temp_ref[0]) / (kd->nplanck - 1) : 1.0;
    T.press_min = std::min(kd->press_ref[0], kd->press_ref[kd->npres - 1]);
    T.press_max = std::max(kd->press_ref[0], kd->press_ref[kd->npres - 1]);
    T.temp_min = kd->temp_ref[0];
    T.temp_max = kd->temp_ref[kd->ntemp - 1];
    T.temp_ref = upload(c, kd->temp_ref, kd->ntemp, &e);
    const size_t nk = (size_t)kd->ngpt * kd->neta * (kd->npres + 1) * kd->ntemp;
    // eta pairs: element (.., eta, g) with its eta + 1 neighbour (0 past the last row)
    auto eta_pair = [&](const double *t, size_t i, int stride) -> std::pair<float, float> {
        const int e = (int)((i / (size_t)stride) % (size_t)kd->neta);
        return {(float)t[i], e + 1 < kd->neta ? (float)t[i + stride] : 0.0f};
    };
    if (want_lw) {
        std::vector<float4> q(nk);
        for (size_t i = 0; i < nk; ++i) {
            const auto k = eta_pair(kd->kmajor, i, kd->ngpt), f = eta_pair(kd->planck_frac, i, kd->ngpt);
            q[i] = make_float4(k.first, k.second, f.first, f.second);
        }
        T.kq = upload(c, q.data(), nk, &e);
        auto p = to_f32(kd->totplnk, (size_t)kd->nplanck * kd->nband); T.totplnk = upload(c, p.data(), p.size(), &e);
        std::vector<float> fit(2 * kd->nband, 0.0f);
        if (kd->optimal_angle_fit) for (int i = 0; i < 2 * kd->nband; ++i) fit[i] = (float)kd->optimal_angle_fit[i];
        else for (int b = 0; b < kd->nband; ++b) fit[2 * b + 1] = 1.66f;
        T.opt_fit = upload(c, fit.data(), fit.size(), &e);
    } else {
        { auto v = to_f32(kd->kmajor, nk); T.kmaj = upload(c, v.data(), nk, &e); }
        const size_t nr = (size_t)kd->ngpt * kd->neta * kd->ntemp;
        auto a = to_f32(kd->krayl_lower, nr); T.kray[0] = upload(c, a.data(), nr, &e);
        auto b = to_f32(kd->krayl_upper, nr); T.kray[1] = upload(c, b.data(), nr, &e);
        c->solar_dev = upload(c, kd->solar_src, kd->ngpt, &e);
    }
    T.vmr_ref = upload(c, kd->vmr_ref, (size_t)2 * (kd->ngas + 1) * kd->ntemp, &e);
    std::vector<int> flav(2 * kd->nflav);
    for (int i = 0; i < 2 * kd->nflav; ++i) flav[i] = kd->flavor[i];
    T.flavor = upload(c, flav.data(), flav.size(), &e);
    {
        // vmr_ref ratio of each flavor's gas pair per (temperature, regime),
        // the same FP64 division the interpolation would do per layer
        const int ng1 = kd->ngas + 1;
        std::vector<double> ratio((size_t)kd->ntemp * kd->nflav * 2);
        for (int jt = 0; jt < kd->ntemp; ++jt)
            for (int f = 0; f < kd->nflav; ++f)
                for (int r = 0; r < 2; ++r) {
                    const int g1 = kd->flavor[2 * f], g2 = kd->flavor[2 * f + 1];
                    ratio[((size_t)jt * kd->nflav + f) * 2 + r] =
                        kd->vmr_ref[((size_t)jt * ng1 + g1) * 2 + r] / kd->vmr_ref[((size_t)jt * ng1 + g2) * 2 + r];
                }
        T.flav_ratio = upload(c, ratio.data(), ratio.size(), &e);
    }
    std::vector<int> gf(2 * kd->ngpt), g2b(kd->ngpt, 0), bl(2 * kd->nband);
    for (int g = 0; g < kd->ngpt; ++g) { gf[2 * g] = kd->gpoint_flavor[2 * g] - 1; gf[2 * g + 1] = kd->gpoint_flavor[2 * g + 1] - 1; }
    for (int b = 0; b < kd->nband; ++b) {
        bl[2 * b] = kd->band_lims_gpt[2 * b] - 1; bl[2 * b + 1] = kd->band_lims_gpt[2 * b + 1] - 1;
        for (int g = bl[2 * b]; g <= bl[2 * b + 1] && g < kd->ngpt; ++g) if (g >= 0) g2b[g] = b;
    }
    T.gpt_flav = upload(c, gf.data(), gf.size(), &e);
    T.g2b = upload(c, g2b.data(), g2b.size(), &e);
    T.band_lims = upload(c, bl.data(), bl.size(), &e);
    // minor absorbers: parameters and per-g CSR lists, per regime
    for (int r = 0; r < 2; ++r) {
        const ccpprad_minor &m = r == 0 ? kd->minor_lower : kd->minor_upper;
        T.nminor[r] = m.n;
        T.ncontrib[r] = m.ncontrib;
        std::vector<int> par(5 * std::max(1, m.n), 0);
        std::vector<std::vector<std::pair<int, int>>> per(kd->ngpt);
        for (int i = 0; i < m.n; ++i) {
            const int gs = m.limits_gpt[2 * i] - 1, ge = m.limits_gpt[2 * i + 1] - 1;
            if (gs < 0 || ge >= kd->ngpt || ge < gs || m.idx_minor[i] < 0 || m.idx_minor[i] > kd->ngas ||
                m.idx_scaling[i] < 0 || m.idx_scaling[i] > kd->ngas ||
                m.kminor_start[i] - 1 + (ge - gs) >= m.ncontrib) {
                delete c;
                seterr(errmsg, errlen, "invalid minor-absorber interval %d", i);
                return 1;
            }
            par[5 * i] = m.idx_minor[i];
            par[5 * i + 1] = m.scales_with_density[i];
            par[5 * i + 2] = m.idx_scaling[i];
            par[5 * i + 3] = m.scale_by_complement[i];
            par[5 * i + 4] = kd->gpoint_flavor[2 * gs + r] - 1;
            for (int g = gs; g <= ge; ++g) per[g].push_back({i, m.kminor_start[i] - 1 + (g - gs)});
        }
        std::vector<int> ptr(kd->ngpt + 1, 0), items;
        for (int g = 0; g < kd->ngpt; ++g) {
            T.maxminor_g = std::max(T.maxminor_g, (int)per[g].size());
            if ((int)per[g].size() > MAX_MINOR_G) {
                delete c;
                seterr(errmsg, errlen, "g-point %d has %d minor absorbers (limit %d)", g + 1, (int)per[g].size(), MAX_MINOR_G);
                return 1;
            }
            ptr[g + 1] = ptr[g] + (int)per[g].size();
            for (auto &p : per[g]) { items.push_back(p.first); items.push_back(p.second); }
        }
        if (items.empty()) items.push_back(0);
        T.minor_par[r] = upload(c, par.data(), par.size(), &e);
        T.minor_gptr[r] = upload(c, ptr.data(), ptr.size(), &e);
        T.minor_items[r] = upload(c, items.data(), items.size(), &e);
        const size_t nkm = (size_t)std::max(1, m.ncontrib) * kd->neta * kd->ntemp;
        if (want_lw) {
            std::vector<float2> km(nkm, make_float2(0.0f, 0.0f));
            if (m.ncontrib)
                for (size_t i = 0; i < nkm; ++i) { const auto p = eta_pair(m.kminor, i, m.ncontrib); km[i] = make_float2(p.first, p.second); }
            T.km2[r] = upload(c, km.data(), km.size(), &e);
        } else {
            std::vector<float> km(nkm, 0.0f);
            if (m.ncontrib) for (size_t i = 0; i < nkm; ++i) km[i] = (float)m.kminor[i];
            T.kmin[r] = upload(c, km.data(), km.size(), &e);
        }
    }
    // McICA: 69069^i mod 2^32
    {
        std::vector<uint32_t> pw(624);
        uint32_t x = 1u;
        for (int i = 0; i < 624; ++i) { pw[i] = x; x *= 69069u; }
        c->mtpow = upload(c, pw.data(), pw.size(), &e);
    }
    // SW albedo band categories (rrtmgp_sw_main.F90:368-387)
    {
        std::vector<int> cat(kd->nband, 3);
        for (int b = 0; b < kd->nband; ++b) {
            const double lo = kd->band_lims_wvn[2 * b];
            if (lo < 12850.0) cat[b] = 0;
            if (lo == 12850.0) { cat[b] = 1; c->ibd = b + 1; }
            if (lo >= 16000.0) cat[b] = 2;
            if (lo == 29000.0) c->ibd_uv = b + 1;
        }
        c->alb_cat = upload(c, cat.data(), cat.size(), &e);
        if (!want_lw && (!c->ibd || !c->ibd_uv)) {
            delete c;
            seterr(errmsg, errlen, "SW band limits lack the 12850 / 29000 cm-1 bands");
            return 1;
        }
    }
    if (lut) {
        if (lut->nband != kd->nband) { delete c; seterr(errmsg, errlen, "cloud LUT band count differs from the k-distribution"); return 1; }
        LutView &L = c->L;
        int icergh = lut->icergh > 0 ? lut->icergh : 1;
        if (icergh > lut->nrghice) { delete c; seterr(errmsg, errlen, "ice roughness category out of range"); return 1; }
        L.pade = lut->pade != 0;
        if (L.pade) {
            // Pade form: coefficients [phase][band][regime][16] and the bounds
            const double *P[2][3] = {{lut->pade_extliq, lut->pade_ssaliq, lut->pade_asyliq},
                                     {lut->pade_extice, lut->pade_ssaice, lut->pade_asyice}};
            const double *B[2][3] = {{lut->pade_sizreg_extliq, lut->pade_sizreg_ssaliq, lut->pade_sizreg_asyliq},
                                     {lut->pade_sizreg_extice, lut->pade_sizreg_ssaice, lut->pade_sizreg_asyice}};
            for (int ph = 0; ph < 2; ++ph)
                for (int f = 0; f < 3; ++f)
                    if (!P[ph][f] || !B[ph][f]) { delete c; seterr(errmsg, errlen, "Pade cloud optics: NULL array"); return 1; }
            const int nb = lut->nband;
            const int nco[3] = {6, 5, 5}, off[3] = {0, 6, 11};
            std::vector<float> pc((size_t)2 * nb * 48, 0.0f);
            for (int ph = 0; ph < 2; ++ph)
                for (int f = 0; f < 3; ++f) {
                    const double *src = P[ph][f] + (ph ? (size_t)nb * 3 * nco[f] * (icergh - 1) : 0);
                    for (int b = 0; b < nb; ++b)
                        for (int r = 0; r < 3; ++r)
                            for (int k = 0; k < nco[f]; ++k)
                                pc[((size_t)(ph * nb + b) * 3 + r) * 16 + off[f] + k] =
                                    (float)src[(size_t)b + (size_t)nb * (r + 3 * (size_t)k)];
                    for (int i = 0; i < 4; ++i) L.pbnd[ph][f][i] = (float)B[ph][f][i];
                    // regime min(floor((re - b2) / b3) + 2, 3) of sizes in [b1, b4]:
                    // b1 <= b2 <= b4 and b2 - b3 <= b1 keep it at 1 or above
                    // without a clamp; the kernel's clamp then never acts
                    const double *b = B[ph][f];
                    if (!(b[2] > 0.0) || !(b[0] <= b[1]) || !(b[1] <= b[3]) || !(b[1] - b[2] <= b[0])) {
                        delete c;
                        seterr(errmsg, errlen, "Pade size-regime bounds (%g, %g, %g, %g) need b1 <= b2 <= b4, b3 > 0 "
                               "and b2 - b3 <= b1", b[0], b[1], b[2], b[3]);
                        return 1;
                    }
                }
            L.pcoef = upload(c, pc.data(), pc.size(), &e);
            L.nband = nb; L.nsize_liq = 0; L.nsize_ice = 0;
            L.liq_lwr = (float)B[0][0][0]; L.liq_upr = (float)B[0][0][3];
            L.ice_lwr = (float)B[1][0][0]; L.ice_upr = (float)B[1][0][3];
            c->have_lut = 1;
        } else {
        L.nband = lut->nband; L.nsize_liq = lut->nsize_liq; L.nsize_ice = lut->nsize_ice;
        L.liq_lwr = (float)lut->radliq_lwr; L.liq_upr = (float)lut->radliq_upr;
        L.ice_lwr = (float)lut->radice_lwr; L.ice_upr = (float)lut->radice_upr;
        L.liq_step = (float)((lut->radliq_upr - lut->radliq_lwr) / (lut->nsize_liq - 1));
        L.ice_step = (float)((lut->radice_upr - lut->radice_lwr) / (lut->nsize_ice - 1));
        L.liq_istep = (float)((lut->nsize_liq - 1) / (lut->radliq_upr - lut->radliq_lwr));
        L.ice_istep = (float)((lut->nsize_ice - 1) / (lut->radice_upr - lut->radice_lwr));
        const size_t nl = (size_t)lut->nsize_liq * lut->nband, ni = (size_t)lut->nsize_ice * lut->nband;
        auto a = to_f32(lut->extliq, nl); L.extliq = upload(c, a.data(), nl, &e);
        auto b = to_f32(lut->ssaliq, nl); L.ssaliq = upload(c, b.data(), nl, &e);
        auto cc = to_f32(lut->asyliq, nl); L.asyliq = upload(c, cc.data(), nl, &e);
        const size_t io = ni * (icergh - 1);
        auto d1 = to_f32(lut->extice + io, ni); L.extice = upload(c, d1.data(), ni, &e);
        auto d2 = to_f32(lut->ssaice + io, ni); L.ssaice = upload(c, d2.data(), ni, &e);
        auto d3 = to_f32(lut->asyice + io, ni); L.asyice = upload(c, d3.data(), ni, &e);
        // single-scattering co-albedo in FP64 before rounding: SW clouds have
        // ssa within 1e-5 of 1, where 1 - (float)ssa keeps only ~2 digits
        std::vector<float> cl(nl), ci(ni);
        for (size_t i = 0; i < nl; ++i) cl[i] = (float)(1.0 - lut->ssaliq[i]);
        for (size_t i = 0; i < ni; ++i) ci[i] = (float)(1.0 - lut->ssaice[io + i]);
        L.coaliq = upload(c, cl.data(), nl, &e);
        L.coaice = upload(c, ci.data(), ni, &e);
        if (sizeof(float) * cloud_lut_floats(L) > 48 * 1024) {
            delete c;
            seterr(errmsg, errlen, "cloud LUT larger than the 48 KiB the cloud kernel stages in LDS");
            return 1;
        }
        c->have_lut = 1;
        }
    }
    if (e != hipSuccess) {
        seterr(errmsg, errlen, "device table upload failed: %s", hipGetErrorString(e));
        ccpprad_finalize(c);
        return 2;
    }
    if (scratch_setup(c->device, errmsg, errlen)) {
        ccpprad_finalize(c);
        return 2;
    }
    *out = c;
    return 0;
}

extern "C" int ccpprad_lw_init(const ccpprad_kdist *kd, const ccpprad_cloud_lut *lut, const char *const *gas_names,
                               int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg, int errlen)
{
    return ctx_init(1, kd, lut, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen);
}

extern "C" int ccpprad_sw_init(const ccpprad_kdist *kd, const ccpprad_cloud_lut *lut, const char *const *gas_names,
                               int ngases, int nLay, int blksz, ccpprad_ctx **ctx, char *errmsg, int errlen)
{
    return ctx_init(0, kd, lut, gas_names, ngases, nLay, blksz, ctx, errmsg, errlen);
}

extern "C" void ccpprad_finalize(ccpprad_ctx *c)
{
    if (!c) return;
    if (c->device >= 0) (void)hipDeviceSynchronize();
    for (void *p : c->allocs) (void)hipFree(p);
    for (Workspace *w : c->ws) { w->release(); delete w; }
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    delete c;
}

extern "C" int ccpprad_lw_set_gas_optics_mlp(ccpprad_ctx *c, ccpprad_mlp *m)
{
    if (!c || !c->is_lw) return 1;
    if (m && c->T.ngpt != 128) return 1;   // the emulator predicts 128 g-points
    c->mlp = m;
    return 0;
}

extern "C" int ccpprad_set_launch_columns(ccpprad_ctx *c, int ncol)
{
    if (!c || ncol < 1) return 1;
    std::lock_guard<std::mutex> lk(c->mu);
    c->chunk = std::min(ncol, c->chunk_max);
    return 0;
}

extern "C" int ccpprad_launch_columns(const ccpprad_ctx *c) { return c ? c->chunk : 0; }

extern "C" int ccpprad_ngpt(const ccpprad_ctx *c) { return c ? c->T.ngpt : 0; }
extern "C" int ccpprad_nband(const ccpprad_ctx *c) { return c ? c->T.nband : 0; }
extern "C" const char *ccpprad_version(void) { return "ccpprad-mi355x 0.1 (gfx950)"; }

// Gauss-Jacobi(0,1) secants/weights for n = 1..4 (n = 1: diffusivity 1.66, 0.5).
static void gauss_table(int n, float *D, float *W)
{
    static const double Dt[4][4] = {{1.66, 0, 0, 0},
                                    {1.1835034190722739, 2.8164965809277263, 0, 0},
                                    {1.0971985826868363, 1.6933850767089306, 4.7094163405312410, 0},
                                    {1.0605625725391020, 1.3828256085317800, 2.4014818012578690, 7.1551300134919480}};
    static const double Wt[4][4] = {{0.5, 0, 0, 0},
                                    {0.3180413817439772, 0.1819586182560228, 0, 0},
                                    {0.2009319137389255, 0.2292411063595887, 0.0698269799014858, 0},
                                    {0.1355069134315015, 0.2034645680102685, 0.1298475476082326, 0.0311809709499974}};
    for (int i = 0; i < MAX_ANG; ++i) { D[i] = (float)Dt[n - 1][i]; W[i] = (float)Wt[n - 1][i]; }
}

static bool need(const void *p) { return p != nullptr; }

// McICA arguments of the LW / SW scheme for the launch starting at column
// (LW) or daylit position (SW) c0: seeds ngpt + column (isubc 1) or
// icseed (isubc 2), the overlap flags, and the mask words of that launch
// (rrtmgp_lw_main.F90:441-493, rrtmgp_sw_main.F90:479-523).
static McicaArgs lw_mcica_args(const ccpprad_ctx *c, const ccpprad_lw_args *a, const unsigned char *blk, int c0,
                               unsigned *mask)
{
    McicaArgs ma{};
    ma.nlay = a->nLay; ma.ngpt = c->T.ngpt; ma.ncol_ld = a->nCol; ma.col0 = c0; ma.blksz = a->rrtmgp_phys_blksz;
    ma.isubc = a->isubc_lw;
    ma.iovr = a->iovr; ma.iovr_max = a->iovr_max; ma.iovr_maxrand = a->iovr_maxrand; ma.iovr_rand = a->iovr_rand;
    ma.iovr_dcorr = a->iovr_dcorr; ma.iovr_exp = a->iovr_exp; ma.iovr_exprand = a->iovr_exprand;
    ma.col_idx = nullptr; ma.blkflag = blk; ma.cld_frac = a->cld_frac; ma.overlap = a->cloud_overlap_param;
    ma.icseed = a->icseed_lw; ma.mtpow = c->mtpow; ma.mask = mask;
    return ma;
}

static McicaArgs sw_mcica_args(const ccpprad_ctx *c, const ccpprad_sw_args *a, const unsigned char *blk, int c0,
                               unsigned *mask)
{
    McicaArgs ma{};
    ma.nlay = a->nLay; ma.ngpt = c->T.ngpt; ma.ncol_ld = a->nCol; ma.col0 = c0; ma.blksz = a->rrtmgp_phys_blksz;
    ma.isubc = a->isubc_sw;
    ma.iovr = a->iovr; ma.iovr_max = a->iovr_max; ma.iovr_maxrand = a->iovr_maxrand; ma.iovr_rand = a->iovr_rand;
    ma.iovr_dcorr = a->iovr_dcorr; ma.iovr_exp = a->iovr_exp; ma.iovr_exprand = a->iovr_exprand;
    ma.col_idx = a->idx; ma.blkflag = blk; ma.cld_frac = a->cld_frac; ma.overlap = a->cloud_overlap_param;
    ma.icseed = a->icseed_sw; ma.mtpow = c->mtpow; ma.mask = mask;
    return ma;
}

// The sampled cloud mask of a scheme call on its own: the block cloud flags,
// then the McICA kernel over every column (LW) or daylit column (SW), in
// launches of the context's column count, into mask[col][lay][MASK_WORDS].
static int mcica_mask_entry(ccpprad_ctx *c, int lw, const ccpprad_lw_args *al, const ccpprad_sw_args *as,
                            unsigned *mask, void *stream, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!c || (!al && !as) || !mask) { seterr(errmsg, errlen, "null context, arguments or mask"); return 1; }
    if (!!c->is_lw != !!lw) { seterr(errmsg, errlen, "context is for the other spectral range"); return 1; }
    const int nL = lw ? al->nLay : as->nLay, nC = lw ? al->nCol : as->nCol;
    const int n = lw ? nC : as->nDay, bs = lw ? al->rrtmgp_phys_blksz : as->rrtmgp_phys_blksz;
    const int isubc = lw ? al->isubc_lw : as->isubc_sw, iovr = lw ? al->iovr : as->iovr;
    const bool expo = lw ? (iovr == al->iovr_exp || iovr == al->iovr_exprand || iovr == al->iovr_dcorr)
                         : (iovr == as->iovr_exp || iovr == as->iovr_exprand || iovr == as->iovr_dcorr);
    if (nL != c->nlay) { seterr(errmsg, errlen, "nLay=%d differs from init (%d)", nL, c->nlay); return 1; }
    if (n <= 0) return 0;
    if (bs < 1 || n % bs) { seterr(errmsg, errlen, "column count %d must be a multiple of rrtmgp_phys_blksz=%d", n, bs); return 1; }
    if (!(lw ? al->cld_frac : as->cld_frac) || (!lw && !as->idx)) { seterr(errmsg, errlen, "required array argument is NULL"); return 1; }
    if (isubc == 2 && !(lw ? al->icseed_lw : as->icseed_sw)) { seterr(errmsg, errlen, "isubc=2 needs icseed"); return 1; }
    if (expo && !(lw ? al->cloud_overlap_param : as->cloud_overlap_param)) {
        seterr(errmsg, errlen, "exponential/decorrelation overlap needs cloud_overlap_param"); return 1;
    }
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(c->mu);
    Workspace *w = ws_for_stream(c, st);
    if (!w) { seterr(errmsg, errlen, "workspace allocation failed"); return 2; }
    HIPCHK(w->w_blk.ensure((size_t)(n / bs) + 1));
    unsigned char *blk = (unsigned char *)w->w_blk.p;
    launch_block_flags(n, nL, nC, lw ? nullptr : as->idx, bs, lw ? al->cld_frac : as->cld_frac, blk, nullptr, st);
    const int CH = std::min(n, c->chunk);
    for (int c0 = 0; c0 < n; c0 += CH) {
        unsigned *m = mask + (size_t)c0 * nL * MASK_WORDS;
        const McicaArgs ma = lw ? lw_mcica_args(c, al, blk, c0, m) : sw_mcica_args(c, as, blk, c0, m);
        launch_mcica(ma, std::min(CH, n - c0), st);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ccpprad_lw_mcica_mask_dev(ccpprad_ctx *c, const ccpprad_lw_args *a, unsigned *mask, void *stream,
                                         char *errmsg, int errlen)
{
    return mcica_mask_entry(c, 1, a, nullptr, mask, stream, errmsg, errlen);
}

// ---------------------------------------------------------------------------
// Allocation-free LW McICA mask: the scheme-block cloud flags live in a
// file-scope device array instead of the per-stream workspace.  Every device
// the library runs on holds its own copy until the process ends; host code
// uses only the address hipGetSymbolAddress returned for the current device
// (taken once per device at context init), never the array's name.  Calls
// that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_MAX_BLOCKS = (size_t)1 << 20;   // one flag byte per scheme block
static __device__ __align__(256) unsigned char scratch[NOALLOC_MAX_BLOCKS];
constexpr int NOALLOC_MAX_DEV = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[NOALLOC_MAX_DEV];

// The scratch address on device dev (current): looked up once per device.
static int scratch_setup(int dev, char *errmsg, int errlen)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) { seterr(errmsg, errlen, "device %d beyond %d", dev, NOALLOC_MAX_DEV); return 2; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        seterr(errmsg, errlen, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        return 2;
    }
    g_scratch_ptr[dev] = (unsigned char *)p;
    return 0;
}

static unsigned char *scratch_ptr(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

extern "C" int ccpprad_lw_mcica_mask_noalloc_dev(ccpprad_ctx *c, const ccpprad_lw_args *a, unsigned *mask,
                                                 void *stream, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!c || !a || !mask) { seterr(errmsg, errlen, "null context, arguments or mask"); return 1; }
    if (!c->is_lw) { seterr(errmsg, errlen, "context is not a LW context"); return 1; }
    const int nL = a->nLay, n = a->nCol, bs = a->rrtmgp_phys_blksz, iovr = a->iovr;
    if (nL != c->nlay) { seterr(errmsg, errlen, "nLay=%d differs from init (%d)", nL, c->nlay); return 1; }
    if (n <= 0) return 0;
    if (bs < 1 || n % bs) { seterr(errmsg, errlen, "column count %d must be a multiple of rrtmgp_phys_blksz=%d", n, bs); return 1; }
    if ((size_t)(n / bs) > NOALLOC_MAX_BLOCKS) {
        seterr(errmsg, errlen, "nCol / rrtmgp_phys_blksz = %d blocks exceeds the allocation-free limit of %zu blocks",
               n / bs, NOALLOC_MAX_BLOCKS);
        return 1;
    }
    if (!a->cld_frac) { seterr(errmsg, errlen, "required array argument is NULL"); return 1; }
    if (a->isubc_lw == 2 && !a->icseed_lw) { seterr(errmsg, errlen, "isubc=2 needs icseed"); return 1; }
    if ((iovr == a->iovr_exp || iovr == a->iovr_exprand || iovr == a->iovr_dcorr) && !a->cloud_overlap_param) {
        seterr(errmsg, errlen, "exponential/decorrelation overlap needs cloud_overlap_param"); return 1;
    }
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    unsigned char *blk = scratch_ptr(dev);
    if (!blk) { seterr(errmsg, errlen, "no context was initialised on device %d", dev); return 1; }
    hipStream_t st = (hipStream_t)stream;
    int CH;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        CH = std::min(n, c->chunk);
    }
    // every flag of this call is written (cleared, then set) before it is read
    launch_block_flags(n, nL, n, nullptr, bs, a->cld_frac, blk, nullptr, st);
    for (int c0 = 0; c0 < n; c0 += CH) {
        const McicaArgs ma = lw_mcica_args(c, a, blk, c0, mask + (size_t)c0 * nL * MASK_WORDS);
        launch_mcica(ma, std::min(CH, n - c0), st);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ccpprad_sw_mcica_mask_dev(ccpprad_ctx *c, const ccpprad_sw_args *a, unsigned *mask, void *stream,
                                         char *errmsg, int errlen)
{
    return mcica_mask_entry(c, 0, nullptr, a, mask, stream, errmsg, errlen);
}

// ---------------------------------------------------------------------------
// LW run, device pointers
// ---------------------------------------------------------------------------
extern "C" int ccpprad_lw_run_dev(ccpprad_ctx *c, const ccpprad_lw_args *a, void *stream, int validate,
                                  char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!c || !a) { seterr(errmsg, errlen, "null context or arguments"); return 1; }
    if (!c->is_lw) { seterr(errmsg, errlen, "context is not a LW context"); return 1; }
    if (!a->doLWrad) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int nC = a->nCol, nL = a->nLay;
    const bool scat = a->doGP_lwscat != 0;
    if (scat && a->use_LW_jacobian) {
        seterr(errmsg, errlen, "doGP_lwscat: rte_lw gives no surface-temperature Jacobian with two-stream optics");
        return 1;
    }
    if (nL != c->nlay) { seterr(errmsg, errlen, "nLay=%d differs from init (%d)", nL, c->nlay); return 1; }
    if (nC < 0) { seterr(errmsg, errlen, "nCol negative"); return 1; }
    if (nC == 0) return 0;
    const int bs = a->rrtmgp_phys_blksz;
    if (bs < 1 || nC % bs) { seterr(errmsg, errlen, "nCol=%d must be a multiple of rrtmgp_phys_blksz=%d", nC, bs); return 1; }
    if (a->nGauss_angles < 1 || a->nGauss_angles > MAX_ANG) { seterr(errmsg, errlen, "nGauss_angles must be 1..4"); return 1; }
    const void *req[] = {a->semis, a->tsfg, a->p_lay, a->p_lev, a->t_lay, a->t_lev, a->vmr_h2o, a->vmr_co2, a->vmr_o3,
                         a->vmr_n2o, a->vmr_ch4, a->vmr_o2, a->cld_frac, a->fluxlwUP_allsky, a->fluxlwDOWN_allsky,
                         a->fluxlwUP_clrsky, a->fluxlwDOWN_clrsky};
    for (const void *p : req) if (!need(p)) { seterr(errmsg, errlen, "required array argument is NULL"); return 1; }
    const bool have_clouds = c->have_lut && a->cld_lwp && a->cld_iwp && a->cld_reliq && a->cld_reice;
    if (!c->have_lut && a->cld_frac) {
        // clouds need the LUT; without it the all-sky fluxes are the clear-sky ones
    }
    if (a->isubc_lw == 2 && !a->icseed_lw) { seterr(errmsg, errlen, "isubc_lw=2 needs icseed_lw"); return 1; }
    if ((a->iovr == a->iovr_exp || a->iovr == a->iovr_exprand || a->iovr == a->iovr_dcorr) && !a->cloud_overlap_param) {
        seterr(errmsg, errlen, "exponential/decorrelation overlap needs cloud_overlap_param"); return 1;
    }
    if (a->use_LW_jacobian && !a->fluxlwUP_jac) { seterr(errmsg, errlen, "use_LW_jacobian needs fluxlwUP_jac"); return 1; }
    const TablesView &T = c->T;
    const int G = T.ngpt, B = T.nband;
    const int CH = std::min(nC, c->chunk);
    const int nblk = nC / bs;
    std::unique_lock<std::mutex> lk(c->mu);
    Workspace *w = ws_for_stream(c, st);
    if (!w) { seterr(errmsg, errlen, "workspace allocation failed"); return 2; }
    HIPCHK(w->w_tau.ensure(sizeof(float) * (size_t)CH * nL * G));
    HIPCHK(w->w_aux.ensure(sizeof(float) * (size_t)CH * nL * G));
    HIPCHK(w->w_mask.ensure(sizeof(unsigned) * (size_t)CH * nL * MASK_WORDS));
    HIPCHK(w->w_ctau.ensure((scat ? sizeof(float3) : sizeof(float)) * (size_t)CH * nL * B));
    HIPCHK(w->w_blk.ensure((size_t)nblk + 1));
    HIPCHK(w->w_colf.ensure((size_t)nC + 1));
    if (c->mlp) {
        HIPCHK(w->w_oob.ensure((size_t)CH * nL));
        HIPCHK(w->w_colm.ensure((size_t)CH + 1));
    }
    if (a->aerlw_tau) HIPCHK(w->w_aer.ensure(sizeof(float) * (size_t)CH * nL * B));
    unsigned char *blk = (unsigned char *)w->w_blk.p;
    launch_fill_i32(w->err, 1, 0, st);
    if (have_clouds) launch_block_flags(nC, nL, nC, nullptr, bs, a->cld_frac, blk, (unsigned char *)w->w_colf.p, st);
    else launch_fill_u8(blk, (size_t)nblk, st);
    VmrPtrs vm;
    vm.v[0] = a->vmr_h2o; vm.v[1] = a->vmr_co2; vm.v[2] = a->vmr_o3;
    vm.v[3] = a->vmr_n2o; vm.v[4] = a->vmr_ch4; vm.v[5] = a->vmr_o2;
    float D[MAX_ANG], W[MAX_ANG];
    gauss_table(a->nGauss_angles, D, W);
    for (int c0 = 0; c0 < nC; c0 += CH) {
        const int nc = std::min(CH, nC - c0);
        float *tau = (float *)w->w_tau.p, *pf = (float *)w->w_aux.p;
        unsigned *mask = (unsigned *)w->w_mask.p;
        // cloud optics by band: absorption depth, or the full two-stream
        // record when the all-sky solver scatters
        float *ctau = scat ? nullptr : (float *)w->w_ctau.p;
        float3 *copt = scat ? (float3 *)w->w_ctau.p : nullptr;
        if (have_clouds) {
            CloudArgs ca{};
            ca.nlay = nL; ca.ncol_ld = nC; ca.col0 = c0; ca.blksz = bs; ca.nband = B; ca.is_lw = 1;
            ca.col_idx = nullptr; ca.blkflag = blk; ca.cld_frac = a->cld_frac;
            ca.lwp = a->cld_lwp; ca.iwp = a->cld_iwp; ca.reliq = a->cld_reliq; ca.reice = a->cld_reice;
            if (a->doGP_sgs_cnv) { ca.cnv[0] = a->cld_cnv_lwp; ca.cnv[1] = a->cld_cnv_iwp; ca.cnv[2] = a->cld_cnv_reliq; ca.cnv[3] = a->cld_cnv_reice; }
            if (a->doGP_sgs_pbl) { ca.pbl[0] = a->cld_pbl_lwp; ca.pbl[1] = a->cld_pbl_iwp; ca.pbl[2] = a->cld_pbl_reliq; ca.pbl[3] = a->cld_pbl_reice; }
            ca.rwp = a->cld_rwp; ca.swp = a->cld_swp; ca.resnow = a->cld_resnow;
            if (!ca.rwp || !ca.swp || !ca.resnow) { seterr(errmsg, errlen, "cld_rwp/cld_swp/cld_resnow required with clouds"); return 1; }
            ca.ctau = ctau; ca.copt = copt; ca.err = w->err;
            { KScope k_(c, K_CLOUD, st); launch_cloud_optics(c->L, ca, nc, st); }
            const McicaArgs ma = lw_mcica_args(c, a, blk, c0, mask);
            { KScope k_(c, K_MCICA, st); launch_mcica(ma, nc, st); }
        }
        if (c->mlp) {
            MlpInferArgs mi{};
            mi.nlay = nL; mi.ncol_ld = nC; mi.col0 = c0; mi.ncols = nc;
            mi.p_lay = a->p_lay; mi.p_lev = a->p_lev; mi.t_lay = a->t_lay;
            mi.vmr[0] = a->vmr_h2o; mi.vmr[1] = a->vmr_o3; mi.vmr[2] = a->vmr_co2;
            mi.vmr[3] = a->vmr_ch4; mi.vmr[4] = a->vmr_n2o;
            mi.tau = tau; mi.pfrac = pf;
            mi.oob = (unsigned char *)w->w_oob.p; mi.colflag = (unsigned char *)w->w_colm.p;
            launch_fill_u8(mi.colflag, (size_t)nc, st);
            { KScope k_(c, K_MLP, st); launch_mlp_infer(c->mlp, mi, st); }
            // layers outside the training hull: table gas optics
            { KScope k_(c, K_GASFB, st); launch_gas_optics(T, nc, nL, nC, nullptr, c0, a->p_lay, a->p_lev, a->t_lay, vm, w->err, tau, pf, st, mi.oob, mi.colflag); }
        } else {
            { KScope k_(c, K_GAS, st); launch_gas_optics(T, nc, nL, nC, nullptr, c0, a->p_lay, a->p_lev, a->t_lay, vm, w->err, tau, pf, st); }
        }
        LwSolveArgs s{};
        s.nlay = nL; s.ncol_ld = nC; s.col0 = c0; s.top_at_1 = a->top_at_1;
        s.tau = tau; s.pfrac = pf; s.mask = mask; s.ctau = ctau; s.copt = copt; s.aer_tau = a->aerlw_tau;
        s.aer1 = a->aerlw_tau ? (const float *)w->w_aer.p : nullptr;
        s.t_lay = a->t_lay; s.t_lev = a->t_lev; s.tsfc = a->tsfg; s.semis = a->semis; s.p_lev = a->p_lev;
        for (int i = 0; i < MAX_ANG; ++i) { s.D[i] = D[i]; s.W[i] = W[i]; }
        s.flux_up = a->fluxlwUP_allsky; s.flux_dn = a->fluxlwDOWN_allsky;
        s.flux_up2 = a->fluxlwUP_radtime; s.flux_dn2 = a->fluxlwDOWN_radtime;
        s.flux_jac = a->use_LW_jacobian ? a->fluxlwUP_jac : nullptr;
        s.htr = a->htrlw;
        if (a->aerlw_tau) { KScope k_(c, K_AER, st); launch_lw_aer_pack(s, nc, B, (float *)w->w_aer.p, st); }
        if (scat) {
            KScope k_(c, K_LWSOLVE, st);
            launch_lw_solve_2str(T, s, nc, have_clouds, st);
        } else {
            KScope k_(c, K_LWSOLVE, st);
            launch_lw_solve(T, s, nc, a->nGauss_angles, have_clouds, a->use_LW_jacobian != 0, false, st);
        }
        if (a->doLWclrsky) {
            LwSolveArgs s2 = s;
            s2.mask = nullptr; s2.ctau = nullptr; s2.copt = nullptr;
            s2.flux_up = a->fluxlwUP_clrsky; s2.flux_dn = a->fluxlwDOWN_clrsky;
            s2.flux_up2 = nullptr; s2.flux_dn2 = nullptr; s2.flux_jac = nullptr;
            s2.htr = a->htrlwc;
            KScope k_(c, K_LWSOLVE, st);
            launch_lw_solve(T, s2, nc, a->nGauss_angles, false, false, a->nGauss_angles == 1, st);
        }
    }
    if (!a->doLWclrsky) {
        launch_fill_f64(a->fluxlwUP_clrsky, (size_t)nC * (nL + 1), 0.0, st);
        launch_fill_f64(a->fluxlwDOWN_clrsky, (size_t)nC * (nL + 1), 0.0, st);
    }
    HIPCHK(hipGetLastError());
    if (validate) {
        int herr = 0;
        HIPCHK(hipMemcpyAsync(&herr, w->err, sizeof(int), hipMemcpyDeviceToHost, st));
        lk.unlock();   // the copy is ordered after this call's kernels
        HIPCHK(hipStreamSynchronize(st));
        if (herr & 1) { seterr(errmsg, errlen, "gas_optics(): p_lay or t_lay outside the table range"); return 3; }
        if (herr & 2) { seterr(errmsg, errlen, "cloud_optics(): effective radius outside the LUT range"); return 3; }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// SW run, device pointers
// ---------------------------------------------------------------------------
extern "C" int ccpprad_sw_run_dev(ccpprad_ctx *c, const ccpprad_sw_args *a, void *stream, int validate,
                                  char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!c || !a) { seterr(errmsg, errlen, "null context or arguments"); return 1; }
    if (c->is_lw) { seterr(errmsg, errlen, "context is not a SW context"); return 1; }
    if (!a->doSWrad) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int nC = a->nCol, nL = a->nLay, nD = a->nDay, nV = nL + 1;
    if (nL != c->nlay) { seterr(errmsg, errlen, "nLay=%d differs from init (%d)", nL, c->nlay); return 1; }
    const void *req[] = {a->fluxswUP_allsky, a->fluxswDOWN_allsky, a->fluxswUP_clrsky, a->fluxswDOWN_clrsky, a->scmpsw, a->cldtausw};
    for (const void *p : req) if (!need(p)) { seterr(errmsg, errlen, "required output array is NULL"); return 1; }
    if (nD <= 0) {
        launch_fill_f64(a->fluxswUP_allsky, (size_t)nC * nV, 0.0, st);
        launch_fill_f64(a->fluxswDOWN_allsky, (size_t)nC * nV, 0.0, st);
        launch_fill_f64(a->fluxswUP_clrsky, (size_t)nC * nV, 0.0, st);
        launch_fill_f64(a->fluxswDOWN_clrsky, (size_t)nC * nV, 0.0, st);
        launch_fill_f64(a->scmpsw, (size_t)6 * nC, 0.0, st);
        if (a->htrsw) launch_fill_f64(a->htrsw, (size_t)nC * nL, 0.0, st);
        if (a->htrswc) launch_fill_f64(a->htrswc, (size_t)nC * nL, 0.0, st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    const int bs = a->rrtmgp_phys_blksz;
    if (bs < 1 || nD % bs) { seterr(errmsg, errlen, "nDay=%d must be a multiple of rrtmgp_phys_blksz=%d", nD, bs); return 1; }
    if (nD > nC) { seterr(errmsg, errlen, "nDay exceeds nCol"); return 1; }
    const void *req2[] = {a->idx, a->coszen, a->sfc_alb_nir_dir, a->sfc_alb_nir_dif, a->sfc_alb_uvvis_dir,
                          a->sfc_alb_uvvis_dif, a->p_lay, a->p_lev, a->t_lay, a->vmr_h2o, a->vmr_co2, a->vmr_o3,
                          a->vmr_n2o, a->vmr_ch4, a->vmr_o2, a->cld_frac};
    for (const void *p : req2) if (!need(p)) { seterr(errmsg, errlen, "required array argument is NULL"); return 1; }
    if (a->iSFC < 1 || a->iSFC > nV) { seterr(errmsg, errlen, "iSFC out of range"); return 1; }
    if (a->isubc_sw == 2 && !a->icseed_sw) { seterr(errmsg, errlen, "isubc_sw=2 needs icseed_sw"); return 1; }
    if ((a->aersw_tau != nullptr) != (a->aersw_ssa != nullptr) || (a->aersw_tau != nullptr) != (a->aersw_g != nullptr)) {
        seterr(errmsg, errlen, "aersw_tau, aersw_ssa and aersw_g: give all three or none"); return 1;
    }
    if ((a->iovr == a->iovr_exp || a->iovr == a->iovr_exprand || a->iovr == a->iovr_dcorr) && !a->cloud_overlap_param) {
        seterr(errmsg, errlen, "exponential/decorrelation overlap needs cloud_overlap_param"); return 1;
    }
    const bool have_clouds = c->have_lut && a->cld_lwp && a->cld_iwp && a->cld_reliq && a->cld_reice;
    std::unique_lock<std::mutex> lk(c->mu);
    Workspace *w = ws_for_stream(c, st);
    if (!w) { seterr(errmsg, errlen, "workspace allocation failed"); return 2; }
    TablesView T = c->T;   // with this call's TOA source
    T.toa_norm = w->toa32;
    T.toa_norm64 = w->toa64;
    const int G = T.ngpt, B = T.nband;
    const int CH = std::min(nD, c->chunk);
    const int nblk = nD / bs;
    HIPCHK(w->w_tau.ensure(sizeof(float) * (size_t)CH * nL * G));
    HIPCHK(w->w_aux.ensure(sizeof(float) * (size_t)CH * nL * G));
    HIPCHK(w->w_mask.ensure(sizeof(unsigned) * (size_t)CH * nL * MASK_WORDS));
    HIPCHK(w->w_ctau.ensure(sizeof(float3) * (size_t)CH * nL * B));
    HIPCHK(w->w_sb.ensure(sizeof(double) * (size_t)nD * 3 * B));
    HIPCHK(w->w_blk.ensure((size_t)nblk + 1));
    HIPCHK(w->w_colf.ensure((size_t)nD + 1));
    HIPCHK(w->w_day.ensure((size_t)nC + 1));
    const bool have_aer = a->aersw_tau != nullptr;
    if (have_aer) HIPCHK(w->w_aer.ensure(sizeof(float3) * (size_t)CH * nL * B));
    unsigned char *blk = (unsigned char *)w->w_blk.p, *colf = (unsigned char *)w->w_colf.p;
    launch_fill_i32(w->err, 1, 0, st);
    launch_fill_i32((w->err + 1), 1, -1, st);
    launch_fill_u8(colf, (size_t)nD, st);
    if (have_clouds) launch_block_flags(nD, nL, nC, a->idx, bs, a->cld_frac, blk, colf, st);
    else launch_fill_u8(blk, (size_t)nblk, st);
    launch_toa_norm(c->solar_dev, G, a->solcon, w->toa64, w->toa32, st);
    // cldtausw: reset by every block of the reference loop, so only the last
    // block's values survive (rrtmgp_sw_main.F90:308,403)
    launch_fill_f64(a->cldtausw, (size_t)nC * nL, 0.0, st);
    if (a->htrsw) launch_fill_f64(a->htrsw, (size_t)nC * nL, 0.0, st);
    if (a->htrswc) launch_fill_f64(a->htrswc, (size_t)nC * nL, 0.0, st);
    VmrPtrs vm;
    vm.v[0] = a->vmr_h2o; vm.v[1] = a->vmr_co2; vm.v[2] = a->vmr_o3;
    vm.v[3] = a->vmr_n2o; vm.v[4] = a->vmr_ch4; vm.v[5] = a->vmr_o2;
    const int last_blk_start = (nblk - 1) * bs;
    for (int c0 = 0; c0 < nD; c0 += CH) {
        const int nc = std::min(CH, nD - c0);
        float *tabs = (float *)w->w_tau.p, *tray = (float *)w->w_aux.p;
        unsigned *mask = (unsigned *)w->w_mask.p;
        float3 *copt = (float3 *)w->w_ctau.p;
        if (have_clouds) {
            CloudArgs ca{};
            ca.nlay = nL; ca.ncol_ld = nC; ca.col0 = c0; ca.blksz = bs; ca.nband = B; ca.is_lw = 0;
            ca.col_idx = a->idx; ca.blkflag = blk; ca.cld_frac = a->cld_frac;
            ca.lwp = a->cld_lwp; ca.iwp = a->cld_iwp; ca.reliq = a->cld_reliq; ca.reice = a->cld_reice;
            if (a->doGP_sgs_cnv) { ca.cnv[0] = a->cld_cnv_lwp; ca.cnv[1] = a->cld_cnv_iwp; ca.cnv[2] = a->cld_cnv_reliq; ca.cnv[3] = a->cld_cnv_reice; }
            if (a->doGP_sgs_pbl) { ca.pbl[0] = a->cld_pbl_lwp; ca.pbl[1] = a->cld_pbl_iwp; ca.pbl[2] = a->cld_pbl_reliq; ca.pbl[3] = a->cld_pbl_reice; }
            ca.rwp = a->cld_rwp; ca.swp = a->cld_swp; ca.resnow = a->cld_resnow;
            if (!ca.rwp || !ca.swp || !ca.resnow) { seterr(errmsg, errlen, "cld_rwp/cld_swp/cld_resnow required with clouds"); return 1; }
            ca.copt = copt; ca.err = w->err;
            // only the last block's band-11 cloud depth survives the per-block reset
            ca.cldtau = (c0 + nc > last_blk_start) ? a->cldtausw : nullptr;
            ca.cldtau_blk = nblk - 1;
            { KScope k_(c, K_CLOUD, st); launch_cloud_optics(c->L, ca, nc, st); }
            const McicaArgs ma = sw_mcica_args(c, a, blk, c0, mask);
            { KScope k_(c, K_MCICA, st); launch_mcica(ma, nc, st); }
        }
        { KScope k_(c, K_GAS, st); launch_gas_optics(T, nc, nL, nC, a->idx, c0, a->p_lay, a->p_lev, a->t_lay, vm, w->err, tabs, tray, st); }
        SwSolveArgs s{};
        s.nlay = nL; s.ncol_ld = nC; s.col0 = c0; s.top_at_1 = a->top_at_1; s.blksz = bs; s.isfc = a->iSFC - 1;
        s.write_clr_zero = !a->doSWclrsky; s.role = 0;
        s.col_idx = a->idx; s.tau_abs = tabs; s.tau_ray = tray; s.mask = mask; s.copt = copt;
        s.blkflag = blk; s.aer_tau = a->aersw_tau; s.aer_ssa = a->aersw_ssa; s.aer_g = a->aersw_g;
        s.aer4 = have_aer ? (const float3 *)w->w_aer.p : nullptr;
        s.coszen = a->coszen; s.alb_nir_dir = a->sfc_alb_nir_dir; s.alb_nir_dif = a->sfc_alb_nir_dif;
        s.alb_uv_dir = a->sfc_alb_uvvis_dir; s.alb_uv_dif = a->sfc_alb_uvvis_dif; s.p_lev = a->p_lev;
        s.alb_cat = c->alb_cat;
        s.up_all = a->fluxswUP_allsky; s.dn_all = a->fluxswDOWN_allsky; s.up_clr = a->fluxswUP_clrsky;
        s.dn_clr = a->fluxswDOWN_clrsky; s.sb = (double *)w->w_sb.p + (size_t)c0 * 3 * B; s.htr = a->htrsw;
        s.htr_clr = a->doSWclrsky ? a->htrswc : nullptr;
        if (have_aer) { KScope k_(c, K_AER, st); launch_sw_aer_pack(s, nc, B, (float3 *)w->w_aer.p, st); }
        { KScope k_(c, K_SWSOLVE, st); launch_sw_solve(T, s, nc, have_clouds, st); }
        if (a->doSWclrsky && have_clouds) {
            SwSolveArgs s2 = s;
            s2.role = 1; s2.write_clr_zero = 0; s2.mask = nullptr; s2.htr = nullptr;
            KScope k_(c, K_SWSOLVE, st);
            launch_sw_solve(T, s2, nc, false, st);
        }
    }
    ScmpArgs sc{};
    sc.ndays = nD; sc.blksz = bs; sc.nband = B; sc.ibd = c->ibd; sc.ibd_uv = c->ibd_uv; sc.do_clr = a->doSWclrsky;
    sc.col_idx = a->idx; sc.blkflag = blk; sc.colflag = colf; sc.sb = (const double *)w->w_sb.p;
    sc.last_wipe = (w->err + 1); sc.scmpsw = a->scmpsw;
    launch_fill_u8((unsigned char *)w->w_day.p, (size_t)nC, st);
    launch_scmpsw(sc, nC, (unsigned char *)w->w_day.p, (w->err + 1), st);
    HIPCHK(hipGetLastError());
    if (validate) {
        int herr = 0;
        HIPCHK(hipMemcpyAsync(&herr, w->err, sizeof(int), hipMemcpyDeviceToHost, st));
        lk.unlock();   // the copy is ordered after this call's kernels
        HIPCHK(hipStreamSynchronize(st));
        if (herr & 1) { seterr(errmsg, errlen, "gas_optics(): p_lay or t_lay outside the table range"); return 3; }
        if (herr & 2) { seterr(errmsg, errlen, "cloud_optics(): effective radius outside the LUT range"); return 3; }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// component entries
// ---------------------------------------------------------------------------
extern "C" int ccpprad_heating_rate_dev(int nCol, int nLay, const double *flux_up, const double *flux_dn,
                                        const double *p_lev, double *htr, void *stream, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (nCol < 0 || nLay < 1) { seterr(errmsg, errlen, "bad dimensions"); return 1; }
    if (!flux_up || !flux_dn || !p_lev || !htr) { seterr(errmsg, errlen, "NULL array"); return 1; }
    launch_heating_rate(nCol, nLay, flux_up, flux_dn, p_lev, htr, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

namespace ccpprad {
void launch_post(const ccpprad_post_args &A, hipStream_t st);                // prep.hip
void launch_cloud_overlap(const ccpprad_cloud_overlap_args &A, hipStream_t st);   // prep.hip
}

extern "C" int ccpprad_post_dev(const ccpprad_post_args *a, void *stream, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!a) { seterr(errmsg, errlen, "null arguments"); return 1; }
    if (a->nCol < 0 || a->nLev < 1) { seterr(errmsg, errlen, "bad dimensions"); return 1; }
    if (a->nCol == 0 || !(a->doLWrad || a->doSWrad)) return 0;
    if (a->iSFC < 1 || a->iSFC > a->nLev + 1 || a->iTOA < 1 || a->iTOA > a->nLev + 1) {
        seterr(errmsg, errlen, "iSFC / iTOA outside 1..nLev+1"); return 1;
    }
    auto all = [](std::initializer_list<const void *> l) {
        for (const void *p : l) if (!p) return false;
        return true;
    };
    if (a->doLWrad && !all({a->fluxlwUP_allsky, a->fluxlwDOWN_allsky, a->fluxlwUP_clrsky, a->fluxlwDOWN_clrsky,
                            a->tsfa, a->htrlw, a->topflw, a->sfcflw, a->sfcdlw, a->sfculw, a->tsflw, a->htrlwu})) {
        seterr(errmsg, errlen, "LW: NULL array"); return 1;
    }
    if (a->doSWrad && !all({a->fluxswUP_allsky, a->fluxswDOWN_allsky, a->fluxswUP_clrsky, a->fluxswDOWN_clrsky,
                            a->scmpsw, a->sfc_alb_nir_dir, a->sfc_alb_nir_dif, a->sfc_alb_uvvis_dir,
                            a->sfc_alb_uvvis_dif, a->topfsw, a->sfcfsw, a->nirbmdi, a->nirdfdi, a->visbmdi,
                            a->visdfdi, a->nirbmui, a->nirdfui, a->visbmui, a->visdfui, a->sfcnsw, a->sfcdsw})) {
        seterr(errmsg, errlen, "SW: NULL array"); return 1;
    }
    if (a->save_diag) {
        if (!a->fluxr || a->nfxr < 45 || !a->cldsa || !a->mtopa || !a->mbota) {
            seterr(errmsg, errlen, "save_diag needs fluxr (nfxr >= 45), cldsa, mtopa and mbota"); return 1;
        }
        if (a->doLWrad && !a->cldtaulw) { seterr(errmsg, errlen, "save_diag needs cldtaulw"); return 1; }
        if (a->doSWrad && (!a->cldtausw || !a->aerodp || a->naerodp < 6 || !a->coszen || !a->coszdg || !a->p_lev)) {
            seterr(errmsg, errlen, "save_diag needs cldtausw, aerodp (>= 6), coszen, coszdg and p_lev"); return 1;
        }
    }
    launch_post(*a, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ccpprad_layer_geometry_dev(int nCol, int nLev, double con_rd, double con_g, double con_fvirt,
                                          const double *p_lev, const double *p_lay, const double *t_lay,
                                          const double *q_lay, double *tv_lay, double *deltaZ,
                                          double *deltaZc, double *deltaP, void *stream, char *errmsg,
                                          int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (nCol < 0 || nLev < 1) { seterr(errmsg, errlen, "bad dimensions"); return 1; }
    if (nCol == 0) return 0;
    if (!p_lev || !p_lay || !t_lay || !q_lay || !deltaZc) { seterr(errmsg, errlen, "NULL array"); return 1; }
    if (!(con_g != 0.0)) { seterr(errmsg, errlen, "con_g must be nonzero"); return 1; }
    GeomArgs A{};
    A.ncol = nCol; A.nlev = nLev;
    A.rdog = con_rd / con_g; A.fvirt = con_fvirt;
    A.p_lev = p_lev; A.p_lay = p_lay; A.t_lay = t_lay; A.q_lay = q_lay;
    A.tv_lay = tv_lay; A.dz = deltaZ; A.dzc = deltaZc; A.dp = deltaP;
    launch_layer_geometry(A, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ccpprad_cloud_overlap_dev(const ccpprad_cloud_overlap_args *a, void *stream, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!a) { seterr(errmsg, errlen, "null arguments"); return 1; }
    if (a->nCol < 0 || a->nLev < 1) { seterr(errmsg, errlen, "bad dimensions"); return 1; }
    if (a->nLev > 65535) { seterr(errmsg, errlen, "nLev above 65535"); return 1; }
    if (a->nCol == 0 || !(a->doLWrad || a->doSWrad)) return 0;
    if (!a->de_lgth || !a->cloud_overlap_param || !a->precip_overlap_param) {
        seterr(errmsg, errlen, "NULL output array"); return 1;
    }
    auto expo = [&](int iovr) { return iovr == a->iovr_dcorr || iovr == a->iovr_exp || iovr == a->iovr_exprand; };
    const bool cnv = a->imfdeepcnv == a->imfdeepcnv_samf || a->imfdeepcnv == a->imfdeepcnv_gf;
    if ((a->idcor == a->idcor_hogan || a->idcor == a->idcor_oreopoulos) && !a->lat) {
        seterr(errmsg, errlen, "the decorrelation length of idcor needs lat"); return 1;
    }
    if ((expo(a->iovr) || (cnv && expo(a->iovr_convcld))) && !a->deltaZc) {
        seterr(errmsg, errlen, "an exponential overlap needs deltaZc"); return 1;
    }
    if (expo(a->iovr) && a->iovr == a->iovr_exprand && !a->cld_frac) {
        seterr(errmsg, errlen, "exponential-random overlap needs cld_frac"); return 1;
    }
    if (cnv && !a->cnv_cloud_overlap_param) {
        seterr(errmsg, errlen, "imfdeepcnv selects SAMF or GF: cnv_cloud_overlap_param is NULL"); return 1;
    }
    if (cnv && expo(a->iovr_convcld) && a->iovr_convcld == a->iovr_exprand && !a->cld_cnv_frac) {
        seterr(errmsg, errlen, "exponential-random convective overlap needs cld_cnv_frac"); return 1;
    }
    launch_cloud_overlap(*a, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ccpprad_prepare_inputs_dev(const ccpprad_ctx *c, int nCol, int nLev, const double *prsi,
                                          const double *prsl, const double *tgrs, const double *qgrs_h2o,
                                          const double *qgrs_o3, const double *tsfc, double *p_lay,
                                          double *p_lev, double *t_lay, double *t_lev, double *q_lay,
                                          double *vmr_h2o, double *vmr_o3, int *top_at_1, void *stream,
                                          char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!c || !c->is_lw) { seterr(errmsg, errlen, "the table range comes from a LW context"); return 1; }
    if (nCol < 0 || nLev < 2) { seterr(errmsg, errlen, "bad dimensions"); return 1; }
    if (!prsi || !prsl || !tgrs || !qgrs_h2o || !tsfc || !p_lay || !p_lev || !t_lay || !t_lev || !vmr_h2o ||
        (qgrs_o3 && !vmr_o3)) {
        seterr(errmsg, errlen, "NULL array"); return 1;
    }
    PrepArgs A{};
    A.ncol = nCol; A.nlev = nLev;
    A.prsi = prsi; A.prsl = prsl; A.tgrs = tgrs; A.qh2o = qgrs_h2o; A.qo3 = qgrs_o3; A.tsfc = tsfc;
    A.p_lay = p_lay; A.p_lev = p_lev; A.t_lay = t_lay; A.t_lev = t_lev; A.q_lay = q_lay;
    A.vmr_h2o = vmr_h2o; A.vmr_o3 = vmr_o3; A.top_at_1 = top_at_1;
    A.press_min = c->T.press_min; A.press_max = c->T.press_max;
    A.temp_min = c->T.temp_min; A.temp_max = c->T.temp_max;
    launch_prepare(A, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

static int gas_optics_entry(ccpprad_ctx *c, int want_lw, int nCol, int nLay, const double *p_lay,
                            const double *p_lev, const double *t_lay, const double *vmr, float *tau, float *aux,
                            void *stream, char *errmsg, int errlen)
{
    if (errmsg && errlen > 0) errmsg[0] = 0;
    if (!c || c->is_lw != want_lw) { seterr(errmsg, errlen, "wrong or null context"); return 1; }
    if (nCol < 0 || nLay < 1) { seterr(errmsg, errlen, "bad dimensions"); return 1; }
    if (!p_lay || !p_lev || !t_lay || !vmr || !tau || !aux) { seterr(errmsg, errlen, "NULL array"); return 1; }
    VmrPtrs vm;
    for (int g = 0; g < MAX_GAS; ++g) vm.v[g] = vmr + (size_t)std::min(g, c->T.ngas - 1) * nCol * nLay;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(c->mu);
    Workspace *w = ws_for_stream(c, st);
    if (!w) { seterr(errmsg, errlen, "workspace allocation failed"); return 2; }
    launch_fill_i32(w->err, 1, 0, st);
    if (want_lw && c->mlp) {
        // the gas-optics stage of the LW run with the emulator attached:
        // MLP inference, then the table path on the layers outside its hull
        if (nLay != c->nlay) { seterr(errmsg, errlen, "nLay=%d differs from init (%d)", nLay, c->nlay); return 1; }
        HIPCHK(w->w_oob.ensure((size_t)nCol * nLay));
        HIPCHK(w->w_colm.ensure((size_t)nCol + 1));
        MlpInferArgs mi{};
        mi.nlay = nLay; mi.ncol_ld = nCol; mi.col0 = 0; mi.ncols = nCol;
        mi.p_lay = p_lay; mi.p_lev = p_lev; mi.t_lay = t_lay;
        mi.vmr[0] = vm.v[0]; mi.vmr[1] = vm.v[2]; mi.vmr[2] = vm.v[1]; mi.vmr[3] = vm.v[4]; mi.vmr[4] = vm.v[3];
        mi.tau = tau; mi.pfrac = aux;
        mi.oob = (unsigned char *)w->w_oob.p; mi.colflag = (unsigned char *)w->w_colm.p;
        launch_fill_u8(mi.colflag, (size_t)nCol, st);
        launch_mlp_infer(c->mlp, mi, st);
        launch_gas_optics(c->T, nCol, nLay, nCol, nullptr, 0, p_lay, p_lev, t_lay, vm, w->err, tau, aux, st, mi.oob,
                          mi.colflag);
    } else {
        launch_gas_optics(c->T, nCol, nLay, nCol, nullptr, 0, p_lay, p_lev, t_lay, vm, w->err, tau, aux, st);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int ccpprad_lw_gas_optics_dev(ccpprad_ctx *c, int nCol, int nLay, const double *p_lay, const double *p_lev,
                                         const double *t_lay, const double *vmr, float *tau, float *pfrac,
                                         void *stream, char *errmsg, int errlen)
{
    return gas_optics_entry(c, 1, nCol, nLay, p_lay, p_lev, t_lay, vmr, tau, pfrac, stream, errmsg, errlen);
}

extern "C" int ccpprad_sw_gas_optics_dev(ccpprad_ctx *c, int nCol, int nLay, const double *p_lay, const double *p_lev,
                                         const double *t_lay, const double *vmr, float *tau_abs, float *tau_ray,
                                         void *stream, char *errmsg, int errlen)
{
    return gas_optics_entry(c, 0, nCol, nLay, p_lay, p_lev, t_lay, vmr, tau_abs, tau_ray, stream, errmsg, errlen);
}
