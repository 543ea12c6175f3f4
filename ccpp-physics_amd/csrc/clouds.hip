// Cloud optics (LUT), precipitation optics and McICA sampling on gfx950.
//
// cloud_optics_kernel restates ty_cloud_optics%cloud_optics (LUT branch,
// compute_all_from_table + combine; call sites rrtmgp_lw_main.F90:366-401,
// rrtmgp_sw_main.F90:396-432), the rain/snow optics of the scheme layer
// (rrtmgp_lw_main.F90:409-431; rrtmgp_sw_main.F90:436-471, constants
// rrtmgp_lw_cloud_optics.F90:63-66, rrtmgp_sw_cloud_optics.F90:59-63,369-381)
// and increment_2stream_by_2stream.  One thread per (column, layer, band), FP32.
//
// mcica_kernel restates the per-column seeding and MT19937 draws of
// rrtmgp_lw_main.F90:441-466 / rrtmgp_sw_main.F90:479-509
// (physics/Radiation/mersenne_twister.f:285-372) and sampled_mask
// (rrtmgp_sampling.F90:84-182).  One wavefront per column: the 624-word MT
// state lives in LDS, the twist runs 64 words at a time (reads of a chunk
// complete before its writes, so every word reads its successor's old value
// and the word 227 places back already updated, exactly as the serial loop
// does), tempered blocks go to a two-slot LDS ring, and the uniform draws are
// converted to FP64 exactly as the Fortran does (non-negative draws pass
// through default REAL).  Masks are bit-identical to the reference's.
#include "kernels.h"

namespace ccpprad {


// Divisions below go through v_rcp_f32 (about 1 ulp): the results are FP32
// optics, and the table position's floor decides only which of two equal
// neighbouring interpolants is taken on an exact grid point.
__device__ __forceinline__ void lut_eval(const float *ext, const float *ssa, const float *coa, const float *asy,
                                         int nsz, float lwr, float istep, float path, float re, float &t, float &ts,
                                         float &ta, float &tsg)
{
    const float x = (re - lwr) * istep;
    int idx = min((int)floorf(x) + 1, nsz - 1);
    float fint = x - (float)(idx - 1);
    float e0 = ext[idx - 1], e1 = ext[idx];
    float s0 = ssa[idx - 1], s1 = ssa[idx];
    float c0 = coa[idx - 1], c1 = coa[idx];
    float a0 = asy[idx - 1], a1 = asy[idx];
    t = path * (e0 + fint * (e1 - e0));
    ts = t * (s0 + fint * (s1 - s0));
    ta = t * (c0 + fint * (c1 - c0));   // = t - ts, without the cancellation
    tsg = ts * (a0 + fint * (a1 - a0));
}

// rte-rrtmgp's pade_eval: numerator degree m, denominator degree n with
// constant term 1, coefficients c[0 .. m + n] (Horner, FP32).
__device__ __forceinline__ float pade_f(const float *c, int m, int n, float re)
{
    float den = c[n + m];
    for (int i = n + m - 1; i >= m + 1; --i) den = fmaf(re, den, c[i]);
    den = fmaf(re, den, 1.0f);
    float num = c[m];
    for (int i = m - 1; i >= 0; --i) num = fmaf(re, num, c[i]);
    return num / den;
}

// the size regime of a radius, rte's min(floor((re - b2) / b3) + 2, 3), 0-based
__device__ __forceinline__ int pade_reg(const float *bd, float re)
{
    return min(max((int)floorf((re - bd[1]) / bd[2]) + 2, 1), 3) - 1;
}

// compute_all_from_pade for one phase (0 liquid, 1 ice) and band:
// t = path ext, absorption t max(0, co-albedo), ts = t - that, tsg = ts g.
__device__ __forceinline__ void pade_phase(const LutView &L, int ph, int b, float path, float re, float &t,
                                           float &ts, float &ta, float &tsg)
{
    const float *pc = L.pcoef + (size_t)(ph * L.nband + b) * 48;
    const float e = pade_f(pc + 16 * pade_reg(L.pbnd[ph][0], re), 2, 3, re);
    const float co = fmaxf(0.0f, pade_f(pc + 16 * pade_reg(L.pbnd[ph][1], re) + 6, 2, 2, re));
    const float g = pade_f(pc + 16 * pade_reg(L.pbnd[ph][2], re) + 11, 2, 2, re);
    t = path * e;
    ta = t * co;
    ts = t * (1.0f - co);
    tsg = ts * g;
}

// Two-stream optics carried with the co-albedo: (tau, ssa, 1 - ssa, g).
struct Opt2 {
    float t, w, p, g;
};

// increment_2stream_by_2stream: b += a, eps = 3*tiny(FP32).  The co-albedo
// combines as the absorption-weighted mean.
__device__ __forceinline__ void inc2s(const Opt2 &a, Opt2 &b)
{
    const float eps = 3.0f * 1.17549435e-38f;
    const float t12 = a.t + b.t;
    const float ts12 = a.t * a.w + b.t * b.w;
    const float ta12 = a.t * a.p + b.t * b.p;
    const float rt = frcp(fmaxf(eps, t12));
    b.g = (a.t * a.w * a.g + b.t * b.w * b.g) * frcp(fmaxf(eps, ts12));
    b.w = ts12 * rt;
    b.p = t12 > eps ? ta12 * rt : 1.0f;
    b.t = t12;
}

// The effective-radius range check lut_cloud makes for a layer's condensate
// (LUT and Pade forms alike: the Pade bounds are the extinction size bounds).
__device__ __forceinline__ bool lut_out_of_range(const LutView &L, float lwp, float iwp, float rl, float ri)
{
    return (lwp > 0.0f && !(rl >= L.liq_lwr && rl <= L.liq_upr)) || (iwp > 0.0f && !(ri >= L.ice_lwr && ri <= L.ice_upr));
}

template <bool PADE>
__device__ __forceinline__ bool lut_cloud(const LutView &L, int b, float lwp, float iwp, float rl, float ri, Opt2 &o)
{
    float lt = 0, lts = 0, lta = 0, ltsg = 0, it = 0, its = 0, ita = 0, itsg = 0;
    bool bad = false;
    if (PADE) {
        if (lwp > 0.0f) {
            bad |= !(rl >= L.liq_lwr && rl <= L.liq_upr);
            pade_phase(L, 0, b, lwp, fminf(fmaxf(rl, L.liq_lwr), L.liq_upr), lt, lts, lta, ltsg);
        }
        if (iwp > 0.0f) {
            bad |= !(ri >= L.ice_lwr && ri <= L.ice_upr);
            pade_phase(L, 1, b, iwp, fminf(fmaxf(ri, L.ice_lwr), L.ice_upr), it, its, ita, itsg);
        }
    } else {
    if (lwp > 0.0f) {
        bad |= !(rl >= L.liq_lwr && rl <= L.liq_upr);
        const size_t o = (size_t)b * L.nsize_liq;
        lut_eval(L.extliq + o, L.ssaliq + o, L.coaliq + o, L.asyliq + o, L.nsize_liq, L.liq_lwr, L.liq_istep, lwp,
                 fminf(fmaxf(rl, L.liq_lwr), L.liq_upr), lt, lts, lta, ltsg);
    }
    if (iwp > 0.0f) {
        bad |= !(ri >= L.ice_lwr && ri <= L.ice_upr);
        const size_t o = (size_t)b * L.nsize_ice;
        lut_eval(L.extice + o, L.ssaice + o, L.coaice + o, L.asyice + o, L.nsize_ice, L.ice_lwr, L.ice_istep, iwp,
                 fminf(fmaxf(ri, L.ice_lwr), L.ice_upr), it, its, ita, itsg);
    }
    }
    o.t = lt + it;
    const float ts = lts + its, ta = lta + ita, tsg = ltsg + itsg;
    const float rt = frcp(fmaxf(1.1920929e-7f, o.t));
    o.w = ts * rt;
    o.p = o.t > 0.0f ? ta * rt : 1.0f;
    o.g = tsg * frcp(fmaxf(1.1920929e-7f, ts));
    return bad;
}

// Rain/snow optics band coefficients (rrtmgp_sw_cloud_optics.F90:59-63);
// bands beyond 14 use the last value.
__constant__ float PREC_B0R[14] = {0.496f, 0.466f, 0.437f, 0.416f, 0.391f, 0.374f, 0.352f, 0.183f, 0.048f, 0.012f, 0.f, 0.f, 0.f, 0.f};
__constant__ float PREC_B0S[14] = {0.460f, 0.460f, 0.460f, 0.460f, 0.460f, 0.460f, 0.460f, 0.460f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
__constant__ float PREC_B1S[14] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.62e-5f, 1.62e-5f, 0.f, 0.f, 0.f, 0.f};
__constant__ float PREC_C0R[14] = {0.980f, 0.975f, 0.965f, 0.960f, 0.955f, 0.952f, 0.950f, 0.944f, 0.894f, 0.884f, 0.883f, 0.883f, 0.883f, 0.883f};
__constant__ float PREC_C0S[14] = {0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.700f, 0.700f, 0.700f, 0.700f};

// Floats of the LUT a block stages in LDS: ext, ssa, co-albedo, asy for
// liquid then ice, [band][size] each.
size_t cloud_lut_floats(const LutView &L)
{
    if (L.pade) return 2 * (size_t)L.nband * 48;
    return 4 * (size_t)L.nband * ((size_t)L.nsize_liq + L.nsize_ice);
}

__device__ __forceinline__ const float *lut_table(const LutView &L, int t)
{
    return t == 0 ? L.extliq : t == 1 ? L.ssaliq : t == 2 ? L.coaliq : t == 3 ? L.asyliq
         : t == 4 ? L.extice : t == 5 ? L.ssaice : t == 6 ? L.coaice : L.asyice;
}

// One thread per (layer, column) item, column fastest, all bands in a loop:
// the CCPP inputs of a wavefront's items are consecutive, they are read once
// for all bands, and its [l][d][b] outputs form one contiguous block.  The
// LUT is staged in LDS once per block, so an item waits on one round of
// global loads.  Items are strided over a grid of a few blocks per CU.
template <bool PADE>
__global__ __launch_bounds__(256) void cloud_optics_kernel(LutView L, CloudArgs A, int ncols, unsigned n)
{
    extern __shared__ float slut[];
    const int nl = L.nband * L.nsize_liq, ni = L.nband * L.nsize_ice, ntot = 4 * (nl + ni);
    LutView S = L;
    if (PADE) {
        const int np = 2 * L.nband * 48;
        for (int k = threadIdx.x; k < np; k += blockDim.x) slut[k] = L.pcoef[k];
        S.pcoef = slut;
    } else {
        constexpr int NLD = 4;   // staging loads in flight per thread
        for (int k0 = threadIdx.x; k0 < ntot; k0 += NLD * blockDim.x) {
            float v[NLD];
#pragma unroll
            for (int u = 0; u < NLD; ++u) {
                const int k = min(k0 + u * (int)blockDim.x, ntot - 1);
                const int t = k < 4 * nl ? k / nl : 4 + (k - 4 * nl) / ni;
                const int o = k < 4 * nl ? k - t * nl : k - 4 * nl - (t - 4) * ni;
                v[u] = lut_table(L, t)[o];
            }
#pragma unroll
            for (int u = 0; u < NLD; ++u)
                if (k0 + u * (int)blockDim.x < ntot) slut[k0 + u * blockDim.x] = v[u];
        }
        S.extliq = slut; S.ssaliq = slut + nl; S.coaliq = slut + 2 * nl; S.asyliq = slut + 3 * nl;
        S.extice = slut + 4 * nl; S.ssaice = S.extice + ni; S.coaice = S.extice + 2 * ni; S.asyice = S.extice + 3 * ni;
    }
    __syncthreads();

    const int nb = A.nband;
    bool bad = false;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int l = (int)(i / (unsigned)ncols), d = (int)(i - (unsigned)l * (unsigned)ncols);
        const int od = A.col0 + d;
        const int c = A.col_idx ? A.col_idx[od] - 1 : od;
        const int blk = od / A.blksz;
        const size_t q = (size_t)c + (size_t)A.ncol_ld * l;
        const size_t o0 = (size_t)i * nb;
        // The solvers read a record only where the McICA mask has a bit, and
        // the mask has none in a block without cloud or in a cloud-free layer
        // (every overlap method); those records are not written.  The SW
        // band-11 depth (cldtausw) is still formed for its block.
        if (!A.blkflag[blk]) continue;
        const double cf = A.cld_frac[q];
        const float lwp = (float)A.lwp[q], iwp = (float)A.iwp[q];
        const float rl = (float)A.reliq[q], ri = (float)A.reice[q];
        if (!(cf > 0.0) && !(!A.is_lw && A.cldtau && blk == A.cldtau_blk)) {
            // the range check of every layer's condensate still applies
            bad |= lut_out_of_range(S, lwp, iwp, rl, ri);
            for (int k = 0; k < 2; ++k) {
                const double *const *src = k == 0 ? A.cnv : A.pbl;
                if (src[0]) bad |= lut_out_of_range(S, (float)src[0][q], (float)src[1][q], (float)src[2][q], (float)src[3][q]);
            }
            continue;
        }
        // precipitation, band independent parts
        float tr = 0.0f, tsn = 0.0f, res = 0.0f;
        bool prec = false;
        if (A.is_lw) {
            if (cf > 1.0e-6) {  // rrtmgp_lw_main.F90:413
                tr = 0.33e-3f * (float)A.rwp[q];
                const double sw = A.swp[q], rs = A.resnow[q];
                tsn = (sw > 0.0 && rs > 10.0) ? 1.5f * 1.05756f * (float)sw / (float)rs : 0.0f;
                prec = true;
            }
        } else {
            if (cf > 1.0e-12) {  // rrtmgp_sw_main.F90:438
                tr = (float)A.rwp[q] * 3.07e-3f;
                const double sw = A.swp[q], rs = A.resnow[q];
                res = (float)rs;
                tsn = (sw > 0.0 && rs > 10.0) ? (float)sw * 1.09087f * (0.0f + 1.5f / (1.0315f * res)) : 0.0f;
                prec = true;
            }
        }
        for (int b = 0; b < nb; ++b) {
            Opt2 o;
            bad |= lut_cloud<PADE>(S, b, lwp, iwp, rl, ri, o);
            if (!A.is_lw && A.cldtau && b == 10 && blk == A.cldtau_blk) A.cldtau[q] = (double)o.t;
            for (int k = 0; k < 2; ++k) {
                const double *const *src = k == 0 ? A.cnv : A.pbl;
                if (!src[0]) continue;
                Opt2 o2;
                bad |= lut_cloud<PADE>(S, b, (float)src[0][q], (float)src[1][q], (float)src[2][q], (float)src[3][q], o2);
                inc2s(o2, o);
            }
            Opt2 pr = {0.0f, 0.0f, 1.0f, 0.0f};
            if (A.is_lw) {
                pr.t = prec ? tr + tsn : 0.0f;
            } else if (prec) {
                const int bb = b < 14 ? b : 13;
                const float sr = tr * (1.0f - PREC_B0R[bb]);
                const float ar = sr * PREC_C0R[bb];
                const float sn = tsn * (1.0f - (PREC_B0S[bb] + PREC_B1S[bb] * 1.0315f * res));
                const float an = sn * PREC_C0S[bb];
                const float tp = fmaxf(1.0e-12f, tr + tsn);
                const float sp = fmaxf(1.0e-12f, sr + sn);
                const float ap = fmaxf(1.0e-12f, ar + an);
                const float asyw = ap * frcp(fmaxf(1.0e-12f, sp));   // < 1: ar, an are below sr, sn
                const float rtp = frcp(tp);
                const float ssaw = fminf(1.0f - 0.000001f, sp * rtp);
                const float omw = fmaxf(1.0f - (1.0f - 0.000001f), (tp - sp) * rtp);   // 1 - ssaw
                const float za1 = asyw * asyw, za2 = ssaw * za1;
                const float rz = frcp(1.0f - za2);
                pr.t = (1.0f - za2) * tp;
                pr.w = (ssaw - za2) * rz;
                pr.p = omw * rz;
                pr.g = asyw * frcp(1.0f + asyw);
            }
            inc2s(pr, o);
            // LW without scattering, 1scl += 2str: absorption depth tau * (1 - ssa);
            // SW and LW with scattering: tau, co-albedo, g in one 12-byte record
            if (A.ctau) A.ctau[o0 + b] = o.t * o.p;
            else A.copt[o0 + b] = make_float3(o.t, o.p, o.g);
        }
    }
    if (bad) atomicOr(A.err, 2);
}

// ---------------------------------------------------------------------------
// McICA
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// x / 4294967295 correctly rounded, for integral 0 <= x < 2^32, without a
// division: q0 = x 2^-32 (1 + 2^-32) is within an ulp of the quotient, the
// residual x - q0 D is exact under the FMA, and one correction step rounds
// it.  Equal to the IEEE quotient for every such x (checked exhaustively:
// tests/test_mt_uniform.py).
__device__ __forceinline__ double div_2p32m1(double x)
{
    const double S = 2.3283064365386963e-10;   // 2^-32
    const double v = x * S;                    // exact
    const double q0 = fma(v, S, v);
    const double rem = fma(-q0, 4294967295.0, x);
    return fma(rem, 1.0 / 4294967295.0, q0);
}

// random_number_t's conversion of a tempered draw to [0, 1]: non-negative
// draws pass through default REAL first
__device__ __forceinline__ double mt_uniform(uint32_t y)
{
    const int32_t ys = (int32_t)y;
    return div_2p32m1(ys < 0 ? (double)ys + 4294967296.0 : (double)(float)ys);
}

__global__ __launch_bounds__(64) void mcica_kernel(McicaArgs A)
{
    extern __shared__ unsigned char smem_raw[];
    uint32_t *mt = (uint32_t *)smem_raw;              // [624]
    double *cf = (double *)(mt + 624);                 // [nlay]
    const int d = blockIdx.x, lane = threadIdx.x;
    const int od = A.col0 + d;
    const int c = A.col_idx ? A.col_idx[od] - 1 : od;
    const int nlay = A.nlay, ngpt = A.ngpt;
    unsigned *mout = A.mask + (size_t)d * nlay * MASK_WORDS;
    if (!A.blkflag[od / A.blksz]) {
        for (int i = lane; i < nlay * MASK_WORDS; i += 64) mout[i] = 0u;
        return;
    }
    const int iovr = A.iovr;
    const bool plain = (iovr == A.iovr_maxrand || iovr == A.iovr_rand || iovr == A.iovr_max);
    const bool expo = (iovr == A.iovr_exp || iovr == A.iovr_exprand);
    const bool dcorr = (iovr == A.iovr_dcorr);
    const bool vmax = (iovr == A.iovr_max);
    for (int l = lane; l < nlay; l += 64) cf[l] = A.cld_frac[(size_t)c + (size_t)A.ncol_ld * l];
    int seed = 0;
    if (A.isubc == 1) seed = ngpt + c + 1;
    else if (A.isubc == 2) seed = A.icseed[c];
    if (seed == 0) seed = 4357;
    for (int i = lane; i < 624; i += 64) mt[i] = (uint32_t)seed * A.mtpow[i];
    __syncthreads();
    // First and last cloudy layer.  No overlap method sets a mask bit outside
    // them (the decorrelation method compares draws in [0, 1] with 1 - cf = 1
    // in clear layers), and every column seeds its own stream, so the stream
    // is not advanced past the last cloudy layer: the words above it are 0,
    // and a column without cloud draws nothing.
    int fst = -1, lst = -1;
    for (int l = 0; l < nlay; ++l)
        if (cf[l] > 0.0) { if (fst < 0) fst = l; lst = l; }
    for (int i = lane + (lst + 1) * MASK_WORDS; i < nlay * MASK_WORDS; i += 64) mout[i] = 0u;
    // twist the state in place, 64 words at a time (reads of a chunk before
    // its writes)
    auto twist = [&]() {
        for (int k0 = 0; k0 < 624; k0 += 64) {
            const int kk = k0 + lane;
            uint32_t nv = 0;
            if (kk < 624) {
                const uint32_t y = (mt[kk] & 0x80000000u) | (mt[kk == 623 ? 0 : kk + 1] & 0x7fffffffu);
                nv = mt[kk < 227 ? kk + 397 : kk - 227] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            __syncthreads();
            if (kk < 624) mt[kk] = nv;
            __syncthreads();
        }
    };
    int cur = -1;  // number of the twisted block the state holds (draws 624 cur ... 624 cur + 623)
    double lr[2] = {0.0, 0.0}, pr[2] = {0.0, 0.0};
    for (int l = 0; l <= lst; ++l) {
        const int base = vmax ? 0 : l * ngpt;
        const int need = (base + ngpt - 1) / 624;
        // the lane's draws (tempered on read): those of the block held now,
        // then, if the layer reaches into the next block, twist and the rest
        uint32_t yv[2] = {0u, 0u};
        if (cur < need) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = base + lane + 64 * h;
                if (lane + 64 * h < ngpt && k / 624 == cur) yv[h] = mt_temper(mt[k - 624 * cur]);
            }
            while (cur < need) { twist(); ++cur; }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = base + lane + 64 * h;
            if (lane + 64 * h < ngpt && k / 624 == cur) yv[h] = mt_temper(mt[k - 624 * cur]);
        }
        const double cfl = cf[l];
        const bool cl = cfl > 0.0;
        const bool clp = l > 0 && cf[l - 1] > 0.0;
        const double thr = 1.0 - cfl;
        unsigned long long bits[2] = {0ull, 0ull};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int g = lane + 64 * h;
            bool bit = false;
            if (g < ngpt && (cl || dcorr)) {
                const double r = mt_uniform(yv[h]);
                if (plain) {
                    if (cl) {
                        if (!clp) lr[h] = r;
                        bit = lr[h] > thr;
                    }
                } else if (expo) {
                    if (cl) {
                        if (clp) {
                            const double rho = A.overlap[(size_t)c + (size_t)A.ncol_ld * (l - 1)];
                            lr[h] = rho * (lr[h] - 0.5) + sqrt(1.0 - rho * rho) * (r - 0.5) + 0.5;
                        } else {
                            lr[h] = r;
                        }
                        bit = lr[h] > thr;
                    }
                } else if (dcorr) {
                    if (l == fst) bit = r > thr;
                    else if (fst >= 0 && l > fst) {
                        const double ov = A.overlap[(size_t)c + (size_t)A.ncol_ld * (l - 1)];
                        bit = (r <= ov) ? (pr[h] > thr) : (r > thr);
                    }
                    pr[h] = r;
                }
            }
            bits[h] = __ballot(bit);
        }
        if (lane == 0) {
            unsigned *w = mout + (size_t)l * MASK_WORDS;
            w[0] = (unsigned)bits[0]; w[1] = (unsigned)(bits[0] >> 32);
            w[2] = (unsigned)bits[1]; w[3] = (unsigned)(bits[1] >> 32);
        }
    }
}

// Per scheme block: any column with total cloud cover above eps
// (rrtmgp_lw_main.F90:353-362, rrtmgp_sw_main.F90:263-273).
__global__ void block_flag_kernel(int ncols, int nlay, int ncol_ld, const int *col_idx, int blksz,
                                  const double *cld_frac, unsigned char *blkflag, unsigned char *colflag)
{
    const int od = blockIdx.x * blockDim.x + threadIdx.x;
    if (od >= ncols) return;
    const int c = col_idx ? col_idx[od] - 1 : od;
    double z0 = 1.0;
    for (int l = 0; l < nlay; ++l) z0 = fmin(z0, 1.0 - cld_frac[(size_t)c + (size_t)ncol_ld * l]);
    if (z0 <= 1.0e-12) z0 = 0.0;
    if (z0 > 1.0 - 1.0e-6) z0 = 1.0;
    const double zcf1 = 1.0 - z0;
    if (colflag) colflag[od] = zcf1 > 1.0e-6;
    if (zcf1 > 1.0e-6) blkflag[od / blksz] = 1;
}

__global__ void zero_u8(unsigned char *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_block_flags(int ncols, int nlay, int ncol_ld, const int *col_idx, int blksz,
                        const double *cld_frac, unsigned char *blkflag, unsigned char *colflag, hipStream_t st)
{
    const int nblk = (ncols + blksz - 1) / blksz;
    hipLaunchKernelGGL(zero_u8, dim3((nblk + 255) / 256), dim3(256), 0, st, blkflag, (size_t)nblk);
    if (ncols > 0)
        hipLaunchKernelGGL(block_flag_kernel, dim3((ncols + 255) / 256), dim3(256), 0, st, ncols, nlay, ncol_ld,
                           col_idx, blksz, cld_frac, blkflag, colflag);
}

void launch_cloud_optics(const LutView &L, const CloudArgs &A, int ncols, hipStream_t st)
{
    const size_t n = (size_t)ncols * A.nlay;
    if (!n) return;
    const size_t nblk = std::min<size_t>((n + 255) / 256, 2048);
    if (L.pade)
        hipLaunchKernelGGL(cloud_optics_kernel<true>, dim3((unsigned)nblk), dim3(256),
                           sizeof(float) * cloud_lut_floats(L), st, L, A, ncols, (unsigned)n);
    else
        hipLaunchKernelGGL(cloud_optics_kernel<false>, dim3((unsigned)nblk), dim3(256),
                           sizeof(float) * cloud_lut_floats(L), st, L, A, ncols, (unsigned)n);
}

void launch_mcica(const McicaArgs &A, int ncols, hipStream_t st)
{
    if (ncols <= 0) return;
    const size_t smem = sizeof(uint32_t) * 624 + sizeof(double) * A.nlay;
    hipLaunchKernelGGL(mcica_kernel, dim3(ncols), dim3(64), smem, st, A);
}

}  // namespace ccpprad
