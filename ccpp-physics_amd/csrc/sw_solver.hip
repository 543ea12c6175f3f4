// SW two-stream + adding solver on gfx950, FP32 transport, FP64 reductions.
//
// Restates rte_sw -> sw_solver_2stream: combine_and_reorder_2str (gas +
// Rayleigh), the aerosol and McICA cloud increments
// (increment_2stream_by_2stream; rrtmgp_sw_main.F90:540-544,603-604),
// sw_two_stream (Meador & Weaver 1980 with Zdunkowski coefficients),
// sw_source_2str (direct beam), adding (Shonk & Hogan 2008),
// ty_fluxes_byband%reduce + the CCPP band sum (rrtmgp_sw_main.F90:558-559,617-618)
// and compute_heating_rate on daylit columns (GFS_rrtmgp_post.F90:249-266).
//
// One block = one daylit column, blockDim 128, lane = g-point.  Two passes
// per lane, with no per-layer scratch:
//   1 bottom-up: layer two-stream coefficients and the adding albedo A and
//                source q = Q / F_dir (normalised by the direct beam, so this
//                pass needs none); A, q kept every 8 layers (LDS)
//   2 top-down : direct beam F_dir by products of T_noscat (as the
//                reference); per segment the layer terms and A, q below each
//                layer rebuilt in registers from the checkpoint at the segment
//                bottom; downward fluxes and each layer's absorbed flux
// The upward level fluxes are the downward ones minus the net flux, which
// falls by each layer's absorbed flux from the TOA down (FP64); the heating
// rate is the absorbed flux over the layer's pressure thickness.  Forming the
// absorbed flux from the layer's absorptances, rather than as the difference
// of two FP32 level fluxes that agree to 1e-4 in a thin layer, keeps the
// heating rate's error relative to itself.
// Optical properties carry the co-albedo (1 - ssa) next to ssa, so the
// absorption of nearly conservative layers is not lost to FP32 rounding.
#include "kernels.h"

#include <type_traits>

namespace ccpprad {


// Two-stream optics carried with the co-albedo p = 1 - w (p = tau_abs / tau).
struct Opt2s {
    float t, w, p, g;
};

struct TwoStream {
    float Rd, Td, Rdir, Tdir, T0;   // T0 = exp(-tau/mu0), the direct transmittance
    // absorptances, formed without cancellation: of the direct beam
    // Adir = 1 - T0 - Rdir - Tdir, of diffuse light Adif = 1 - Rd - Td
    float Adir, Adif;
};

// (e^{-a} - e^{-b}) / d for b - a = d * s, s > 0, from ea = e^{-a} and
// eb = e^{-b}, without cancellation: e^{-a} s phi(d s) with
// phi(x) = (1 - e^{-x}) / x from its series where |d s| < 0.25 (truncation
// below 2e-9 relative), the direct difference (whose terms then differ by a
// factor e^{0.25} or more) elsewhere.
__device__ __forceinline__ float diff_ratio(float ea, float eb, float d, float s)
{
    const float x = d * s;
    const float phi = 1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f - x * (1.0f / 120.0f -
                      x * (1.0f / 720.0f - x * (1.0f / 5040.0f))))));
    return fabsf(x) < 0.25f ? ea * s * phi : (ea - eb) * frcp(d);
}

// sw_two_stream for one layer in FP32 (Meador & Weaver 1980, Zdunkowski
// coefficients).  The reference's direct-beam brackets are sums of O(1) terms
// that cancel to O(k tau) under a 1/k prefactor, with a removable singularity
// at k mu0 = 1.  With E1 = 1 - e^{-k tau}, E2 = 1 - e^{-2 k tau},
// F = 1 - e^{-tau/mu0} and G = 1 - e^{-(k + 1/mu0) tau} (each accurate to a
// few ulp) the brackets are rewritten two ways, both exact in the algebra:
//  * away from k mu0 = 1 the O(1) terms cancel explicitly,
//      Rdir: (1+k mu)(a2-k g3) E2 + 2 k (g3 - a2 mu) G
//      Tdir: 2 k (g4 + mu a1)(E1 - F) + (1-k mu)(a1-k g4) E2 (1 - F)
//    so both carry their factor k and keep FP32 relative accuracy;
//  * within 1 % of k mu0 = 1 the factor delta = 1 - k mu0 is divided out of
//    both brackets and of 1 - (k mu0)^2, with
//      T_R = (e^{-2k tau} - e^{-(k+1/mu0) tau}) / delta,
//      T_T = (e^{-k tau} - e^{-tau/mu0}) / delta,
//    both from the exponentials already formed (diff_ratio),
//    Rdir = w RT / (1+k mu) [a2 (2G - E2 - 2 T_R) + k g3 (E2 + 2 T_R)]
//    Tdir = -w RT / (1+k mu) [(a1 - k g4) E2 (1 - F) - 2 k (g4 + mu a1) T_T].
// gamma1 - gamma2 = 2 (1 - w) comes from the co-albedo p.  Away from
// k mu0 = 1 one reciprocal serves RT = 1 / X and RT / (1 - (k mu0)^2).
// The layer's absorptances: of diffuse light
//   1 - Rd - Td = RT (k (1 + e^2) - 2 k e + (g1 - g2) E2) = RT (k E1^2 + 2 p E2),
// a sum of positive terms; of the direct beam F - Rdir - Tdir, which the
// clamps keep >= 0 and whose rounding is relative to F.
//
// G0: the layer's asymmetry is 0 (gas and Rayleigh only).  Then
// g3 = g4 = 1/2 and a1 = a2 = (g1 + g2) / 2, and the terms that vanish are
// left out; every remaining operation is the one the general form performs
// with g = 0 (the dropped ones are exact: x * 0 + y = y, and
// g1 / 2 + g2 / 2 rounds as (g1 + g2) / 2), so both forms give the same
// bits.  Contraction is off so the two forms round alike.
template <bool G0>
__device__ __forceinline__ TwoStream two_stream32(float mu0, float mu0_inv, float c3, const Opt2s &o)
{
#pragma clang fp contract(off)
    const float tau = o.t, w = o.w, p = o.p, g = o.g;
    // (8 - w (5 + 3 g)) / 4, 3 w (1 - g) / 4, (2 - 3 mu0 g) / 4 and 1 - g3
    // (c3 = 3 mu0 / 4 is the column's)
    const float g1 = fmaf(-w, G0 ? 1.25f : fmaf(0.75f, g, 1.25f), 2.0f);
    const float w34 = 0.75f * w;
    const float g2 = G0 ? w34 : fmaf(-w34, g, w34);
    const float g3 = G0 ? 0.5f : fmaf(-c3, g, 0.5f);
    const float g4 = G0 ? 0.5f : fmaf(c3, g, 0.5f);
    const float g12 = g1 + g2;
    const float a1 = G0 ? 0.5f * g12 : fmaf(g1, g4, g2 * g3);
    const float a2 = G0 ? a1 : fmaf(g1, g3, g2 * g4);
    const float k = __builtin_amdgcn_sqrtf(fmaxf((2.0f * p) * g12, 1.0e-12f));
    const float kmu = k * mu0;
    const float dd = fmaf(-kmu, kmu, 1.0f);
    const float s = tau * mu0_inv;
    float e, E1, T0, F;
    exp_pair(tau * k, e, E1);
    exp_pair(s, T0, F);
    const float E2 = E1 * (2.0f - E1);
    const float G = fmaf(-E1, F, E1 + F);
    const float X = fmaf(k, fmaf(e, e, 1.0f), g1 * E2);
    const bool reg = fabsf(dd) >= 1.0e-2f;
    const float r = frcp(reg ? X * dd : X);
    const float RT = reg ? r * dd : r;
    float Rdir, Tdir;
    if (reg) {
        const float RT2 = w * r;
        const float br = fmaf((1.0f + kmu) * fmaf(-k, g3, a2), E2, ((2.0f * k) * fmaf(-a2, mu0, g3)) * G);
        const float bt = fmaf((2.0f * k) * fmaf(mu0, a1, g4), E1 - F,
                              (((1.0f - kmu) * fmaf(-k, g4, a1)) * E2) * (1.0f - F));
        Rdir = RT2 * br;
        Tdir = -RT2 * bt;
    } else {
        const float delta = 1.0f - kmu;
        // e^{-2 k tau} = e^2, e^{-(k tau + s)} = e T0, e^{-k tau} = e, e^{-s} = T0
        const float TR = diff_ratio(e * e, e * T0, delta, s);
        const float TT = diff_ratio(e, T0, delta, s);
        const float pre = (w * RT) * frcp(1.0f + kmu);
        Rdir = pre * fmaf(a2, fmaf(2.0f, G - TR, -E2), (k * g3) * fmaf(2.0f, TR, E2));
        Tdir = -pre * fmaf((fmaf(-k, g4, a1) * E2), 1.0f - F, -(((2.0f * k) * fmaf(mu0, a1, g4)) * TT));
    }
    TwoStream rs;
    // max(0, min(x, hi)) for hi >= 0, one v_med3 each
    rs.Rdir = __builtin_amdgcn_fmed3f(Rdir, 0.0f, F);
    rs.Tdir = __builtin_amdgcn_fmed3f(Tdir, 0.0f, F - rs.Rdir);
    rs.Rd = (RT * g2) * E2;
    rs.Td = (RT * (2.0f * k)) * e;
    rs.T0 = T0;
    rs.Adif = RT * fmaf(k * E1, E1, (2.0f * p) * E2);
    rs.Adir = (F - rs.Rdir) - rs.Tdir;
    return rs;
}

// Layers per segment.  The column is walked in segments of SW_SEG layers;
// per lane only the direct beam at each segment top and the adding albedo /
// source at each segment boundary are kept (LDS), and a segment's layer
// terms are recomputed in registers where they are needed.
constexpr int SW_SEG = 8;   // 2 SW_SEG = 16 level sums per segment, reduced together

// checkpoints kept in LDS: segment boundaries 1 .. nseg - 1 (the TOA and
// surface values stay in registers)
static __host__ __device__ inline int sw_nckpt(int nseg) { return nseg > 1 ? nseg - 1 : 0; }

size_t sw_solve_smem(int nlay, int nband)
{
    const int nseg = (nlay + SW_SEG - 1) / SW_SEG;
    return sizeof(double) * (2 * (size_t)(nlay + 1) + (size_t)nlay + 6 * (size_t)nband) +
           sizeof(float) * 2 * (size_t)sw_nckpt(nseg) * 128 + 2 * sizeof(unsigned long long);
}

template <bool CLOUDS, bool AER>
__global__ __launch_bounds__(128, 4) void sw_solve_kernel(TablesView T, SwSolveArgs A)
{
    // LDS: [2][nlev] level sums and [nlay] layer absorption sums of both
    // wavefronts, [2 waves][3][nband] band partials, then per lane
    // [2][nseg - 1] checkpoints: the adding albedo A (slot X) and the
    // normalised source q (slot Q) at segment boundaries
    extern __shared__ double red[];
    const int d = blockIdx.x;
    const int od = A.col0 + d;
    const int c = A.col_idx ? A.col_idx[od] - 1 : od;
    const int g = threadIdx.x, wave = g >> 6, lane = g & 63;
    const int nlay = A.nlay, nlev = nlay + 1, ld = A.ncol_ld, ngpt = T.ngpt, nb = T.nband;
    const int nseg = (nlay + SW_SEG - 1) / SW_SEG;
    const bool act = g < ngpt;
    const int gg = act ? g : 0;
    const int b = T.g2b[gg];
    const bool cloudy = CLOUDS && A.blkflag[od / A.blksz];
    // the clear-sky pass after an all-sky one: a block without clouds already
    // has its clear-sky fluxes, band sums and heating rate from the all-sky
    // pass (the same clear-segment arithmetic, so the same bits)
    if (A.role == 1 && A.blkflag && !A.blkflag[od / A.blksz]) return;
    const int top = A.top_at_1;
    const int lev_toa = top ? 0 : nlay;
    // level and layer sums of both wavefronts: each slot gets exactly two
    // FP64 adds onto 0, so the total is the same in either order
    double *rup = red, *rdn = red + nlev, *rab = red + 2 * (size_t)nlev;
    double *rb = rab + nlay;  // [wave][3][nband]
    float *cX = (float *)(rb + 6 * (size_t)nb);   // checkpoint of boundary k at k - 1
    float *cQ = cX + (size_t)sw_nckpt(nseg) * 128;
    unsigned long long *sflag = (unsigned long long *)(cQ + (size_t)sw_nckpt(nseg) * 128);  // [2 waves]
    for (int v = g; v < 2 * nlev + nlay; v += blockDim.x) red[v] = 0.0;
    __syncthreads();

    const double mu0d = A.coszen[c];
    const float mu0 = (float)mu0d;
    const float mu0_inv = (float)(1.0 / mu0d);
    const float c3 = 0.75f * mu0;
    const int cat = A.alb_cat[b];
    float adir = 0.0f, adif = 0.0f;
    if (cat == 0) { adir = (float)A.alb_nir_dir[c]; adif = (float)A.alb_nir_dif[c]; }
    else if (cat == 1) {
        adir = (float)(0.5 * (A.alb_nir_dir[c] + A.alb_uv_dir[c]));
        adif = (float)(0.5 * (A.alb_nir_dif[c] + A.alb_uv_dif[c]));
    } else if (cat == 2) { adir = (float)A.alb_uv_dir[c]; adif = (float)A.alb_uv_dif[c]; }

    auto lay = [&](int k) { return top ? k : nlay - 1 - k; };       // k counts layers from the TOA
    auto lev = [&](int k) { return top ? k : nlay - k; };           // level k is the top of layer k

    // Column bases (uniform) with 32-bit in-column offsets.
    const float *ta_c = A.tau_abs + (size_t)d * nlay * ngpt, *tr_c = A.tau_ray + (size_t)d * nlay * ngpt;
    const unsigned *mk_c = CLOUDS ? A.mask + (size_t)d * nlay * MASK_WORDS : nullptr;
    // cloud records [layer][column][band]; the grid has one block per column
    const float3 *co_c = CLOUDS ? A.copt + (size_t)d * nb : nullptr;
    const float3 *ae_c = AER ? A.aer4 + (size_t)d * nlay * nb : nullptr;   // [l][band]
    const unsigned co_ls = gridDim.x * (unsigned)nb;
    const unsigned mbit = 1u << (gg & 31), mword = (unsigned)gg >> 5;

    // Segments with a cloudy g-point in any layer: bit s of segc for segment
    // s (segments from 64 on count as cloudy).  Clear segments skip the mask
    // and cloud loads; both passes walk the same segments.
    unsigned long long segc = 0;
    if (CLOUDS && cloudy) {
        unsigned long long mine = 0;
        for (int k0 = wave * 64; k0 < nlay; k0 += 128) {
            const int k = k0 + lane;
            bool f = false;
            if (k < nlay) {
                const uint4 m = *(const uint4 *)(mk_c + (unsigned)lay(k) * MASK_WORDS);
                f = (m.x | m.y | m.z | m.w) != 0u;
            }
            const unsigned long long bal = __ballot(f);
            for (int i = 0; i < 64 / SW_SEG; ++i) {
                const int sg = k0 / SW_SEG + i;
                if (sg < 64 && ((bal >> (SW_SEG * i)) & ((1ull << SW_SEG) - 1))) mine |= 1ull << sg;
            }
        }
        if (lane == 0) sflag[wave] = mine;
        __syncthreads();
        const unsigned long long both = sflag[0] | sflag[1];
        segc = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(both >> 32)) << 32) |
               (unsigned)__builtin_amdgcn_readfirstlane((unsigned)both);
    }
    auto seg_cloudy = [&](int sg) { return CLOUDS && cloudy && (sg >= 64 || ((segc >> sg) & 1ull)); };

    const float EPS3 = 3.0f * 1.17549435e-38f;
    // combine_and_reorder_2str + aerosol and McICA cloud increments
    // (increment_2stream_by_2stream), accumulated as extinction, scattering,
    // absorption and scattering * g depths and normalised once.
    // One segment's optics.  Every load of the segment is issued before the
    // first use, with no branch between them: layers past the column end
    // read the last layer again (their results are not used), and cloud
    // values are selected by the mask bit rather than loaded under it.
    auto load_seg = [&](int sg, Opt2s *o, auto cl) {
        float ta[SW_SEG], tr[SW_SEG];
        float3 ae[SW_SEG];
        unsigned mw[SW_SEG];
        float3 cv[SW_SEG];
#pragma unroll
        for (int j = 0; j < SW_SEG; ++j) {
            const int l = lay(min(sg * SW_SEG + j, nlay - 1));
            const unsigned og = (unsigned)l * ngpt + gg;
            ta[j] = ldu(ta_c, og); tr[j] = ldu(tr_c, og);
            if (AER) ae[j] = ldu(ae_c, (unsigned)l * nb + b);
            if (cl.value) {
                mw[j] = ldu(mk_c, (unsigned)l * MASK_WORDS + mword);
                cv[j] = ldu(co_c, (unsigned)l * co_ls + b);
            }
        }
#pragma unroll
        for (int j = 0; j < SW_SEG; ++j) {
            float te = ta[j] + tr[j], ts = tr[j], tab = ta[j], tsg = 0.0f;
            if (AER) { te = (ae[j].x + ae[j].y) + te; ts += ae[j].x; tab += ae[j].y; tsg += ae[j].z; }
            if (cl.value) {
                const bool m = mw[j] & mbit;
                const float tc = m ? cv[j].x : 0.0f, pc = m ? cv[j].y : 0.0f, gc = m ? cv[j].z : 0.0f;
                const float sc = tc - tc * pc;
                te = tc + te; ts += sc; tab += tc * pc; tsg += sc * gc;
            }
            o[j].t = te;
            const float rte = frcp(fmaxf(EPS3, te));
            o[j].w = ts * rte;
            o[j].p = te > EPS3 ? tab * rte : 1.0f;
            o[j].g = tsg * frcp(fmaxf(EPS3, ts));
        }
    };
    const float toa = T.toa_norm[gg];

    // ---- pass 1, bottom-up: adding albedo A and the direct-beam source
    // normalised by the direct beam at its level, q = Q / F_dir.  Q is linear
    // in F_dir, so q follows a local recurrence,
    //   q_k = Rdir_k + Td_k den_k (T0_k q_{k+1} + A_{k+1} Tdir_k),
    // bounded by 1 (the upwelling a direct beam can feed), and this pass
    // needs no direct beam.  A and q are kept at segment boundaries.
    float Ab = adif, qb = adir;
    auto pass1_seg = [&](int s, auto cl) {
        Opt2s o[SW_SEG];
        load_seg(s, o, cl);
#pragma unroll
        for (int j = SW_SEG - 1; j >= 0; --j) {
            const int k = s * SW_SEG + j;
            if (k >= nlay) continue;
            const TwoStream ts = two_stream32<!decltype(cl)::value && !AER>(mu0, mu0_inv, c3, o[j]);
            const float den = frcp(1.0f - ts.Rd * Ab);
            const float Aa = ts.Rd + ts.Td * ts.Td * Ab * den;
            const float qa = ts.Rdir + ts.Td * den * (qb * ts.T0 + Ab * ts.Tdir);
            Ab = Aa; qb = qa;
        }
    };
    // Pass 2's terms of one segment, from A, q below it (a, q, advanced to
    // the segment top): per layer, diffuse down below = al * above + be *
    // F_dir above, the direct transmittance T0, and the absorbed flux
    // ab = c1 * (diffuse down above) + c2 * F_dir above.
    float al[SW_SEG], be[SW_SEG], T0[SW_SEG], c1[SW_SEG], c2[SW_SEG];
    auto pass2_terms = [&](int s, auto cl, float &a, float &q) {
        Opt2s o[SW_SEG];
        load_seg(s, o, cl);
#pragma unroll
        for (int j = SW_SEG - 1; j >= 0; --j) {
            const int k = s * SW_SEG + j;
            if (k >= nlay) continue;
            const TwoStream ts = two_stream32<!decltype(cl)::value && !AER>(mu0, mu0_inv, c3, o[j]);
            const float den = frcp(1.0f - ts.Rd * a);
            // dn_dif below = (Td dn_dif + Rd Q + Tdir F_dir) den with Q = q T0 F_dir
            al[j] = ts.Td * den;
            be[j] = (ts.Rd * q * ts.T0 + ts.Tdir) * den;
            T0[j] = ts.T0;
            // absorbed = Adir F_dir + Adif (dn_dif above + up below), with
            // up below = a dn_dif below + q T0 F_dir: sums of non-negative
            // terms, so the absorbed flux keeps its relative accuracy
            c1[j] = ts.Adif * fmaf(a, al[j], 1.0f);
            c2[j] = fmaf(ts.Adif, fmaf(a, be[j], q * ts.T0), ts.Adir);
            const float qa = ts.Rdir + ts.Td * den * (q * ts.T0 + a * ts.Tdir);
            a = ts.Rd + ts.Td * ts.Td * a * den;
            q = qa;
        }
    };
    for (int s = nseg - 1; s >= 1; --s) {
        if (seg_cloudy(s)) pass1_seg(s, std::true_type{});
        else pass1_seg(s, std::false_type{});
        cX[(size_t)(s - 1) * 128 + g] = Ab; cQ[(size_t)(s - 1) * 128 + g] = qb;
    }
    // the top segment, last of pass 1 and first of pass 2: its pass-2 terms
    // are formed here once, and A, q reach the TOA
    if (seg_cloudy(0)) pass2_terms(0, std::true_type{}, Ab, qb);
    else pass2_terms(0, std::false_type{}, Ab, qb);

    // ---- pass 2, top-down: direct beam by products of T_noscat (as the
    // reference) and fluxes.  Per segment the layer terms and A, q below each
    // layer are rebuilt from the checkpoint at the segment bottom.
    float dnd = 0.0f;        // diffuse down at TOA
    float fdir = toa * mu0;  // direct beam at the current level
    {
        const float up = dnd * Ab + qb * fdir;   // A, q at the TOA: the end of pass 1
        const double su = wave_sum(act ? (double)up : 0.0);
        const double sdn = wave_sum(act ? (double)(dnd + fdir) : 0.0);
        if (lane == 0) { atomicAdd(&rup[lev_toa], su); atomicAdd(&rdn[lev_toa], sdn); }
    }
    // direct and total down flux at the surface level (TOA values if the
    // surface level is the TOA)
    float sfc_dn = fdir, sfc_dir = fdir;
    for (int s = 0; s < nseg; ++s) {
        if (s > 0) {
            // the segment's terms, from the checkpoint at its bottom (the
            // surface below the last segment)
            float a = adif, q = adir;
            if (s + 1 < nseg) { a = cX[(size_t)s * 128 + g]; q = cQ[(size_t)s * 128 + g]; }
            if (seg_cloudy(s)) pass2_terms(s, std::true_type{}, a, q);
            else pass2_terms(s, std::false_type{}, a, q);
        }
        // the segment's downward fluxes at the levels below its layers
        // (fl[0..7]) and each layer's absorbed flux, the net flux at its top
        // minus that at its bottom (fl[8..15]), summed over the g-points
        // together.  The absorbed flux comes from the layer's absorptances,
        // so the heating rate of a thin layer does not rest on the difference
        // of two nearly equal FP32 fluxes; the upward level fluxes follow
        // from it in the epilogue.
        float fl[2 * SW_SEG];
#pragma unroll
        for (int j = 0; j < SW_SEG; ++j) {
            fl[j] = 0.0f; fl[SW_SEG + j] = 0.0f;
            const int k = s * SW_SEG + j;
            if (k >= nlay) continue;
            const float fd = T0[j] * fdir;            // direct beam at the level below
            const float ab = fmaf(dnd, c1[j], fdir * c2[j]);
            dnd = al[j] * dnd + be[j] * fdir;
            const float dn = dnd + fd;
            if (lev(k + 1) == A.isfc) { sfc_dn = dn; sfc_dir = fd; }
            fdir = fd;
            if (act) { fl[j] = dn; fl[SW_SEG + j] = ab; }
        }
        const double sum = wave_sum_n<2 * SW_SEG>(fl);
        const int idx = wave_sum_idx<2 * SW_SEG>(lane), k = s * SW_SEG + (idx & (SW_SEG - 1));
        if (wave_sum_writer<2 * SW_SEG>(lane) && k < nlay)
            atomicAdd(idx < SW_SEG ? &rdn[lev(k + 1)] : &rab[lay(k)], sum);
    }
    {
        // per-band surface sums for scmpsw
        const double vdn = act ? (double)sfc_dn : 0.0, vdir = act ? (double)sfc_dir : 0.0;
        for (int bb = 0; bb < nb; ++bb) {
            const bool mine = act && b == bb;
            const double s1 = wave_sum(mine ? vdn : 0.0), s2 = wave_sum(mine ? vdir : 0.0);
            if (lane == 0) { rb[(size_t)wave * 3 * nb + bb] = s1; rb[(size_t)wave * 3 * nb + nb + bb] = s2; }
        }
    }
    __syncthreads();
    // ---- epilogue: the upward fluxes below the TOA from the net flux,
    // net(below layer) = net(above) - absorbed(layer), in FP64
    if (g == 0) {
        double net = rdn[lev_toa] - rup[lev_toa];
        for (int k = 0; k < nlay; ++k) {
            net -= rab[lay(k)];
            rup[lev(k + 1)] = rdn[lev(k + 1)] - net;
        }
    }
    __syncthreads();
    const bool clr_too = (A.role == 1) || !cloudy;  // fluxes equal the clear-sky ones
    for (int v = g; v < nlev; v += blockDim.x) {
        const double up = rup[v], dn = rdn[v];
        const size_t q = (size_t)c + (size_t)ld * v;
        if (A.role == 0) { A.up_all[q] = up; A.dn_all[q] = dn; }
        if (clr_too) { A.up_clr[q] = up; A.dn_clr[q] = dn; }
        else if (A.write_clr_zero) { A.up_clr[q] = 0.0; A.dn_clr[q] = 0.0; }
    }
    for (int bb = g; bb < nb; bb += blockDim.x) {
        const double sdn = rb[bb] + rb[3 * nb + bb], sdir = rb[nb + bb] + rb[3 * nb + nb + bb];
        double *o = A.sb + (size_t)d * 3 * nb;
        if (A.role == 0) { o[bb] = sdn; o[nb + bb] = sdir; }
        if (clr_too) o[2 * nb + bb] = sdn;
    }
    double *h_all = A.role == 0 ? A.htr : nullptr, *h_clr = (clr_too || A.role == 1) ? A.htr_clr : nullptr;
    if (h_all || h_clr) {
        // compute_heating_rate: (net flux at the layer top - at its bottom)
        // g / (cp dp), the flux difference taken as the layer's absorbed flux
        for (int l = g; l < nlay; l += blockDim.x) {
            const double p0 = A.p_lev[(size_t)c + (size_t)ld * l], p1 = A.p_lev[(size_t)c + (size_t)ld * (l + 1)];
            const double hr = rab[l] * GRAV / (CP_DRY * fabs(p1 - p0));
            if (h_all) h_all[(size_t)c + (size_t)ld * l] = hr;
            if (h_clr) h_clr[(size_t)c + (size_t)ld * l] = hr;
        }
    }
}

void launch_sw_solve(const TablesView &T, const SwSolveArgs &A, int ncols, bool clouds, hipStream_t st)
{
    if (ncols <= 0) return;
    const size_t smem = sw_solve_smem(A.nlay, T.nband);
    const bool aer = A.aer4 != nullptr;
    if (clouds && aer)
        hipLaunchKernelGGL((sw_solve_kernel<true, true>), dim3(ncols), dim3(128), smem, st, T, A);
    else if (clouds)
        hipLaunchKernelGGL((sw_solve_kernel<true, false>), dim3(ncols), dim3(128), smem, st, T, A);
    else if (aer)
        hipLaunchKernelGGL((sw_solve_kernel<false, true>), dim3(ncols), dim3(128), smem, st, T, A);
    else
        hipLaunchKernelGGL((sw_solve_kernel<false, false>), dim3(ncols), dim3(128), smem, st, T, A);
}

// scmpsw (cmpfsw_type) from the surface band fluxes, reproducing
// rrtmgp_sw_main.F90:562-672 including its whole-array reset in the
// cloudy-block branch without clear-sky fluxes (:587-591).

__global__ void scmpsw_kernel(ScmpArgs A)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= A.ndays) return;
    const int c = A.col_idx[d] - 1, blk = d / A.blksz, nb = A.nband;
    const bool cloudy = A.blkflag[blk];
    const bool clr = !cloudy || A.do_clr;
    const double *sb = A.sb + (size_t)d * 3 * nb;
    double scl[6] = {0, 0, 0, 0, 0, 0}, sal[6] = {0, 0, 0, 0, 0, 0};
    if (clr) {
        for (int bb = 0; bb < nb; ++bb) {
            const double fdir = sb[2 * nb + bb];
            if (bb + 1 < A.ibd) scl[2] += fdir;
            if (bb + 1 == A.ibd) { scl[2] += fdir * 0.5; scl[4] += fdir * 0.5; }
            if (bb + 1 > A.ibd) scl[4] += fdir;
        }
        scl[0] = sb[2 * nb + A.ibd_uv - 1];
    }
    double o[6];
    if (cloudy) {
        for (int bb = 0; bb < nb; ++bb) {
            const double fdir = sb[nb + bb], fdif = sb[bb] - sb[nb + bb];
            if (bb + 1 < A.ibd) { sal[2] += fdir; sal[3] += fdif; }
            if (bb + 1 == A.ibd) { sal[2] += fdir * 0.5; sal[3] += fdif * 0.5; sal[4] += fdir * 0.5; sal[5] += fdif * 0.5; }
            if (bb + 1 > A.ibd) { sal[4] += fdir; sal[5] += fdif; }
        }
        const bool cc = A.colflag[d];
        const double *src = cc ? sal : scl;
        o[2] = src[2]; o[3] = src[3]; o[4] = src[4]; o[5] = src[5];
        o[0] = cc ? sb[A.ibd_uv - 1] : (clr ? sb[2 * nb + A.ibd_uv - 1] : 0.0);
        o[1] = clr ? sb[2 * nb + A.ibd_uv - 1] : 0.0;
    } else {
        o[2] = scl[2]; o[3] = scl[3]; o[4] = scl[4]; o[5] = scl[5];
        o[0] = scl[0]; o[1] = scl[0];
    }
    if (*A.last_wipe > blk) for (int i = 0; i < 6; ++i) o[i] = 0.0;
    for (int i = 0; i < 6; ++i) A.scmpsw[(size_t)6 * c + i] = o[i];
}

// Night columns: zeroed by the whole-array reset if any happened.
__global__ void scmpsw_night_kernel(int ncol, const unsigned char *daymap, const int *last_wipe, double *scmpsw)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol || daymap[c] || *last_wipe < 0) return;
    for (int i = 0; i < 6; ++i) scmpsw[(size_t)6 * c + i] = 0.0;
}

// last_wipe = max block index b with (cloudy block and no clear-sky request)
__global__ void sw_wipe_kernel(int nblk, int do_clr, const unsigned char *blkflag, int *last_wipe)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    if (blkflag[b] && !do_clr) atomicMax(last_wipe, b);
}

__global__ void daymap_kernel(int nday, const int *idx, unsigned char *daymap)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < nday) daymap[idx[d] - 1] = 1;
}

void launch_scmpsw(const ScmpArgs &A, int ncol, unsigned char *daymap, int *last_wipe, hipStream_t st)
{
    const int nblk = (A.ndays + A.blksz - 1) / A.blksz;
    hipLaunchKernelGGL(sw_wipe_kernel, dim3((nblk + 255) / 256), dim3(256), 0, st, nblk, A.do_clr, A.blkflag, last_wipe);
    hipLaunchKernelGGL(scmpsw_kernel, dim3((A.ndays + 255) / 256), dim3(256), 0, st, A);
    hipLaunchKernelGGL(daymap_kernel, dim3((A.ndays + 255) / 256), dim3(256), 0, st, A.ndays, A.col_idx, daymap);
    hipLaunchKernelGGL(scmpsw_night_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, ncol, daymap, last_wipe, A.scmpsw);
}

// Aerosol optics of a chunk in the solver's order.  The CCPP arrays are
// (ncol, nlay, nband) with the column fastest; the solver reads, per layer,
// the column's nband values together.  Block = (64 columns, one layer): the
// FP64 reads run along the columns, the FP32 records go out through LDS as
// each column's contiguous run of nband records.  A record holds the
// aerosol's increments to the scattering, absorption and scattering x
// asymmetry depths (increment_2stream_by_2stream); their first two add up
// to its extinction depth.
__global__ __launch_bounds__(256) void sw_aer_pack_kernel(SwSolveArgs A, int ncols, int nb, float3 *out)
{
    extern __shared__ float3 tile[];   // [64 columns][nb]
    const int d0 = blockIdx.x * 64, l = blockIdx.y, nlay = A.nlay, ld = A.ncol_ld;
    for (int i = threadIdx.x; i < 64 * nb; i += blockDim.x) {
        const int b = i >> 6, dd = i & 63, d = d0 + dd;
        if (d >= ncols) continue;
        const int od = A.col0 + d;
        const int c = A.col_idx ? A.col_idx[od] - 1 : od;
        const size_t q = (size_t)c + (size_t)ld * (l + (size_t)nlay * b);
        const double sd = A.aer_ssa[q];
        const float at = (float)A.aer_tau[q], aw = (float)sd, ap = (float)(1.0 - sd), ag = (float)A.aer_g[q];
        const float s1 = at * aw;
        tile[dd * nb + b] = make_float3(s1, at * ap, s1 * ag);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * nb; i += blockDim.x) {
        const int dd = i / nb, b = i - dd * nb, d = d0 + dd;
        if (d < ncols) out[((size_t)d * nlay + l) * nb + b] = tile[i];
    }
}

void launch_sw_aer_pack(const SwSolveArgs &A, int ncols, int nband, float3 *out, hipStream_t st)
{
    if (ncols <= 0 || A.nlay <= 0) return;
    hipLaunchKernelGGL(sw_aer_pack_kernel, dim3((ncols + 63) / 64, A.nlay), dim3(256),
                       sizeof(float3) * 64 * (size_t)nband, st, A, ncols, nband, out);
}

}  // namespace ccpprad
