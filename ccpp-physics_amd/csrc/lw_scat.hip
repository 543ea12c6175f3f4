// LW two-stream solver with scattering on gfx950 (the doGP_lwscat path):
// FP32 transport, FP64 band/broadband reduction and FP64 heating rate.
//
// Restates rte_lw with two-stream optical properties -> lw_solver_2stream:
// lw_combine_sources, lw_two_stream (Fu et al. 1997 coefficients, diffusivity
// secant 1.66), lw_source_2str (Toon et al. 1989 linear-in-tau source) and
// adding, as called at physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:552-575,
// with the clear sky incremented into the McICA-sampled clouds
// (increment_2stream_by_1scalar, :554-555), the CCPP band sum (:604-605) and
// compute_heating_rate (GFS_rrtmgp_post.F90:159-171).
//
// One block = one column, blockDim 128, lane = g-point, in two passes with
// no per-layer scratch:
//   1 bottom-up: layer terms and the adding albedo A and upward source S
//                below each level; A, S kept every 8 layers (LDS)
//   2 top-down : per segment the layer terms and A, S below each layer
//                rebuilt in registers from the checkpoint at the segment
//                bottom, then the diffuse fluxes.
// The level sums of both wavefronts go into one zeroed LDS slot per level
// (two FP64 adds onto 0: the order does not change the result).
#include "kernels.h"

namespace ccpprad {

constexpr int L2_SEG = 8;   // 2 L2_SEG = 16 level sums per segment, reduced together

struct Lw2Layer {
    float R, T, su, sd;   // diffuse reflectance, transmittance, sources leaving the top / bottom
};

// One layer of lw_two_stream + lw_source_2str: total depth t, single
// scattering albedo w and co-albedo p = 1 - w (both from the optical depths,
// so p keeps its digits as w -> 1), asymmetry g, level sources st (top) and
// sb (bottom).  k^2 = (g1 - g2)(g1 + g2) with g1 - g2 = D p and
// g1 + g2 = D (1 - w g).  The source term Z (1 + R - T), with
// Z = (sb - st) / (t (g1 + g2)), is large over thin layers and cancels; with
// E1 = 1 - e^{-kt}, E2 = 1 - e^{-2kt} = E1 (2 - E1) and phi = E1 / (kt),
//   1 + R - T = RT (k E1^2 + (g1 + g2) E2),   RT = 1 / (k (1 + e^{-2kt}) + g1 E2)
//   Z (1 + R - T) = (sb - st) (RT k) phi (k^2 t phi / (g1 + g2) + 2 - E1),
// which stays O(sb - st) for any t.  phi comes from its series
// sum (-x)^n / (n+1)! below x = kt = 0.25 (7 terms, truncation below 2e-9).
__device__ __forceinline__ Lw2Layer lw2_layer(float t, float w, float p, float g, float st, float sb)
{
    constexpr float D = 1.66f, PI = 3.14159265358979f;
    const float g1 = D * (1.0f - 0.5f * w * (1.0f + g));
    const float g2 = D * 0.5f * w * (1.0f - g);
    const float gs = D * (1.0f - w * g);
    const float k = __builtin_amdgcn_sqrtf(fmaxf(D * D * p * (1.0f - w * g), 1.0e-12f));
    const float x = k * t;
    const float e = exp_neg(x);
    const float ps = 1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f - x * (1.0f / 120.0f -
                     x * (1.0f / 720.0f - x * (1.0f / 5040.0f))))));
    const bool small = x < 0.25f;
    const float E1 = small ? x * ps : 1.0f - e;
    const float phi = small ? ps : E1 * frcp(x);
    const float E2 = E1 * (2.0f - E1);
    const float RT = frcp(k * (1.0f + e * e) + g1 * E2);
    Lw2Layer r;
    r.R = RT * g2 * E2;
    r.T = RT * 2.0f * k * e;
    const float C = (sb - st) * (RT * k) * phi * (k * x * phi * frcp(gs) + 2.0f - E1);
    const bool src = t > 1.0e-8f;   // lw_source_2str: no source below this depth
    r.su = src ? PI * (C + st * (1.0f - r.R) - r.T * sb) : 0.0f;
    r.sd = src ? PI * (-C + sb * (1.0f - r.R) - r.T * st) : 0.0f;
    return r;
}

// checkpoints kept in LDS: segment boundaries 1 .. nseg - 1
static __host__ __device__ inline int l2_nckpt(int nseg) { return nseg > 1 ? nseg - 1 : 0; }

// LDS per block: [2][nlev] FP64 level sums, the column's band Planck function
// at the levels [nlev][nband] and the surface [nband], then per lane the
// albedo and source checkpoints [2][nseg - 1][128].
size_t lw_solve_2str_smem(int nlay, int nband)
{
    const int nseg = (nlay + L2_SEG - 1) / L2_SEG;
    return sizeof(double) * 2 * (size_t)(nlay + 1) + sizeof(float) * (size_t)nband * (nlay + 2) +
           sizeof(float) * 2 * (size_t)l2_nckpt(nseg) * 128;
}

// At 127 layers a block takes 25.6 KB of LDS: 6 blocks per CU, 3 wavefronts
// per SIMD, the residency the waves-per-EU attribute gives the compiler
// (the LDS size is known only at launch).
template <bool CLOUDS>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3, 3))) void lw_solve_2str_kernel(TablesView T,
                                                                                                        LwSolveArgs A)
{
    extern __shared__ double red[];
    const int d = blockIdx.x;
    const int c = A.col0 + d;
    const int g = threadIdx.x, lane = g & 63;
    const int nlay = A.nlay, nlev = nlay + 1, ld = A.ncol_ld, ngpt = T.ngpt, nb = T.nband;
    const int nseg = (nlay + L2_SEG - 1) / L2_SEG;
    const bool act = g < ngpt;
    const int gg = act ? g : 0;
    const int b = T.g2b[gg];
    const int top = A.top_at_1;
    double *rup = red, *rdn = red + nlev;
    float *Blev = (float *)(red + 2 * (size_t)nlev), *Bsfc = Blev + (size_t)nlev * nb;
    float *cA = Bsfc + nb, *cS = cA + (size_t)l2_nckpt(nseg) * 128;
    for (int v = g; v < 2 * nlev; v += blockDim.x) red[v] = 0.0;

    // ---- the column's band Planck function at the levels and the surface
    {
        const float tmin = (float)T.temp_ref_min, rdel = (float)(1.0 / T.totplnk_delta);
        const int ntot = (nlev + 1) * nb;
        for (int i = g; i < ntot; i += blockDim.x) {
            const int row = i / nb, bb = i - row * nb;
            const double tk = row < nlev ? A.t_lev[(size_t)c + (size_t)ld * row] : A.tsfc[c];
            const float tf = (float)tk;
            const float v0 = (tf - tmin) * rdel;
            const int iv = (int)v0;
            const float frac = v0 - (float)iv;
            const int idx = d_min(T.nplanck - 1, d_max(1, iv + 1));
            const float *tab = T.totplnk + (size_t)bb * T.nplanck;
            const float a0 = tab[idx - 1], a1 = tab[idx];
            Blev[i] = a0 + frac * (a1 - a0);   // Blev and Bsfc are consecutive
        }
        __syncthreads();
    }

    const double se = A.semis[c];
    const float emis = (se > 1.0e-6 && se <= 1.0) ? (float)se : 1.0f;
    auto lay = [&](int k) { return top ? k : nlay - 1 - k; };   // k counts layers from the TOA
    auto lev = [&](int k) { return top ? k : nlay - k; };       // level k is the top of layer k
    const int lev_toa = lev(0);

    // column bases (uniform) with 32-bit in-column offsets
    const float *tau_c = A.tau + (size_t)d * nlay * ngpt;
    const float *pf_c = A.pfrac + (size_t)d * nlay * ngpt;
    const unsigned *mk_c = CLOUDS ? A.mask + (size_t)d * nlay * MASK_WORDS : nullptr;
    // cloud records [layer][column][band]; the grid has one block per column
    const float3 *co_c = CLOUDS ? A.copt + (size_t)d * nb : nullptr;
    const unsigned co_ls = gridDim.x * (unsigned)nb;
    const unsigned mbit = 1u << (gg & 31), mword = (unsigned)gg >> 5;
    const bool aer = A.aer1 != nullptr;
    const float *ae_c = aer ? A.aer1 + (size_t)d * nlay * nb : nullptr;   // [l][band]
    const float EPS3 = 3.0f * 1.17549435e-38f;

    // One segment's layer terms.  Every load is issued before the first use,
    // with no branch between them (layers past the column end read the last
    // layer again; their results are not used; aerosol loads, when present,
    // go first so every later wait covers them).  The Planck fractions of the
    // layers either side of the segment give its edge level sources.
    auto seg_terms = [&](int s, Lw2Layer *o) {
        float tau[L2_SEG], pf[L2_SEG + 2], cx[L2_SEG], cy[L2_SEG], cz[L2_SEG];
        unsigned mw[L2_SEG];
        float at[L2_SEG];
        if (aer) {
#pragma unroll
            for (int j = 0; j < L2_SEG; ++j) {
                const int l = lay(min(s * L2_SEG + j, nlay - 1));
                at[j] = ldu(ae_c, (unsigned)l * nb + b);
            }
        }
#pragma unroll
        for (int j = -1; j <= L2_SEG; ++j) {
            const int k = min(max(s * L2_SEG + j, 0), nlay - 1);
            pf[j + 1] = ldu(pf_c, (unsigned)lay(k) * ngpt + gg);
        }
#pragma unroll
        for (int j = 0; j < L2_SEG; ++j) {
            const int l = lay(min(s * L2_SEG + j, nlay - 1));
            tau[j] = ldu(tau_c, (unsigned)l * ngpt + gg);
            if (CLOUDS) {
                mw[j] = ldu(mk_c, (unsigned)l * MASK_WORDS + mword);
                const float3 v = ldu(co_c, (unsigned)l * co_ls + b);
                cx[j] = v.x; cy[j] = v.y; cz[j] = v.z;
            }
        }
#pragma unroll
        for (int j = 0; j < L2_SEG; ++j) {
            const int k = s * L2_SEG + j;
            // increment_2stream_by_1scalar: the clear sky (gas + aerosol,
            // absorbing) added to the sampled cloud
            float tcl = tau[j];
            if (aer) tcl += at[j];
            float tc = 0.0f, pc = 1.0f, gc = 0.0f;
            if (CLOUDS) {
                const bool m = mw[j] & mbit;
                tc = m ? cx[j] : 0.0f; pc = m ? cy[j] : 1.0f; gc = m ? cz[j] : 0.0f;
            }
            const float te = tcl + tc, tsc = tc - tc * pc, tab = tcl + tc * pc;
            const bool pos = te > EPS3;
            const float rte = frcp(pos ? te : 1.0f);
            const float w = pos ? tsc * rte : 0.0f, p = pos ? tab * rte : 1.0f;
            // lw_combine_sources: the geometric mean of the two layers' level
            // sources between layers, the layer's own at the TOA and surface
            const float p0 = pf[j + 1];
            const float mt = k == 0 ? p0 : __builtin_amdgcn_sqrtf(pf[j] * p0);
            const float mb = k >= nlay - 1 ? p0 : __builtin_amdgcn_sqrtf(p0 * pf[j + 2]);
            const int kk = min(k, nlay - 1);
            const float st = Blev[lev(kk) * nb + b] * mt, sb = Blev[lev(kk + 1) * nb + b] * mb;
            o[j] = lw2_layer(te, w, p, gc, st, sb);
        }
    };

    // ---- surface: albedo 1 - emis, source pi emis pfrac(surface layer) B(T_sfc)
    const float asfc = 1.0f - emis;
    const float ssfc = 3.14159265358979f * emis * (pf_c[(unsigned)lay(nlay - 1) * ngpt + gg] * Bsfc[b]);

    // Pass 2's terms of one segment from A, S below it (a, q, advanced to the
    // segment top): per layer, diffuse down below = al * above + be, and A,
    // S below it.
    float al[L2_SEG], be[L2_SEG], Abl[L2_SEG], Sbl[L2_SEG];
    auto pass2_terms = [&](int s, float &a, float &q) {
        Lw2Layer o[L2_SEG];
        seg_terms(s, o);
#pragma unroll
        for (int j = L2_SEG - 1; j >= 0; --j) {
            const int k = s * L2_SEG + j;
            if (k >= nlay) continue;
            const float den = frcp(1.0f - o[j].R * a);
            al[j] = o[j].T * den;
            be[j] = (o[j].R * q + o[j].sd) * den;
            Abl[j] = a; Sbl[j] = q;
            const float an = o[j].R + o[j].T * o[j].T * a * den;
            q = o[j].su + o[j].T * den * (q + a * o[j].sd);
            a = an;
        }
    };

    // ---- pass 1, bottom-up: adding albedo A and upward source S below each level
    float Ab = asfc, Sb = ssfc;
    for (int s = nseg - 1; s >= 1; --s) {
        Lw2Layer o[L2_SEG];
        seg_terms(s, o);
#pragma unroll
        for (int j = L2_SEG - 1; j >= 0; --j) {
            const int k = s * L2_SEG + j;
            if (k >= nlay) continue;
            const float den = frcp(1.0f - o[j].R * Ab);
            const float An = o[j].R + o[j].T * o[j].T * Ab * den;
            const float Sn = o[j].su + o[j].T * den * (Sb + Ab * o[j].sd);
            Ab = An; Sb = Sn;
        }
        cA[(size_t)(s - 1) * 128 + g] = Ab; cS[(size_t)(s - 1) * 128 + g] = Sb;
    }
    // the top segment, last of pass 1 and first of pass 2: its pass-2 terms
    // are formed here once, and A, S reach the TOA
    pass2_terms(0, Ab, Sb);

    // ---- pass 2, top-down: no diffuse flux enters at the TOA
    {
        const double su = wave_sum(act ? (double)Sb : 0.0);   // up at the TOA: the end of pass 1
        if (lane == 0) atomicAdd(&rup[lev_toa], su);
    }
    float dn = 0.0f;
    for (int s = 0; s < nseg; ++s) {
        if (s > 0) {
            float a = asfc, q = ssfc;   // the surface below the last segment
            if (s + 1 < nseg) { a = cA[(size_t)s * 128 + g]; q = cS[(size_t)s * 128 + g]; }
            pass2_terms(s, a, q);
        }
        // fluxes at the levels below the segment's layers: up in fl[0..7],
        // down in fl[8..15], summed over the g-points together
        float fl[2 * L2_SEG];
#pragma unroll
        for (int j = 0; j < L2_SEG; ++j) {
            fl[j] = 0.0f; fl[L2_SEG + j] = 0.0f;
            const int k = s * L2_SEG + j;
            if (k >= nlay) continue;
            dn = al[j] * dn + be[j];
            const float up = dn * Abl[j] + Sbl[j];
            if (act) { fl[j] = up; fl[L2_SEG + j] = dn; }
        }
        const double sum = wave_sum_n<2 * L2_SEG>(fl);
        const int idx = wave_sum_idx<2 * L2_SEG>(lane), k = s * L2_SEG + (idx & (L2_SEG - 1));
        if (wave_sum_writer<2 * L2_SEG>(lane) && k < nlay) atomicAdd(&(idx < L2_SEG ? rup : rdn)[lev(k + 1)], sum);
    }
    __syncthreads();
    // ---- epilogue: fluxes, heating rate (FP64)
    for (int v = g; v < nlev; v += blockDim.x) {
        const double up = rup[v], dnv = rdn[v];
        const size_t q = (size_t)c + (size_t)ld * v;
        A.flux_up[q] = up;
        A.flux_dn[q] = dnv;
        if (A.flux_up2) A.flux_up2[q] = up;
        if (A.flux_dn2) A.flux_dn2[q] = dnv;
    }
    if (A.htr) {
        for (int l = g; l < nlay; l += blockDim.x) {
            const double u0 = rup[l], u1 = rup[l + 1];
            const double d0 = rdn[l], d1 = rdn[l + 1];
            const double p0 = A.p_lev[(size_t)c + (size_t)ld * l], p1 = A.p_lev[(size_t)c + (size_t)ld * (l + 1)];
            A.htr[(size_t)c + (size_t)ld * l] = (u1 - u0 - d1 + d0) * GRAV / (CP_DRY * (p1 - p0));
        }
    }
}

void launch_lw_solve_2str(const TablesView &T, const LwSolveArgs &A, int ncols, bool clouds, hipStream_t st)
{
    if (ncols <= 0) return;
    const size_t smem = lw_solve_2str_smem(A.nlay, T.nband);
    if (clouds) hipLaunchKernelGGL((lw_solve_2str_kernel<true>), dim3(ncols), dim3(128), smem, st, T, A);
    else hipLaunchKernelGGL((lw_solve_2str_kernel<false>), dim3(ncols), dim3(128), smem, st, T, A);
}

}  // namespace ccpprad
