// LW no-scattering solver with Gaussian quadrature on gfx950, FP32 transport,
// FP64 g-point/band reduction and FP64 heating rate.
//
// Restates rte_lw -> lw_solver_noscat(_GaussQuad) -> lw_source_noscat +
// lw_transport_noscat (called at physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:511-525
// and :584-599), compute_optimal_angles (:508-509), the McICA absorption
// increment tau += tau_c*(1-ssa_c) (:579-580), the aerosol increment (:502-504),
// ty_fluxes_byband%reduce + the CCPP band sum (:529-530, :604-605) and
// compute_heating_rate (GFS_rrtmgp_post.F90:159-171).
//
// One block = one column, blockDim = 128 (two wavefronts), lane = g-point.
// The block first tabulates the column's band Planck function at every layer,
// level and the surface in LDS (one interpolation per (row, band) instead of
// one per g-point).  Each lane then carries the NANG angle intensities through
// a top-down sweep (downward intensity) and a bottom-up sweep (upward
// intensity, Jacobian); tau and the Planck fraction are re-read in the second
// sweep rather than stored, so the HBM traffic is two reads of tau/pfrac (8 B
// per cell each) plus the FP64 level fluxes.
// At every level each wavefront sums its lanes' FP32 fluxes in FP64 and adds
// the partial into one zeroed LDS slot per level (two adds onto 0, so the
// result does not depend on their order); the epilogue writes the
// (nCol, nLay+1) fluxes and differences them into the heating rate.
#include "kernels.h"

namespace ccpprad {


// Transmittance e = e^{-x}, absorptance E = 1 - e^{-x} and the
// linear-in-tau source weight f = E/x - e for x >= 0.  Below x = 0.25, f comes
// from its Taylor series sum_{n>=1} (-1)^{n+1} n x^n / (n+1)! (7 terms,
// truncation below 3e-9 relative; the direct form cancels in FP32 there) and
// E = x (f + e) with no cancellation; above, E = 1 - e and f = E/x - e.
__device__ __forceinline__ void lw_trans(float x, float &e, float &E, float &f)
{
    e = exp_neg(x);
    const float fs = x * (0.5f - x * (1.0f / 3.0f - x * (0.125f - x * (1.0f / 30.0f - x * (1.0f / 144.0f -
                     x * (1.0f / 840.0f - x * (1.0f / 5760.0f)))))));
    const bool small = x < 0.25f;
    const float Eb = 1.0f - e;
    f = small ? fs : Eb * frcp(x) - e;
    E = small ? x * (fs + e) : Eb;
}

constexpr int LW_BATCH = 8;   // also the number of level sums reduced together

// LDS per block: [3][nlev] FP64 level sums (up, down, Jacobian) of both
// wavefronts, then the column's band Planck function at layers
// [nlay][nband], levels [nlev][nband] and the surface [2][nband] (at T_sfc
// and T_sfc + 1 K).
size_t lw_solve_smem(int nlay, int nband)
{
    return sizeof(double) * 3 * (size_t)(nlay + 1) + sizeof(float) * (size_t)nband * (2 * (size_t)nlay + 3);
}

// At 127 layers and 16 bands a block takes 19.5 KB of LDS, so a CU holds 8
// blocks: 4 wavefronts per SIMD.  The LDS size is known only at launch, so
// the waves-per-EU attribute tells the compiler that residency; scheduled
// for it (more loads in flight, up to 128 VGPRs) the 3-angle all-sky solver
// takes 22.7 instead of 24.0 ms per 262,144-column step.
template <int NANG, bool CLOUDS, bool JAC, bool OPTANG, bool AER>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 4))) void lw_solve_kernel(TablesView T,
                                                                                                   LwSolveArgs A)
{
    extern __shared__ double red[];
    const int d = blockIdx.x;
    const int c = A.col0 + d;
    const int g = threadIdx.x;
    const int wave = g >> 6, lane = g & 63;
    const int nlay = A.nlay, nlev = nlay + 1, ld = A.ncol_ld, ngpt = T.ngpt, nb = T.nband;
    const bool act = g < ngpt;
    const int gg = act ? g : 0;
    const int b = T.g2b[gg];
    // level sums of both wavefronts: each slot gets exactly two FP64 adds onto
    // 0, so the total is the same in either order
    double *rup = red, *rdn = red + nlev, *rjc = red + 2 * (size_t)nlev;
    float *Blay = (float *)(red + 3 * (size_t)nlev), *Blev = Blay + (size_t)nlay * nb, *Bsfc = Blev + (size_t)nlev * nb;
    for (int v = g; v < 3 * nlev; v += blockDim.x) red[v] = 0.0;

    // ---- the column's band Planck function (interpolate1D of totplnk),
    // shared by the g-points of a band
    {
        const float tmin = (float)T.temp_ref_min, rdel = (float)(1.0 / T.totplnk_delta);
        const int ntot = (2 * nlay + 3) * nb;
        for (int i = g; i < ntot; i += blockDim.x) {
            const int row = i / nb, bb = i - row * nb;
            double tk;
            if (row < nlay) tk = A.t_lay[(size_t)c + (size_t)ld * row];
            else if (row < nlay + nlev) tk = A.t_lev[(size_t)c + (size_t)ld * (row - nlay)];
            else tk = A.tsfc[c] + (row == 2 * nlay + 2 ? 1.0 : 0.0);
            const float tf = (float)tk;
            const float v0 = (tf - tmin) * rdel;
            const int iv = (int)v0;
            const float frac = v0 - (float)iv;
            const int idx = d_min(T.nplanck - 1, d_max(1, iv + 1));
            const float *tab = T.totplnk + (size_t)bb * T.nplanck;
            const float a0 = tab[idx - 1], a1 = tab[idx];
            Blay[i] = a0 + frac * (a1 - a0);   // Blay, Blev, Bsfc are consecutive
        }
        __syncthreads();
    }

    double se = A.semis[c];
    const float emis = (se > 1.0e-6 && se <= 1.0) ? (float)se : 1.0f;
    // column bases (uniform) with 32-bit in-column offsets
    const float *tau_c = A.tau + (size_t)d * nlay * ngpt;
    const float *pf_c = A.pfrac + (size_t)d * nlay * ngpt;
    const unsigned *mk_c = CLOUDS ? A.mask + (size_t)d * nlay * MASK_WORDS : nullptr;
    // cloud depths [layer][column][band]; the grid has one block per column
    const float *ct_c = CLOUDS ? A.ctau + (size_t)d * nb : nullptr;
    const unsigned ct_ls = gridDim.x * (unsigned)nb;
    const unsigned mbit = 1u << (gg & 31), mword = (unsigned)gg >> 5;
    const float *ae_c = AER ? A.aer1 + (size_t)d * nlay * nb : nullptr;   // [l][band]

    // Per-layer inputs, loaded for LW_BATCH layers at a time with no load
    // depending on another (the cloud depth is fetched unconditionally and
    // selected by the mask bit), so a batch costs one memory latency.
    struct LayerIn { float tau, pf, ct, aer; unsigned mw; };
    auto load_layer = [&](int l, LayerIn &x) {
        x.tau = ldu(tau_c, (unsigned)l * ngpt + gg);
        x.pf = ldu(pf_c, (unsigned)l * ngpt + gg);
        x.aer = AER ? ldu(ae_c, (unsigned)l * nb + b) : 0.0f;
        if (CLOUDS) { x.mw = ldu(mk_c, (unsigned)l * MASK_WORDS + mword); x.ct = ldu(ct_c, (unsigned)l * ct_ls + b); }
    };
    auto layer_tau = [&](const LayerIn &x) -> float {
        float t = x.tau;
        if (AER) t += x.aer;
        if (CLOUDS) { if (x.mw & mbit) t += x.ct; }
        return t;
    };

    float D[NANG], W2pi[NANG];
#pragma unroll
    for (int a = 0; a < NANG; ++a) { D[a] = A.D[a]; W2pi[a] = 6.283185307179586f * A.W[a]; }
    if (OPTANG) {
        // compute_optimal_angles: Ds = fit1 * exp(-column tau) + fit2, weight 0.5
        float s = 0.0f;
        for (int l = 0; l < nlay; ++l) { LayerIn x; load_layer(l, x); s += layer_tau(x); }
        D[0] = T.opt_fit[2 * b] * expf(-s) + T.opt_fit[2 * b + 1];
        W2pi[0] = 6.283185307179586f * 0.5f;
    }

    const int top = A.top_at_1;
    const int lev_toa = top ? 0 : nlay, lev_sfc = top ? nlay : 0;
    // ---- top-down sweep: downward intensity
    float I[NANG];
#pragma unroll
    for (int a = 0; a < NANG; ++a) I[a] = 0.0f;
    for (int k0 = 0; k0 < nlay; k0 += LW_BATCH) {
        LayerIn x[LW_BATCH];
#pragma unroll
        for (int j = 0; j < LW_BATCH; ++j)
            if (k0 + j < nlay) load_layer(top ? k0 + j : nlay - 1 - (k0 + j), x[j]);
        float fl[LW_BATCH];   // g-point fluxes at the batch's lower levels, summed together
#pragma unroll
        for (int j = 0; j < LW_BATCH; ++j) {
            fl[j] = 0.0f;
            const int k = k0 + j;
            if (k >= nlay) continue;
            const int l = top ? k : nlay - 1 - k;
            const int lev_below = top ? l + 1 : l;
            const float tau = layer_tau(x[j]);
            const float pf = x[j].pf;
            const float lay = pf * Blay[l * nb + b], ldn = pf * Blev[lev_below * nb + b];
            const float dsrc = lay - ldn;
            float fsum = 0.0f;
#pragma unroll
            for (int a = 0; a < NANG; ++a) {
                const float xa = tau * D[a];
                float t, omt, f;
                lw_trans(xa, t, omt, f);
                I[a] = t * I[a] + (omt * ldn + 2.0f * f * dsrc);
                fsum += W2pi[a] * I[a];
            }
            if (act) fl[j] = fsum;
        }
        const double sum = wave_sum_n<LW_BATCH>(fl);
        const int k = k0 + wave_sum_idx<LW_BATCH>(lane);
        if (wave_sum_writer<LW_BATCH>(lane) && k < nlay) atomicAdd(&rdn[top ? k + 1 : nlay - 1 - k], sum);
    }
    // ---- surface
    const int sfc_lay = top ? nlay - 1 : 0;
    const float pfs = pf_c[(unsigned)sfc_lay * ngpt + gg];
    const float B0 = Bsfc[b];
    const float sfc_src = pfs * B0;
    float J[NANG];
    {
        float fsum = 0.0f, jsum = 0.0f;
#pragma unroll
        for (int a = 0; a < NANG; ++a) {
            I[a] = I[a] * (1.0f - emis) + emis * sfc_src;
            fsum += W2pi[a] * I[a];
            if (JAC) {
                J[a] = emis * (pfs * (Bsfc[nb + b] - B0));
                jsum += W2pi[a] * J[a];
            }
        }
        const double ws = wave_sum(act ? (double)fsum : 0.0);
        if (lane == 0) atomicAdd(&rup[lev_sfc], ws);
        if (JAC) {
            const double wj = wave_sum(act ? (double)jsum : 0.0);
            if (lane == 0) atomicAdd(&rjc[lev_sfc], wj);
        }
    }
    // ---- bottom-up sweep: upward intensity (and Jacobian)
    for (int k0 = 0; k0 < nlay; k0 += LW_BATCH) {
        LayerIn x[LW_BATCH];
#pragma unroll
        for (int j = 0; j < LW_BATCH; ++j)
            if (k0 + j < nlay) load_layer(top ? nlay - 1 - (k0 + j) : k0 + j, x[j]);
        float fl[2 * LW_BATCH];   // g-point fluxes (and Jacobians) at the batch's upper levels
#pragma unroll
        for (int j = 0; j < LW_BATCH; ++j) {
            fl[j] = 0.0f; fl[LW_BATCH + j] = 0.0f;
            const int k = k0 + j;
            if (k >= nlay) continue;
            const int l = top ? nlay - 1 - k : k;
            const int lev_above = top ? l : l + 1;
            const float tau = layer_tau(x[j]);
            const float pf = x[j].pf;
            const float lay = pf * Blay[l * nb + b], lup = pf * Blev[lev_above * nb + b];
            const float dsrc = lay - lup;
            float fsum = 0.0f, jsum = 0.0f;
#pragma unroll
            for (int a = 0; a < NANG; ++a) {
                const float xa = tau * D[a];
                float t, omt, f;
                lw_trans(xa, t, omt, f);
                I[a] = t * I[a] + (omt * lup + 2.0f * f * dsrc);
                fsum += W2pi[a] * I[a];
                if (JAC) { J[a] = t * J[a]; jsum += W2pi[a] * J[a]; }
            }
            if (act) { fl[j] = fsum; fl[LW_BATCH + j] = jsum; }
        }
        if (JAC) {
            const double sum = wave_sum_n<2 * LW_BATCH>(fl);
            const int idx = wave_sum_idx<2 * LW_BATCH>(lane), k = k0 + (idx & (LW_BATCH - 1));
            if (wave_sum_writer<2 * LW_BATCH>(lane) && k < nlay)
                atomicAdd(&(idx < LW_BATCH ? rup : rjc)[top ? nlay - 1 - k : k + 1], sum);
        } else {
            const double sum = wave_sum_n<LW_BATCH>(*(const float(*)[LW_BATCH])fl);
            const int k = k0 + wave_sum_idx<LW_BATCH>(lane);
            if (wave_sum_writer<LW_BATCH>(lane) && k < nlay) atomicAdd(&rup[top ? nlay - 1 - k : k + 1], sum);
        }
    }
    __syncthreads();
    // ---- epilogue: write fluxes, heating rate (FP64)
    for (int v = g; v < nlev; v += blockDim.x) {
        const double up = rup[v];
        const double dn = rdn[v];
        const size_t q = (size_t)c + (size_t)ld * v;
        A.flux_up[q] = up;
        A.flux_dn[q] = dn;
        if (A.flux_up2) A.flux_up2[q] = up;
        if (A.flux_dn2) A.flux_dn2[q] = dn;
        if (JAC && A.flux_jac) A.flux_jac[q] = rjc[v];
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

template <int NANG, bool CLOUDS, bool JAC, bool OPTANG>
static void launch_one(const TablesView &T, const LwSolveArgs &A, int ncols, hipStream_t st)
{
    const size_t smem = lw_solve_smem(A.nlay, T.nband);
    if (A.aer1)
        hipLaunchKernelGGL((lw_solve_kernel<NANG, CLOUDS, JAC, OPTANG, true>), dim3(ncols), dim3(128), smem, st, T, A);
    else
        hipLaunchKernelGGL((lw_solve_kernel<NANG, CLOUDS, JAC, OPTANG, false>), dim3(ncols), dim3(128), smem, st, T, A);
}

template <int NANG>
static void launch_nang(const TablesView &T, const LwSolveArgs &A, int ncols, bool clouds, bool jac,
                        hipStream_t st)
{
    if (clouds) {
        if (jac) launch_one<NANG, true, true, false>(T, A, ncols, st);
        else launch_one<NANG, true, false, false>(T, A, ncols, st);
    } else {
        if (jac) launch_one<NANG, false, true, false>(T, A, ncols, st);
        else launch_one<NANG, false, false, false>(T, A, ncols, st);
    }
}

void launch_lw_solve(const TablesView &T, const LwSolveArgs &A, int ncols, int nang, bool clouds,
                     bool jac, bool optang, hipStream_t st)
{
    if (ncols <= 0) return;
    if (optang) {
        launch_one<1, false, false, true>(T, A, ncols, st);
        return;
    }
    switch (nang) {
    case 1: launch_nang<1>(T, A, ncols, clouds, jac, st); break;
    case 2: launch_nang<2>(T, A, ncols, clouds, jac, st); break;
    case 3: launch_nang<3>(T, A, ncols, clouds, jac, st); break;
    default: launch_nang<4>(T, A, ncols, clouds, jac, st); break;
    }
}

// compute_heating_rate on its own (GFS_rrtmgp_post.F90:167-171), FP64.
__global__ void heating_rate_kernel(int ncol, int nlay, const double *up, const double *dn,
                                    const double *p_lev, double *htr)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ncol * nlay) return;
    const size_t c = i % ncol, l = i / ncol;
    const size_t q0 = c + (size_t)ncol * l, q1 = q0 + ncol;
    htr[q0] = (up[q1] - up[q0] - dn[q1] + dn[q0]) * GRAV / (CP_DRY * (p_lev[q1] - p_lev[q0]));
}

void launch_heating_rate(int ncol, int nlay, const double *up, const double *dn, const double *p_lev,
                         double *htr, hipStream_t st)
{
    const size_t n = (size_t)ncol * nlay;
    if (!n) return;
    hipLaunchKernelGGL(heating_rate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ncol, nlay,
                       up, dn, p_lev, htr);
}

// The SW heating rate of GFS_rrtmgp_post_run (GFS_rrtmgp_post.F90:260-266):
// zero, then the daylit rows.  Each row's value is the same expression as
// heating_rate_kernel's, so it equals a call on the gathered daylit rows.
__global__ void mark_days_kernel(int nday, const int *idx, unsigned char *day)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nday) day[idx[i] - 1] = 1;
}

__global__ void heating_rate_day_kernel(int ncol, int nlay, const unsigned char *day, const double *up,
                                        const double *dn, const double *p_lev, double *htr)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ncol * nlay) return;
    const size_t c = i % ncol, l = i / ncol;
    const size_t q0 = c + (size_t)ncol * l, q1 = q0 + ncol;
    htr[q0] = day[c] ? (up[q1] - up[q0] - dn[q1] + dn[q0]) * GRAV / (CP_DRY * (p_lev[q1] - p_lev[q0])) : 0.0;
}

// day: nCol flags the caller has zeroed; idx: nday 1-based columns in 1..ncol.
void launch_mark_days(int nday, const int *idx, unsigned char *day, hipStream_t st)
{
    if (nday <= 0) return;
    hipLaunchKernelGGL(mark_days_kernel, dim3((unsigned)((nday + 255) / 256)), dim3(256), 0, st, nday, idx, day);
}

void launch_heating_rate_day(int ncol, int nlay, const unsigned char *day, const double *up, const double *dn,
                             const double *p_lev, double *htr, hipStream_t st)
{
    const size_t n = (size_t)ncol * nlay;
    if (!n) return;
    hipLaunchKernelGGL(heating_rate_day_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ncol, nlay,
                       day, up, dn, p_lev, htr);
}

// Aerosol depths of a chunk in the solvers' order.  The CCPP array is
// (ncol, nlay, nband) with the column fastest; the solvers read, per layer,
// the column's nband values together.  Block = (64 columns, one layer): the
// FP64 reads run along the columns, the FP32 values go out through LDS as
// each column's contiguous run of nband values.
__global__ __launch_bounds__(256) void lw_aer_pack_kernel(LwSolveArgs A, int ncols, int nb, float *out)
{
    extern __shared__ float tile1[];   // [64 columns][nb]
    const int d0 = blockIdx.x * 64, l = blockIdx.y, nlay = A.nlay, ld = A.ncol_ld;
    for (int i = threadIdx.x; i < 64 * nb; i += blockDim.x) {
        const int b = i >> 6, dd = i & 63, d = d0 + dd;
        if (d < ncols) tile1[dd * nb + b] = (float)A.aer_tau[(size_t)(A.col0 + d) + (size_t)ld * (l + (size_t)nlay * b)];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * nb; i += blockDim.x) {
        const int dd = i / nb, b = i - dd * nb, d = d0 + dd;
        if (d < ncols) out[((size_t)d * nlay + l) * nb + b] = tile1[i];
    }
}

void launch_lw_aer_pack(const LwSolveArgs &A, int ncols, int nband, float *out, hipStream_t st)
{
    if (ncols <= 0 || A.nlay <= 0) return;
    hipLaunchKernelGGL(lw_aer_pack_kernel, dim3((ncols + 63) / 64, A.nlay), dim3(256),
                       sizeof(float) * 64 * (size_t)nband, st, A, ncols, nband, out);
}

}  // namespace ccpprad
