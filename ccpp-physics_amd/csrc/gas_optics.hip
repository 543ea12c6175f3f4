// k-distribution gas optics on gfx950.
//
// Restates, in a layout of its own, what ty_gas_optics_rrtmgp%gas_optics does
// inside rte-rrtmgp (called at physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:337-345
// and rrtmgp_sw_main.F90:346-353): get_col_dry, interpolation,
// gas_optical_depths_major (interpolate3D), gas_optical_depths_minor
// (interpolate2D), compute_Planck_source's Planck-fraction interpolation,
// and for SW compute_tau_rayleigh.
//
// One block = LCHUNK layers of one column, 128 threads; the layer chunk is the
// slowest grid index so that co-resident blocks share k-table rows in L2:
//   phase A0 (FP64): per layer  col_dry, col_gas, jtemp/ftemp, jpress/fpress, tropo
//   phase A1 (FP64): per (layer, flavor, T-corner) col_mix, jeta, fminor;
//                    per (layer, minor interval) the absorber scaling
//   phase B  (FP32): lane = g-point; 8 major-table corners (+8 Planck-fraction
//                    corners LW, +4 Rayleigh corners SW) and the minor
//                    intervals covering the g-point.  The discrete decisions
//                    (indices, tropopause) are therefore taken in FP64 exactly
//                    as in the FP64 reference; only the interpolation sums are FP32.
// Output tau / pfrac / tau_ray : [col][lay][gpt] FP32.
#include "kernels.h"

namespace ccpprad {

constexpr int LCHUNK = 16;

struct GasOpticsSmem {
    double col[LCHUNK][MAX_GAS + 1];
    double ftemp[LCHUNK];
    double play[LCHUNK], tlay[LCHUNK];
    int jtemp[LCHUNK], jpress[LCHUNK], itropo[LCHUNK];
    float fpress[LCHUNK];
    float cray[LCHUNK];
    float cm[LCHUNK][MAX_FLAV][2];
    int je[LCHUNK][MAX_FLAV][2];
    float fm[LCHUNK][MAX_FLAV][2][2];
    float ms[LCHUNK][MAX_MINOR + 1];   // [MAX_MINOR] = 0: scale of unused item slots
};

// A minor absorber item of one g-point: contributor column (16 bits),
// interval (8 bits), flavor (8 bits).
__device__ __forceinline__ int mi_ci(unsigned v) { return (int)(v & 0xFFFFu); }
__device__ __forceinline__ int mi_iv(unsigned v) { return (int)((v >> 16) & 0xFFu); }
__device__ __forceinline__ int mi_fv(unsigned v) { return (int)(v >> 24); }

// Gathers through a uniform table base and a 32-bit byte offset.
__device__ __forceinline__ float ldt(const float *base, unsigned idx)
{
    return *(const float *)((const char *)base + (idx << 2));
}

template <bool LW, int NM>
__global__ __launch_bounds__(128) void gas_optics_kernel(
    TablesView T, int nlay, int ncol_ld, const int *__restrict__ col_idx, int col0,
    const double *__restrict__ p_lay, const double *__restrict__ p_lev,
    const double *__restrict__ t_lay, VmrPtrs vmr, float *__restrict__ tau,
    float *__restrict__ aux /* LW pfrac, SW tau_ray */, int *__restrict__ err,
    const unsigned char *__restrict__ row_mask, const unsigned char *__restrict__ col_mask)
{
    __shared__ GasOpticsSmem s;
    // Block = (layer chunk, column), chunk slowest: blocks resident at the same
    // time work on the same pressure range and share that slice of the
    // k-tables in L2.
    const int ncols = gridDim.x / ((nlay + LCHUNK - 1) / LCHUNK);
    const int chunk = blockIdx.x / ncols;
    const int d = blockIdx.x - chunk * ncols;  // column within this launch
    const int gc = col_idx ? col_idx[col0 + d] - 1 : col0 + d;  // global column
    const int tid = threadIdx.x;
    if (col_mask && !col_mask[d]) return;   // uniform per block
    if (row_mask) {
        // fallback launch: only chunks holding a flagged layer do any work
        bool any = false;
        for (int l = chunk * LCHUNK; l < min(nlay, (chunk + 1) * LCHUNK); ++l) any |= row_mask[(size_t)d * nlay + l] != 0;
        if (!any) return;                    // uniform per block
    }
    const int ngpt = T.ngpt, neta = T.neta, np1 = T.npres + 1, nflav = T.nflav;
    const double tmin = T.temp_ref_min, dT = T.temp_ref_delta;

    // ---- per-lane constants: flavors per regime and the minor-absorber
    // items covering this g-point (the lane's g-point is tid)
    const bool act = tid < ngpt;
    const int g = act ? tid : 0;
    const int gfl0 = T.gpt_flav[2 * g], gfl1 = T.gpt_flav[2 * g + 1];
    // Per-lane minor items: NM slots per regime (NM = the table's most items
    // covering one g-point); unused slots point at interval MAX_MINOR, whose
    // scale is 0, and contributor 0, so every load is unconditional.
    unsigned mit[2][NM > 0 ? NM : 1];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int i0 = T.minor_gptr[r][g], n = T.minor_gptr[r][g + 1] - i0;
#pragma unroll
        for (int ii = 0; ii < NM; ++ii) {
            unsigned v = (unsigned)MAX_MINOR << 16;
            if (ii < n) {
                const int iv = T.minor_items[r][2 * (i0 + ii)], ci = T.minor_items[r][2 * (i0 + ii) + 1];
                v = (unsigned)ci | ((unsigned)iv << 16) | ((unsigned)T.minor_par[r][5 * iv + 4] << 24);
            }
            mit[r][ii] = v;
        }
    }
    const float *km_r[2] = {T.kminor[0], T.kminor[1]};
    const unsigned nc_r[2] = {(unsigned)T.ncontrib[0], (unsigned)T.ncontrib[1]};
    const unsigned dp_ = (unsigned)(neta * ngpt), de = (unsigned)ngpt;

    {
        const int l0 = chunk * LCHUNK;
        const int nl = min(LCHUNK, nlay - l0);
        // ---- phase A0: per layer scalars (FP64)
        if (tid < nl) {
            const int l = l0 + tid;
            const size_t q = (size_t)gc + (size_t)ncol_ld * l;
            double pl = p_lay[q], tl = t_lay[q];
            if (!(pl >= T.press_min && pl <= T.press_max && tl >= T.temp_min && tl <= T.temp_max))
                atomicOr(err, 1);
            double dp = fabs(p_lev[q] - p_lev[q + ncol_ld]);
            double vh = vmr.v[0][q];  // gas 1 is h2o
            double fact = 1.0 / (1.0 + vh);
            double m_air = (M_DRY + M_H2O * vh) * fact;
            double col_dry = 10.0 * dp * AVOGAD / (1000.0 * m_air * 100.0 * GRAV) * fact;
            s.col[tid][0] = col_dry;
            for (int gs = 1; gs <= T.ngas; ++gs) s.col[tid][gs] = vmr.v[gs - 1][q] * col_dry;
            int jt = (int)((tl - (tmin - dT)) / dT);
            jt = min(T.ntemp - 1, max(1, jt));
            s.jtemp[tid] = jt;
            s.ftemp[tid] = (tl - T.temp_ref[jt - 1]) / dT;
            double lpl = log(pl);
            double locp = 1.0 + (T.press_ref_log1 - lpl) / T.press_ref_log_delta;
            int jp = (int)locp;
            jp = min(T.npres - 1, max(1, jp));
            s.jpress[tid] = jp;
            s.fpress[tid] = (float)(locp - (double)jp);
            s.itropo[tid] = (lpl > T.press_ref_trop_log) ? 1 : 2;
            s.play[tid] = pl;
            s.tlay[tid] = tl;
            s.cray[tid] = (float)(s.col[tid][1] + col_dry);
            s.ms[tid][MAX_MINOR] = 0.0f;
        }
        __syncthreads();
        // ---- phase A1: flavor interpolation items and minor scalings (FP64)
        const int nflav_items = nl * nflav * 2;
        const int nmin_max = max(T.nminor[0], T.nminor[1]);
        const int nitems = nflav_items + nl * nmin_max;
        for (int it = tid; it < nitems; it += blockDim.x) {
            if (it < nflav_items) {
                const int j = it / (nflav * 2), r = it - j * (nflav * 2), f = r >> 1, t = r & 1;
                const int g1 = T.flavor[2 * f], g2 = T.flavor[2 * f + 1];
                const int itr = s.itropo[j], jt = s.jtemp[j];
                const double ratio = T.flav_ratio[((size_t)(jt - 1 + t) * nflav + f) * 2 + (itr - 1)];
                const double cm = s.col[j][g1] + ratio * s.col[j][g2];
                const double eta = (cm > 2.0 * 2.2250738585072014e-308) ? s.col[j][g1] / cm : 0.5;
                const double loceta = eta * (double)(neta - 1);
                const int jeta = min((int)loceta + 1, neta - 1);
                const double feta = loceta - floor(loceta);   // fmod(loceta, 1) for loceta >= 0
                const double ftt = t == 0 ? (1.0 - s.ftemp[j]) : s.ftemp[j];
                s.cm[j][f][t] = (float)cm;
                s.je[j][f][t] = jeta;
                s.fm[j][f][t][0] = (float)((1.0 - feta) * ftt);
                s.fm[j][f][t][1] = (float)(feta * ftt);
            } else {
                const int r = it - nflav_items, j = r / nmin_max, i = r - j * nmin_max;
                const int reg = s.itropo[j] - 1;
                if (i < T.nminor[reg]) {
                    const int *mp = T.minor_par[reg] + 5 * i;
                    double sc = s.col[j][mp[0]];
                    if (mp[1]) {
                        sc *= 0.01 * s.play[j] / s.tlay[j];
                        if (mp[2] > 0) {
                            double vf = 1.0 / s.col[j][0];
                            double df = 1.0 / (1.0 + s.col[j][1] * vf);
                            if (mp[3]) sc *= (1.0 - s.col[j][mp[2]] * vf * df);
                            else sc *= s.col[j][mp[2]] * vf * df;
                        }
                    }
                    s.ms[j][i] = (float)sc;
                }
            }
        }
        __syncthreads();
        // ---- phase B: g-point gathers (FP32), two layers at a time: every
        // table load of both layers (major, Planck fraction or Rayleigh, minor)
        // is issued before any arithmetic; layer-uniform indices are scalar
        if (act) {
            for (int j0 = 0; j0 < nl; j0 += 2) {
                const bool two = j0 + 1 < nl;
                float k[2][2][4], q[2][2][4], kr[2][2][2], km[2][NM > 0 ? NM : 1][2][2];
                float w[2][2][2], cmv[2][2], fp1v[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u == 1 && !two) break;
                    const int j = j0 + u;
                    const int itr = __builtin_amdgcn_readfirstlane(s.itropo[j]), reg = itr - 1;
                    const int f = reg == 0 ? gfl0 : gfl1;
                    const unsigned jt0 = (unsigned)__builtin_amdgcn_readfirstlane(s.jtemp[j] - 1);
                    const unsigned jp0 = (unsigned)__builtin_amdgcn_readfirstlane(s.jpress[j] + itr - 2);
                    fp1v[u] = s.fpress[j];
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int je0 = s.je[j][f][t] - 1;
                        w[u][t][0] = s.fm[j][f][t][0];
                        w[u][t][1] = s.fm[j][f][t][1];
                        cmv[u][t] = s.cm[j][f][t];
                        const unsigned base = (((jt0 + t) * np1 + jp0) * neta + je0) * ngpt + g;
                        k[u][t][0] = ldt(T.kmajor, base); k[u][t][1] = ldt(T.kmajor, base + dp_);
                        k[u][t][2] = ldt(T.kmajor, base + de); k[u][t][3] = ldt(T.kmajor, base + dp_ + de);
                        if (LW) {
                            q[u][t][0] = ldt(T.pfrac_tab, base); q[u][t][1] = ldt(T.pfrac_tab, base + dp_);
                            q[u][t][2] = ldt(T.pfrac_tab, base + de); q[u][t][3] = ldt(T.pfrac_tab, base + dp_ + de);
                        } else {
                            const unsigned b0 = ((jt0 + t) * neta + je0) * ngpt + g;
                            kr[u][t][0] = ldt(T.krayl[reg], b0); kr[u][t][1] = ldt(T.krayl[reg], b0 + ngpt);
                        }
                    }
                    const unsigned nc = nc_r[reg];
                    const float *kmt = km_r[reg];
#pragma unroll
                    for (int ii = 0; ii < NM; ++ii) {
                        const unsigned v = reg == 0 ? mit[0][ii] : mit[1][ii];
                        const int fv = mi_fv(v);
                        const unsigned ci = (unsigned)mi_ci(v);
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            const unsigned b0 = ((jt0 + t) * neta + (unsigned)(s.je[j][fv][t] - 1)) * nc + ci;
                            km[u][ii][t][0] = ldt(kmt, b0);
                            km[u][ii][t][1] = ldt(kmt, b0 + nc);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u == 1 && !two) break;
                    const int j = j0 + u, l = l0 + j;
                    const int reg = __builtin_amdgcn_readfirstlane(s.itropo[j]) - 1;
                    const float fp1 = fp1v[u], fp0 = 1.0f - fp1;
                    float tg = 0.0f, pf = 0.0f, kv = 0.0f;
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const float fe0 = w[u][t][0], fe1 = w[u][t][1];
                        const float st = fp0 * fe0 * k[u][t][0] + fp1 * fe0 * k[u][t][1] + fp0 * fe1 * k[u][t][2] +
                                         fp1 * fe1 * k[u][t][3];
                        tg += cmv[u][t] * st;
                        if (LW) pf += fp0 * fe0 * q[u][t][0] + fp1 * fe0 * q[u][t][1] + fp0 * fe1 * q[u][t][2] +
                                      fp1 * fe1 * q[u][t][3];
                        else kv += fe0 * kr[u][t][0] + fe1 * kr[u][t][1];
                    }
                    // minor absorbers of this layer's regime covering g
#pragma unroll
                    for (int ii = 0; ii < NM; ++ii) {
                        const unsigned v = reg == 0 ? mit[0][ii] : mit[1][ii];
                        const int fv = mi_fv(v);
                        float kvm = 0.0f;
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            kvm += s.fm[j][fv][t][0] * km[u][ii][t][0] + s.fm[j][fv][t][1] * km[u][ii][t][1];
                        tg += s.ms[j][mi_iv(v)] * kvm;
                    }
                    if (row_mask && !row_mask[(size_t)d * nlay + l]) continue;
                    const size_t o = ((size_t)d * nlay + l) * ngpt + g;
                    tau[o] = tg;
                    aux[o] = LW ? pf : kv * s.cray[j];
                }
            }
        }
    }
}

template <bool LW>
static void launch_nm(const TablesView &T, int nblocks, int nlay, int ncol_ld, const int *col_idx, int col0,
                      const double *p_lay, const double *p_lev, const double *t_lay, const VmrPtrs &vmr,
                      float *tau, float *aux, int *err, hipStream_t st, const unsigned char *row_mask,
                      const unsigned char *col_mask)
{
#define GO_LAUNCH(NMV)                                                                                        \
    hipLaunchKernelGGL((gas_optics_kernel<LW, NMV>), dim3(nblocks), dim3(128), 0, st, T, nlay, ncol_ld, col_idx, \
                       col0, p_lay, p_lev, t_lay, vmr, tau, aux, err, row_mask, col_mask)
    const int nm = T.maxminor_g;
    if (nm <= 0) GO_LAUNCH(0);
    else if (nm == 1) GO_LAUNCH(1);
    else if (nm == 2) GO_LAUNCH(2);
    else if (nm <= 4) GO_LAUNCH(4);
    else GO_LAUNCH(MAX_MINOR_G);
#undef GO_LAUNCH
}

void launch_gas_optics(const TablesView &T, int ncols, int nlay, int ncol_ld, const int *col_idx,
                       int col0, const double *p_lay, const double *p_lev, const double *t_lay,
                       const VmrPtrs &vmr, float *tau, float *aux, int *err, hipStream_t st,
                       const unsigned char *row_mask, const unsigned char *col_mask)
{
    if (ncols <= 0 || nlay <= 0) return;
    const int nblocks = ncols * ((nlay + LCHUNK - 1) / LCHUNK);
    if (T.is_lw)
        launch_nm<true>(T, nblocks, nlay, ncol_ld, col_idx, col0, p_lay, p_lev, t_lay, vmr, tau, aux, err, st,
                        row_mask, col_mask);
    else
        launch_nm<false>(T, nblocks, nlay, ncol_ld, col_idx, col0, p_lay, p_lev, t_lay, vmr, tau, aux, err, st,
                         row_mask, col_mask);
}

}  // namespace ccpprad
