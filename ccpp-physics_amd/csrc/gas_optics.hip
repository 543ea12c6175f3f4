// k-distribution gas optics on gfx950.
//
// Restates, in a layout of its own, what ty_gas_optics_rrtmgp%gas_optics does
// inside rte-rrtmgp (called at physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:337-345
// and rrtmgp_sw_main.F90:346-353): get_col_dry, interpolation,
// gas_optical_depths_major (interpolate3D), gas_optical_depths_minor
// (interpolate2D), compute_Planck_source's Planck-fraction interpolation,
// and for SW compute_tau_rayleigh.
//
// One block = one column, 128 threads.  Layers go in chunks of LCHUNK:
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

constexpr int LCHUNK = 8;

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
    float ms[LCHUNK][MAX_MINOR];
};

template <bool LW>
__global__ __launch_bounds__(128) void gas_optics_kernel(
    TablesView T, int nlay, int ncol_ld, const int *__restrict__ col_idx, int col0,
    const double *__restrict__ p_lay, const double *__restrict__ p_lev,
    const double *__restrict__ t_lay, VmrPtrs vmr, float *__restrict__ tau,
    float *__restrict__ aux /* LW pfrac, SW tau_ray */, int *__restrict__ err,
    const unsigned char *__restrict__ row_mask, const unsigned char *__restrict__ col_mask)
{
    __shared__ GasOpticsSmem s;
    const int d = blockIdx.x;                 // column within this launch
    const int gc = col_idx ? col_idx[col0 + d] - 1 : col0 + d;  // global column
    const int tid = threadIdx.x;
    if (col_mask && !col_mask[d]) return;   // uniform per block
    const int ngpt = T.ngpt, neta = T.neta, np1 = T.npres + 1, ng1 = T.ngas + 1;
    const double tmin = T.temp_ref_min, dT = T.temp_ref_delta;

    // per-lane constants
    int gfl[2] = {0, 0};
    if (tid < ngpt) { gfl[0] = T.gpt_flav[2 * tid]; gfl[1] = T.gpt_flav[2 * tid + 1]; }

    for (int l0 = 0; l0 < nlay; l0 += LCHUNK) {
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
            for (int g = 1; g <= T.ngas; ++g) s.col[tid][g] = vmr.v[g - 1][q] * col_dry;
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
        }
        __syncthreads();
        // ---- phase A1: flavor interpolation items and minor scalings (FP64)
        const int nflav_items = nl * T.nflav * 2;
        const int nitems = nflav_items + nl * MAX_MINOR;
        for (int it = tid; it < nitems; it += blockDim.x) {
            if (it < nflav_items) {
                int j = it / (T.nflav * 2), r = it % (T.nflav * 2), f = r >> 1, t = r & 1;
                int g1 = T.flavor[2 * f], g2 = T.flavor[2 * f + 1];
                int itr = s.itropo[j], jt = s.jtemp[j];
                double ratio = T.vmr_ref[((size_t)(jt - 1 + t) * ng1 + g1) * 2 + (itr - 1)] /
                               T.vmr_ref[((size_t)(jt - 1 + t) * ng1 + g2) * 2 + (itr - 1)];
                double cm = s.col[j][g1] + ratio * s.col[j][g2];
                double eta = (cm > 2.0 * 2.2250738585072014e-308) ? s.col[j][g1] / cm : 0.5;
                double loceta = eta * (double)(neta - 1);
                int jeta = min((int)loceta + 1, neta - 1);
                double feta = fmod(loceta, 1.0);
                double ftt = t == 0 ? (1.0 - s.ftemp[j]) : s.ftemp[j];
                s.cm[j][f][t] = (float)cm;
                s.je[j][f][t] = jeta;
                s.fm[j][f][t][0] = (float)((1.0 - feta) * ftt);
                s.fm[j][f][t][1] = (float)(feta * ftt);
            } else {
                int r = it - nflav_items, j = r / MAX_MINOR, i = r % MAX_MINOR;
                int reg = s.itropo[j] - 1;
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
        // ---- phase B: g-point gathers (FP32)
        for (int g = tid; g < ngpt; g += blockDim.x) {
            for (int j = 0; j < nl; ++j) {
                const int l = l0 + j;
                const int itr = s.itropo[j];
                const int f = (g == tid) ? (itr == 1 ? gfl[0] : gfl[1]) : T.gpt_flav[2 * g + itr - 1];
                const int jt0 = s.jtemp[j] - 1;
                const int jp0 = s.jpress[j] + itr - 2;  // 0-based lower pressure row
                const float fp1 = s.fpress[j], fp0 = 1.0f - fp1;
                float tg = 0.0f, pf = 0.0f;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const int je0 = s.je[j][f][t] - 1;
                    const float fe0 = s.fm[j][f][t][0], fe1 = s.fm[j][f][t][1];
                    const size_t base = (((size_t)(jt0 + t) * np1 + jp0) * neta + je0) * ngpt + g;
                    const size_t dp_ = (size_t)neta * ngpt, de = (size_t)ngpt;
                    float k00 = T.kmajor[base], k10 = T.kmajor[base + dp_];
                    float k01 = T.kmajor[base + de], k11 = T.kmajor[base + dp_ + de];
                    float st = fp0 * fe0 * k00 + fp1 * fe0 * k10 + fp0 * fe1 * k01 + fp1 * fe1 * k11;
                    tg += s.cm[j][f][t] * st;
                    if (LW) {
                        float q00 = T.pfrac_tab[base], q10 = T.pfrac_tab[base + dp_];
                        float q01 = T.pfrac_tab[base + de], q11 = T.pfrac_tab[base + dp_ + de];
                        pf += fp0 * fe0 * q00 + fp1 * fe0 * q10 + fp0 * fe1 * q01 + fp1 * fe1 * q11;
                    }
                }
                // minor absorbers of this regime covering g
                const int reg = itr - 1;
                const int *gp = T.minor_gptr[reg];
                const int i0 = gp[g], i1 = gp[g + 1];
                const int nc = T.ncontrib[reg];
                const float *km = T.kminor[reg];
                for (int ii = i0; ii < i1; ++ii) {
                    const int iv = T.minor_items[reg][2 * ii], ci = T.minor_items[reg][2 * ii + 1];
                    const int fv = T.minor_par[reg][5 * iv + 4];
                    float kv = 0.0f;
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int je0 = s.je[j][fv][t] - 1;
                        const size_t b0 = ((size_t)(jt0 + t) * neta + je0) * nc + ci;
                        kv += s.fm[j][fv][t][0] * km[b0] + s.fm[j][fv][t][1] * km[b0 + nc];
                    }
                    tg += s.ms[j][iv] * kv;
                }
                const size_t o = ((size_t)d * nlay + l) * ngpt + g;
                if (row_mask && !row_mask[(size_t)d * nlay + l]) continue;
                tau[o] = tg;
                if (LW) {
                    aux[o] = pf;
                } else {
                    const float *kr = T.krayl[reg];
                    float kv = 0.0f;
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int je0 = s.je[j][f][t] - 1;
                        const size_t b0 = ((size_t)(jt0 + t) * neta + je0) * ngpt + g;
                        kv += s.fm[j][f][t][0] * kr[b0] + s.fm[j][f][t][1] * kr[b0 + ngpt];
                    }
                    aux[o] = kv * s.cray[j];
                }
            }
        }
        __syncthreads();
    }
}

void launch_gas_optics(const TablesView &T, int ncols, int nlay, int ncol_ld, const int *col_idx,
                       int col0, const double *p_lay, const double *p_lev, const double *t_lay,
                       const VmrPtrs &vmr, float *tau, float *aux, int *err, hipStream_t st,
                       const unsigned char *row_mask, const unsigned char *col_mask)
{
    if (ncols <= 0) return;
    if (T.is_lw)
        hipLaunchKernelGGL(gas_optics_kernel<true>, dim3(ncols), dim3(128), 0, st, T, nlay, ncol_ld,
                           col_idx, col0, p_lay, p_lev, t_lay, vmr, tau, aux, err, row_mask, col_mask);
    else
        hipLaunchKernelGGL(gas_optics_kernel<false>, dim3(ncols), dim3(128), 0, st, T, nlay, ncol_ld,
                           col_idx, col0, p_lay, p_lev, t_lay, vmr, tau, aux, err, row_mask, col_mask);
}

}  // namespace ccpprad
