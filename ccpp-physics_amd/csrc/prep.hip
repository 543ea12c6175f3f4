// The callers either side of the radiation schemes, on gfx950 (FP64), so a
// host that keeps its state on the GPU need not round-trip the fields
// through the CPU.  First the input preparation of GFS_rrtmgp_pre_run:
//   q_lay = max(qgrs(:,:,1), 1e-6)                   GFS_rrtmgp_pre.F90:236-237
//   p_lev = prsi, p_lay = prsl, t_lay = tgrs         :239-246
//   t_lay, p_lay bounded into the k-distribution's   :248-263
//     range (min + epsilon / max - epsilon)
//   t_lev by cmp_tlev (radiation_tools.F90:16-84),   :266-274
//     bounded the same way
//   vmr_h2o = q / (1 - q) amd / amw (0 where q = 1)  :370
//   vmr_o3 = o3 amd / amo3 with o3 = max(con_epsqs,  :349-352, 371
//     qgrs(:,:,i_o3)) (0 where o3 <= 0)
// One thread per column (the level temperatures combine neighbouring
// layers); threads of a wavefront are consecutive columns, so every read and
// write of the column-fastest fields is coalesced.
#include "kernels.h"
#include "../../include/ccpprad.h"

#include <cfloat>

namespace ccpprad {

constexpr double AMD = 28.9644, AMW = 18.0154, AMO3 = 47.9982;
constexpr double AMDW = AMD / AMW, AMDO3 = AMD / AMO3;
constexpr double CON_EPSQS = 1.0e-10;

__global__ __launch_bounds__(256) void prepare_kernel(PrepArgs A)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncol) return;
    const int nc = A.ncol, nl = A.nlev;
    auto at = [&](int k) { return (size_t)c + (size_t)nc * k; };
    // vertical order from the first column, as the reference
    const bool top = A.prsi[0] < A.prsi[(size_t)nc * (nl - 1)];
    if (c == 0 && A.top_at_1) *A.top_at_1 = top ? 1 : 0;
    auto bound_t = [&](double t) {
        if (t <= A.temp_min) t = A.temp_min + DBL_EPSILON;
        if (t >= A.temp_max) t = A.temp_max - DBL_EPSILON;
        return t;
    };
    // one pass over the layers: the bounded layer fields, the mixing ratios,
    // p_lev, and cmp_tlev's level between this layer and the previous one
    // (linear in log p between the layer centres), carrying the previous
    // layer's log pressure and temperature
    double la_prev = 0.0, t_prev = 0.0;
    for (int k = 0; k < nl; ++k) {
        const size_t q = at(k);
        double t = A.tgrs[q], p = A.prsl[q];
        const double pv = A.prsi[q];
        if (t <= A.temp_min) t = A.temp_min + DBL_EPSILON;
        if (p <= A.press_min) p = A.press_min + DBL_EPSILON;
        if (t >= A.temp_max) t = A.temp_max - DBL_EPSILON;
        if (p >= A.press_max) p = A.press_max - DBL_EPSILON;
        A.t_lay[q] = t;
        A.p_lay[q] = p;
        A.p_lev[q] = pv;
        double qv = A.qh2o[q];
        if (qv < 1.0e-6) qv = 1.0e-6;
        if (A.q_lay) A.q_lay[q] = qv;
        A.vmr_h2o[q] = qv != 1.0 ? (qv / (1.0 - qv)) * AMDW : 0.0;
        if (A.qo3) {
            const double o = fmax(CON_EPSQS, A.qo3[q]);
            A.vmr_o3[q] = o > 0.0 ? o * AMDO3 : 0.0;
        }
        const double la = log(p);
        if (k > 0) {
            if (top) {
                // the layer above is the previous one
                A.t_lev[q] = bound_t(t + (t_prev - t) * (log(pv) - la) / (la_prev - la));
            } else {
                const double lb = log(k == nl - 1 ? fmax(A.press_min, pv) : pv);
                A.t_lev[q] = bound_t(t_prev + (t - t_prev) * (lb - la_prev) / (la - la_prev));
            }
        }
        la_prev = la;
        t_prev = t;
    }
    A.p_lev[at(nl)] = A.prsi[at(nl)];
    // the levels beyond the first and last layers: the surface temperature
    // or the outermost layer's
    A.t_lev[at(0)] = bound_t(top ? A.t_lay[at(0)] : A.tsfc[c]);
    A.t_lev[at(nl)] = bound_t(top ? A.tsfc[c] : t_prev);
}

void launch_prepare(const PrepArgs &A, hipStream_t st)
{
    if (A.ncol <= 0) return;
    hipLaunchKernelGGL(prepare_kernel, dim3((A.ncol + 255) / 256), dim3(256), 0, st, A);
}

// GFS_rrtmgp_post_run's flux bookkeeping (GFS_rrtmgp_post.F90:173-389, less
// the heating rates): thread per column.  The statements keep the
// reference's order, so fluxr accumulates bit for bit as the host loops do.
__global__ __launch_bounds__(256) void post_kernel(ccpprad_post_args A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.nCol) return;
    const int nc = A.nCol, nl = A.nLev;
    const size_t it = (size_t)i + (size_t)nc * (A.iTOA - 1), is = (size_t)i + (size_t)nc * (A.iSFC - 1);
    auto fx = [&](int j) -> double & { return A.fluxr[(size_t)i + (size_t)nc * (j - 1)]; };   // fluxr(i, j)
    auto ilev = [&](const int *m, int j) { return m[(size_t)i + (size_t)nc * (j - 1)]; };     // (nCol, 3)
    if (A.doLWrad) {
        double *tf = A.topflw + 2 * (size_t)i, *sf = A.sfcflw + 4 * (size_t)i;
        tf[0] = A.fluxlwUP_allsky[it];
        tf[1] = A.fluxlwUP_clrsky[it];
        sf[0] = A.fluxlwUP_allsky[is];
        sf[1] = A.fluxlwUP_clrsky[is];
        sf[2] = A.fluxlwDOWN_allsky[is];
        sf[3] = A.fluxlwDOWN_clrsky[is];
        A.tsflw[i] = A.tsfa[i];
        A.sfcdlw[i] = sf[2];
        A.sfculw[i] = sf[0];
        for (int k = 0; k < nl; ++k) A.htrlwu[(size_t)i + (size_t)nc * k] = A.htrlw[(size_t)i + (size_t)nc * k];
        if (A.save_diag) {
            fx(1) = fx(1) + A.fhlwr * A.fluxlwUP_allsky[it];
            fx(19) = fx(19) + A.fhlwr * A.fluxlwDOWN_allsky[is];
            fx(20) = fx(20) + A.fhlwr * A.fluxlwUP_allsky[is];
            fx(28) = fx(28) + A.fhlwr * A.fluxlwUP_clrsky[it];
            fx(30) = fx(30) + A.fhlwr * A.fluxlwDOWN_clrsky[is];
            fx(33) = fx(33) + A.fhlwr * A.fluxlwUP_clrsky[is];
            for (int j = 1; j <= 3; ++j) {
                const double tem0d = A.raddt * A.cldsa[(size_t)i + (size_t)nc * (j - 1)];
                // cloud top / base levels outside 1..nLev (the reference assumes
                // valid ones) sum nothing instead of reading out of bounds
                const int itop = min(ilev(A.mtopa, j), nl), ibtc = max(ilev(A.mbota, j), 1);
                double tem2 = 0.0;
                for (int k = ibtc; k <= itop; ++k) tem2 = tem2 + A.cldtaulw[(size_t)i + (size_t)nc * (k - 1)];
                fx(46 - j) = fx(46 - j) + tem0d * (1.0 - exp(-tem2));
            }
        }
    }
    if (A.doSWrad) {
        double *tf = A.topfsw + 3 * (size_t)i, *sf = A.sfcfsw + 4 * (size_t)i;
        const double *cm = A.scmpsw + 6 * (size_t)i;
        if (A.nDay > 0) {
            tf[0] = A.fluxswUP_allsky[it];
            tf[2] = A.fluxswUP_clrsky[it];
            tf[1] = A.fluxswDOWN_allsky[it];
            sf[0] = A.fluxswUP_allsky[is];
            sf[2] = A.fluxswUP_clrsky[is];
            sf[1] = A.fluxswDOWN_allsky[is];
            sf[3] = A.fluxswDOWN_clrsky[is];
            A.nirbmdi[i] = cm[2];
            A.nirdfdi[i] = cm[3];
            A.visbmdi[i] = cm[4];
            A.visdfdi[i] = cm[5];
            A.nirbmui[i] = cm[2] * A.sfc_alb_nir_dir[i];
            A.nirdfui[i] = cm[3] * A.sfc_alb_nir_dif[i];
            A.visbmui[i] = cm[4] * A.sfc_alb_uvvis_dir[i];
            A.visdfui[i] = cm[5] * A.sfc_alb_uvvis_dif[i];
        } else {
            for (int k = 0; k < 4; ++k) sf[k] = 0.0;
            for (int k = 0; k < 3; ++k) tf[k] = 0.0;
            A.nirbmdi[i] = 0.0; A.nirdfdi[i] = 0.0; A.visbmdi[i] = 0.0; A.visdfdi[i] = 0.0;
            A.nirbmui[i] = 0.0; A.nirdfui[i] = 0.0; A.visbmui[i] = 0.0; A.visdfui[i] = 0.0;
        }
        A.sfcnsw[i] = sf[1] - sf[0];
        A.sfcdsw[i] = sf[1];
        if (A.save_diag) {
            for (int j = 1; j <= 6; ++j) fx(33 + j) = A.aerodp[(size_t)i + (size_t)nc * (j - 1)];
            if (A.coszen[i] > 0.0) {
                const double tem0d = A.fhswr * A.coszdg[i] / A.coszen[i];
                fx(2) = fx(2) + tf[0] * tem0d;
                fx(3) = fx(3) + sf[0] * tem0d;
                fx(4) = fx(4) + sf[1] * tem0d;
                fx(21) = fx(21) + cm[0] * tem0d;
                fx(22) = fx(22) + cm[1] * tem0d;
                fx(23) = fx(23) + tf[1] * tem0d;
                fx(24) = fx(24) + A.visbmdi[i] * tem0d;
                fx(25) = fx(25) + A.visdfdi[i] * tem0d;
                fx(26) = fx(26) + A.nirbmdi[i] * tem0d;
                fx(27) = fx(27) + A.nirdfdi[i] * tem0d;
                fx(29) = fx(29) + tf[2] * tem0d;
                fx(31) = fx(31) + sf[2] * tem0d;
                fx(32) = fx(32) + sf[3] * tem0d;
            }
            fx(17) = fx(17) + A.raddt * A.cldsa[(size_t)i + (size_t)nc * 3];
            fx(18) = fx(18) + A.raddt * A.cldsa[(size_t)i + (size_t)nc * 4];
            for (int j = 1; j <= 3; ++j) {
                const double tem0d = A.raddt * A.cldsa[(size_t)i + (size_t)nc * (j - 1)];
                // levels clamped into 1..nLev+1 for the pressures, 1..nLev for the sums
                const int mt = min(max(ilev(A.mtopa, j), 1), nl + 1), mb = min(max(ilev(A.mbota, j), 1), nl + 1);
                const double ptop = A.p_lev[(size_t)i + (size_t)nc * (mt - 1)];
                const double pbtc = A.p_lev[(size_t)i + (size_t)nc * (mb - 1)];
                const int itop = min(ilev(A.mtopa, j), nl), ibtc = max(ilev(A.mbota, j), 1);
                fx(8 - j) = fx(8 - j) + tem0d;
                fx(11 - j) = fx(11 - j) + tem0d * ptop;
                fx(14 - j) = fx(14 - j) + tem0d * pbtc;
                fx(17 - j) = fx(17 - j) + tem0d * ptop;
                double tem1 = 0.0;
                for (int k = ibtc; k <= itop; ++k) tem1 = tem1 + A.cldtausw[(size_t)i + (size_t)nc * (k - 1)];
                fx(43 - j) = fx(43 - j) + tem0d * tem1;
            }
        }
    }
}

void launch_post(const ccpprad_post_args &A, hipStream_t st)
{
    if (A.nCol <= 0) return;
    hipLaunchKernelGGL(post_kernel, dim3((A.nCol + 255) / 256), dim3(256), 0, st, A);
}

// GFS_rrtmgp_pre_run's virtual temperature and layer thicknesses
// (GFS_rrtmgp_pre.F90:289,294-341): thread per column, one pass from the
// surface up carrying the boundary height and the centre height of the layer
// below, so each layer's thickness between centres is formed as it is
// reached.  Heights accumulate from the surface in the reference's order.
__global__ __launch_bounds__(256) void geometry_kernel(GeomArgs A)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncol) return;
    const int nc = A.ncol, nl = A.nlev;
    auto at = [&](int k) { return (size_t)c + (size_t)nc * k; };
    const bool top = A.p_lev[0] < A.p_lev[(size_t)nc * (nl - 1)];   // from the first column
    double hb_lo = 0.0, hc_lo = 0.0;   // boundary height below the layer, centre height of the layer below
    double lb = log(A.p_lev[at(top ? nl : 0)]);   // log pressure at the layer's bottom level
    for (int i = 0; i < nl; ++i) {
        const int k = top ? nl - 1 - i : i;          // layer index, surface first
        const int kt = top ? k : k + 1;              // its top level
        const size_t q = at(k);
        const double tv = A.t_lay[q] * (1.0 + A.fvirt * A.q_lay[q]);
        const double lt = log(A.p_lev[at(kt)]);
        const double dz = A.rdog * fabs(lt - lb) * tv;
        const double hb = hb_lo + dz;
        const double pfac = fabs(lb - log(A.p_lay[q])) / fabs(lb - lt);
        const double hc = hb_lo + pfac * (hb - hb_lo);
        A.dzc[q] = hc - (i == 0 ? hb_lo : hc_lo);
        if (A.tv_lay) A.tv_lay[q] = tv;
        if (A.dz) A.dz[q] = dz;
        if (A.dp) A.dp[q] = fabs(A.p_lev[at(k + 1)] - A.p_lev[q]);
        hb_lo = hb;
        hc_lo = hc;
        lb = lt;                                      // the next layer's bottom level
    }
}

void launch_layer_geometry(const GeomArgs &A, hipStream_t st)
{
    if (A.ncol <= 0) return;
    hipLaunchKernelGGL(geometry_kernel, dim3((A.ncol + 255) / 256), dim3(256), 0, st, A);
}

// GFS_rrtmgp_cloud_overlap_run (GFS_rrtmgp_cloud_overlap.F90:97-137) with
// cmp_dcorr_lgth and get_alpha_exper (radiation_cloud_overlap.F90:28-145):
// thread per (column, layer), grid (columns / 256, layers).  Each thread
// forms its column's decorrelation length (a handful of flops) instead of
// reading it back, so the launch has no dependency between threads; the
// layer-1 threads write it.
__device__ __forceinline__ double dcorr_length(const ccpprad_cloud_overlap_args &A, int c)
{
    // each operation rounded as the reference writes it (no FMA), so the
    // length is the reference's to the bit; unsuffixed literals of the
    // reference take its default real's precision
#pragma clang fp contract(off)
    const bool r8 = A.literal_r8 != 0;
    double L = 0.0;
    if (A.idcor == A.idcor_hogan) {
        const double lo = r8 ? 0.6 : (double)0.6f, a = r8 ? 2.78 : (double)2.78f, b = r8 ? 4.6 : (double)4.6f;
        L = fmax(lo, a - b * fabs(A.lat[c] / A.con_pi));                                    // Hogan eq. 13
    }
    if (A.idcor == A.idcor_oreopoulos) {
        constexpr double am1 = 1.4315, am2 = 2.1219, am4 = -25.584, amr = 7.0;
        const double latd = A.lat[c] * (180.0 / A.con_pi);
        const double am3 = A.julian > 181.0 ? (-4.0 * amr * (A.julian - 272.0)) / (double)A.yearlen   // eq. 11a
                                            : (4.0 * amr * (A.julian - 91.0)) / (double)A.yearlen;    // eq. 11b
        const double x = latd - am3;
        L = am1 + am2 * exp(-(x * x) / (am4 * am4));                                        // eq. 10
    }
    if (A.idcor == A.idcor_con) L = A.dcorr_con;
    return L;
}

__device__ __forceinline__ bool exponential_overlap(const ccpprad_cloud_overlap_args &A, int iovr)
{
    return iovr == A.iovr_dcorr || iovr == A.iovr_exp || iovr == A.iovr_exprand;
}

__global__ __launch_bounds__(256) void overlap_kernel(ccpprad_cloud_overlap_args A)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (c >= A.nCol) return;
    const size_t q = (size_t)c + (size_t)A.nCol * k;
    const double L = dcorr_length(A, c);
    if (k == 0) A.de_lgth[c] = L;
    const double km = A.literal_r8 ? 0.001 : (double)0.001f;   // deltaZc * 0.001: m -> km
    auto alpha = [&](int iovr, const double *cf) {
        if (k == 0) return 0.0;
        double a = exp(-(A.deltaZc[q] * km) / L);
        // exponential-random: a clear layer after a cloudy one decorrelates
        if (iovr == A.iovr_exprand && cf[q] == 0.0 && cf[q - (size_t)A.nCol] > 0.0) a = 0.0;
        return a;
    };
    const double a = exponential_overlap(A, A.iovr) ? alpha(A.iovr, A.cld_frac) : 0.0;
    A.cloud_overlap_param[q] = a;
    A.precip_overlap_param[q] = a;
    if (A.imfdeepcnv == A.imfdeepcnv_samf || A.imfdeepcnv == A.imfdeepcnv_gf)
        A.cnv_cloud_overlap_param[q] =
            exponential_overlap(A, A.iovr_convcld) ? alpha(A.iovr_convcld, A.cld_cnv_frac) : 0.0;
}

void launch_cloud_overlap(const ccpprad_cloud_overlap_args &A, hipStream_t st)
{
    if (A.nCol <= 0 || A.nLev <= 0) return;
    hipLaunchKernelGGL(overlap_kernel, dim3((A.nCol + 255) / 256, A.nLev), dim3(256), 0, st, A);
}

}  // namespace ccpprad
