// Radiation input preparation of the caller, GFS_rrtmgp_pre_run, on gfx950
// (FP64), so a host that keeps its state on the GPU need not round-trip the
// fields through the CPU before the radiation schemes:
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
    for (int k = 0; k < nl; ++k) {
        const size_t q = at(k);
        double t = A.tgrs[q], p = A.prsl[q];
        if (t <= A.temp_min) t = A.temp_min + DBL_EPSILON;
        if (p <= A.press_min) p = A.press_min + DBL_EPSILON;
        if (t >= A.temp_max) t = A.temp_max - DBL_EPSILON;
        if (p >= A.press_max) p = A.press_max - DBL_EPSILON;
        A.t_lay[q] = t;
        A.p_lay[q] = p;
        double qv = A.qh2o[q];
        if (qv < 1.0e-6) qv = 1.0e-6;
        if (A.q_lay) A.q_lay[q] = qv;
        A.vmr_h2o[q] = qv != 1.0 ? (qv / (1.0 - qv)) * AMDW : 0.0;
        if (A.qo3) {
            const double o = fmax(CON_EPSQS, A.qo3[q]);
            A.vmr_o3[q] = o > 0.0 ? o * AMDO3 : 0.0;
        }
    }
    for (int k = 0; k <= nl; ++k) A.p_lev[at(k)] = A.prsi[at(k)];
    // cmp_tlev on the bounded layer fields: linear in log p between the
    // layer centres; the level beyond the last layer takes the surface or
    // top layer temperature
    if (top) {
        A.t_lev[at(0)] = bound_t(A.t_lay[at(0)]);
        double la_up = log(A.p_lay[at(0)]), t_up = A.t_lay[at(0)];
        for (int k = 1; k < nl; ++k) {
            const double la = log(A.p_lay[at(k)]), t = A.t_lay[at(k)];
            const double lb = log(A.prsi[at(k)]);
            A.t_lev[at(k)] = bound_t(t + (t_up - t) * (lb - la) / (la_up - la));
            la_up = la; t_up = t;
        }
        A.t_lev[at(nl)] = bound_t(A.tsfc[c]);
    } else {
        A.t_lev[at(0)] = bound_t(A.tsfc[c]);
        double la_lo = log(A.p_lay[at(0)]), t_lo = A.t_lay[at(0)];
        for (int k = 1; k < nl; ++k) {
            const double la = log(A.p_lay[at(k)]), t = A.t_lay[at(k)];
            const double pv = A.prsi[at(k)];
            const double lb = log(k == nl - 1 ? fmax(A.press_min, pv) : pv);
            A.t_lev[at(k)] = bound_t(t_lo + (t - t_lo) * (lb - la_lo) / (la - la_lo));
            la_lo = la; t_lo = t;
        }
        A.t_lev[at(nl)] = bound_t(A.t_lay[at(nl - 1)]);
    }
}

void launch_prepare(const PrepArgs &A, hipStream_t st)
{
    if (A.ncol <= 0) return;
    hipLaunchKernelGGL(prepare_kernel, dim3((A.ncol + 255) / 256), dim3(256), 0, st, A);
}

}  // namespace ccpprad
