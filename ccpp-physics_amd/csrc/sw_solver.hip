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
// One block = one daylit column, blockDim 128, lane = g-point.  Three sweeps
// per lane:
//   1 top-down : direct beam F_dir by products of T_noscat (as the reference),
//                stored per level (FP32 scratch)
//   2 bottom-up: layer two-stream coefficients (FP64), direct-beam sources,
//                adding albedo A and source Q per level; R_dif, T_dif and the
//                downward direct source are kept per layer (FP32 scratch)
//   3 top-down : diffuse down/up fluxes from the stored terms; F_dn += F_dir
// Optical properties carry the co-albedo (1 - ssa) next to ssa, so the
// absorption of nearly conservative layers is not lost to FP32 rounding.
#include "kernels.h"

namespace ccpprad {


// Two-stream optics carried with the co-albedo p = 1 - w (p = tau_abs / tau).
struct Opt2s {
    float t, w, p, g;
};

// increment_2stream_by_2stream: b += a (eps = 3*tiny(FP32)).
__device__ __forceinline__ void inc2s_p(float t1, float w1, float p1, float g1, Opt2s &o)
{
    const float eps = 3.0f * 1.17549435e-38f;
    const float t12 = t1 + o.t;
    const float ts12 = t1 * w1 + o.t * o.w;
    const float ta12 = t1 * p1 + o.t * o.p;
    o.g = (t1 * w1 * g1 + o.t * o.w * o.g) / fmaxf(eps, ts12);
    o.w = ts12 / fmaxf(eps, t12);
    o.p = t12 > eps ? ta12 / t12 : 1.0f;
    o.t = t12;
}

struct TwoStream {
    float Rd, Td, Rdir, Tdir;
};

// sw_two_stream for one layer, evaluated in FP64: the direct-beam
// reflectance/transmittance brackets cancel to O(k tau) while their prefactor
// grows like 1/k, which FP32 cannot carry to the 0.05 W/m2 bar.  gamma1 -
// gamma2 = 2 (1 - w) is formed from the co-albedo.  Outputs are rounded to
// FP32 for the adding recurrences.
__device__ __forceinline__ TwoStream two_stream64(double mu0, const Opt2s &o)
{
    const double tau = o.t, w = o.w, p = o.p, g = o.g;
    const double g1 = (8.0 - w * (5.0 + 3.0 * g)) * 0.25;
    const double g2 = 3.0 * (w * (1.0 - g)) * 0.25;
    const double g3 = (2.0 - 3.0 * mu0 * g) * 0.25;
    const double g4 = 1.0 - g3;
    const double a1 = g1 * g4 + g2 * g3;
    const double a2 = g1 * g3 + g2 * g4;
    const double k = sqrt(fmax(2.0 * p * (g1 + g2), 1.0e-12));
    const double e = exp(-tau * k);
    const double e2 = e * e;
    const double ome2 = -expm1(-2.0 * tau * k);
    const double RT = 1.0 / (k * (1.0 + e2) + g1 * ome2);
    const double x0 = tau / mu0;
    const double T0 = exp(-x0), omT0 = -expm1(-x0);
    const double kmu = k * mu0, kg3 = k * g3, kg4 = k * g4;
    const double dd = 1.0 - kmu * kmu;
    const double RT2 = w * RT / (fabs(dd) >= 2.220446049250313e-16 ? dd : 2.220446049250313e-16);
    double Rdir = RT2 * ((1.0 - kmu) * (a2 + kg3) - (1.0 + kmu) * (a2 - kg3) * e2 - 2.0 * (kg3 - a2 * kmu) * e * T0);
    double Tdir = -RT2 * ((1.0 + kmu) * (a1 + kg4) * T0 - (1.0 - kmu) * (a1 - kg4) * e2 * T0 - 2.0 * (kg4 + a1 * kmu) * e);
    Rdir = fmax(0.0, fmin(Rdir, omT0));
    Tdir = fmax(0.0, fmin(Tdir, omT0 - Rdir));
    TwoStream r;
    r.Rd = (float)(RT * g2 * ome2);
    r.Td = (float)(RT * 2.0 * k * e);
    r.Rdir = (float)Rdir;
    r.Tdir = (float)Tdir;
    return r;
}

template <bool CLOUDS>
__global__ __launch_bounds__(128) void sw_solve_kernel(TablesView T, SwSolveArgs A)
{
    extern __shared__ double red[];  // [2 waves][2][nlev] + [2][3][nband] band partials
    const int d = blockIdx.x;
    const int od = A.col0 + d;
    const int c = A.col_idx ? A.col_idx[od] - 1 : od;
    const int g = threadIdx.x, wave = g >> 6, lane = g & 63;
    const int nlay = A.nlay, nlev = nlay + 1, ld = A.ncol_ld, ngpt = T.ngpt, nb = T.nband;
    const bool act = g < ngpt;
    const int gg = act ? g : 0;
    const int b = T.g2b[gg];
    const bool cloudy = CLOUDS && A.blkflag[od / A.blksz];
    const int top = A.top_at_1;
    const int lev_toa = top ? 0 : nlay, lev_sfc = top ? nlay : 0;
    double *rup = red + (size_t)wave * 2 * nlev, *rdn = rup + nlev;
    double *rb = red + 4 * (size_t)nlev;  // [wave][3][nband]

    const double mu0d = A.coszen[c];
    const float mu0 = (float)mu0d;
    const float mu0_inv = 1.0f / mu0;
    const int cat = A.alb_cat[b];
    float adir = 0.0f, adif = 0.0f;
    if (cat == 0) { adir = (float)A.alb_nir_dir[c]; adif = (float)A.alb_nir_dif[c]; }
    else if (cat == 1) {
        adir = (float)(0.5 * (A.alb_nir_dir[c] + A.alb_uv_dir[c]));
        adif = (float)(0.5 * (A.alb_nir_dif[c] + A.alb_uv_dif[c]));
    } else if (cat == 2) { adir = (float)A.alb_uv_dir[c]; adif = (float)A.alb_uv_dif[c]; }

    const size_t cl = (size_t)d * nlay;
    // combine_and_reorder_2str + aerosol and McICA cloud increments
    auto optics = [&](int l) -> Opt2s {
        const float ta = A.tau_abs[(cl + l) * ngpt + gg], tr = A.tau_ray[(cl + l) * ngpt + gg];
        Opt2s o;
        o.t = ta + tr;
        if (o.t > 2.0f * 1.17549435e-38f) { o.w = tr / o.t; o.p = ta / o.t; } else { o.w = 0.0f; o.p = 1.0f; }
        o.g = 0.0f;
        if (A.aer_tau) {
            const size_t q = (size_t)c + (size_t)ld * (l + (size_t)nlay * b);
            const float aw = (float)A.aer_ssa[q];
            inc2s_p((float)A.aer_tau[q], aw, (float)(1.0 - A.aer_ssa[q]), (float)A.aer_g[q], o);
        }
        if (cloudy) {
            const unsigned mw = A.mask[(cl + l) * MASK_WORDS + (gg >> 5)];
            const bool m = (mw >> (gg & 31)) & 1u;
            const size_t q = (cl + l) * nb + b;
            const float cp = m ? A.cssa[q] : 1.0f;
            inc2s_p(m ? A.ctau[q] : 0.0f, 1.0f - cp, cp, m ? A.cg[q] : 0.0f, o);
        }
        return o;
    };
    const size_t sl = (size_t)d * nlev * ngpt, sy = (size_t)d * nlay * ngpt;
    float *sF = A.sF + sl, *sA = A.sA + sl, *sQ = A.sQ + sl;
    float *sR = A.sR + sy, *sT = A.sT + sy, *sD = A.sD + sy;
    const float toa = T.toa_norm[gg];

    // ---- sweep 1: direct beam, top-down (products of T_noscat)
    float fdir = toa * mu0;
    sF[(size_t)lev_toa * ngpt + gg] = fdir;
    for (int k = 0; k < nlay; ++k) {
        const int l = top ? k : nlay - 1 - k;
        const int lev_below = top ? l + 1 : l;
        const Opt2s o = optics(l);
        fdir = expf(-o.t * mu0_inv) * fdir;
        sF[(size_t)lev_below * ngpt + gg] = fdir;
    }
    // ---- sweep 2: layer two-stream (FP64) and adding albedo / source, bottom-up
    float Ab = adif, Qb = fdir * adir;
    sA[(size_t)lev_sfc * ngpt + gg] = Ab;
    sQ[(size_t)lev_sfc * ngpt + gg] = Qb;
    for (int k = 0; k < nlay; ++k) {
        const int l = top ? nlay - 1 - k : k;
        const int lev_above = top ? l : l + 1;
        const TwoStream o = two_stream64(mu0d, optics(l));
        const float fa = sF[(size_t)lev_above * ngpt + gg];
        const float su = o.Rdir * fa, sd = o.Tdir * fa;
        const float den = 1.0f / (1.0f - o.Rd * Ab);
        const float Aa = o.Rd + o.Td * o.Td * Ab * den;
        const float Qa = su + o.Td * den * (Qb + Ab * sd);
        sA[(size_t)lev_above * ngpt + gg] = Aa;
        sQ[(size_t)lev_above * ngpt + gg] = Qa;
        sR[(size_t)l * ngpt + gg] = o.Rd;
        sT[(size_t)l * ngpt + gg] = o.Td;
        sD[(size_t)l * ngpt + gg] = sd;
        Ab = Aa; Qb = Qa;
    }
    // ---- sweep 3: fluxes, top-down
    float dnd = 0.0f;                    // diffuse down at TOA
    {
        const float up = dnd * Ab + Qb;  // Ab, Qb hold the TOA values
        const double su = wave_sum(act ? (double)up : 0.0);
        const double sdn = wave_sum(act ? (double)(dnd + toa * mu0) : 0.0);
        if (lane == 0) { rup[lev_toa] = su; rdn[lev_toa] = sdn; }
    }
    for (int k = 0; k < nlay; ++k) {
        const int l = top ? k : nlay - 1 - k;
        const int lev_below = top ? l + 1 : l;
        const float Rd = sR[(size_t)l * ngpt + gg], Td = sT[(size_t)l * ngpt + gg], sd = sD[(size_t)l * ngpt + gg];
        const float Abl = sA[(size_t)lev_below * ngpt + gg], Qbl = sQ[(size_t)lev_below * ngpt + gg];
        const float den = 1.0f / (1.0f - Rd * Abl);
        dnd = (Td * dnd + Rd * Qbl + sd) * den;
        const float up = dnd * Abl + Qbl;
        const float fd = sF[(size_t)lev_below * ngpt + gg];
        const float dn = dnd + fd;
        const double su = wave_sum(act ? (double)up : 0.0);
        const double sdn = wave_sum(act ? (double)dn : 0.0);
        if (lane == 0) { rup[lev_below] = su; rdn[lev_below] = sdn; }
        if (lev_below == A.isfc) {
            // per-band surface sums for scmpsw
            const double vdn = act ? (double)dn : 0.0, vdir = act ? (double)fd : 0.0;
            for (int bb = 0; bb < nb; ++bb) {
                const bool mine = act && b == bb;
                const double s1 = wave_sum(mine ? vdn : 0.0), s2 = wave_sum(mine ? vdir : 0.0);
                if (lane == 0) { rb[(size_t)wave * 3 * nb + bb] = s1; rb[(size_t)wave * 3 * nb + nb + bb] = s2; }
            }
        }
    }
    if (A.isfc == lev_toa) {
        for (int bb = 0; bb < nb; ++bb) {
            const bool mine = act && b == bb;
            const double s1 = wave_sum(mine ? (double)(toa * mu0) : 0.0), s2 = s1;
            if (lane == 0) { rb[(size_t)wave * 3 * nb + bb] = s1; rb[(size_t)wave * 3 * nb + nb + bb] = s2; }
        }
    }
    __syncthreads();
    // ---- epilogue
    double *r0 = red, *r1 = red + 2 * (size_t)nlev;
    const bool clr_too = (A.role == 1) || !cloudy;  // fluxes equal the clear-sky ones
    for (int v = g; v < nlev; v += blockDim.x) {
        const double up = r0[v] + r1[v], dn = r0[nlev + v] + r1[nlev + v];
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
    double *hout[2] = {A.role == 0 ? A.htr : nullptr, (clr_too || A.role == 1) ? A.htr_clr : nullptr};
    for (int h = 0; h < 2; ++h) {
        if (!hout[h]) continue;
        for (int l = g; l < nlay; l += blockDim.x) {
            const double u0 = r0[l] + r1[l], u1 = r0[l + 1] + r1[l + 1];
            const double d0 = r0[nlev + l] + r1[nlev + l], d1 = r0[nlev + l + 1] + r1[nlev + l + 1];
            const double p0 = A.p_lev[(size_t)c + (size_t)ld * l], p1 = A.p_lev[(size_t)c + (size_t)ld * (l + 1)];
            hout[h][(size_t)c + (size_t)ld * l] = (u1 - u0 - d1 + d0) * GRAV / (CP_DRY * (p1 - p0));
        }
    }
}

void launch_sw_solve(const TablesView &T, const SwSolveArgs &A, int ncols, bool clouds, hipStream_t st)
{
    if (ncols <= 0) return;
    const size_t smem = sizeof(double) * (4 * (size_t)(A.nlay + 1) + 6 * (size_t)T.nband);
    if (clouds)
        hipLaunchKernelGGL(sw_solve_kernel<true>, dim3(ncols), dim3(128), smem, st, T, A);
    else
        hipLaunchKernelGGL(sw_solve_kernel<false>, dim3(ncols), dim3(128), smem, st, T, A);
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

}  // namespace ccpprad
