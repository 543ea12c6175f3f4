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
//   2 bottom-up: two-stream layer optics, direct-beam sources, adding
//                albedo A and source Q per level (FP32 scratch)
//   3 top-down : layer optics again, diffuse down/up fluxes; F_dn += F_dir
// The layer optics are recomputed from the 8 B/cell gas optics instead of
// being stored.  FP32 cancellation in thin layers is avoided with expm1 for
// 1 - exp(-2k tau) and 1 - T_noscat.
#include "kernels.h"

namespace ccpprad {


__device__ __forceinline__ void inc2s_f(float t1, float s1, float g1, float &t2, float &s2, float &g2)
{
    const float eps = 3.0f * 1.17549435e-38f;
    const float t12 = t1 + t2;
    const float ts12 = t1 * s1 + t2 * s2;
    g2 = (t1 * s1 * g1 + t2 * s2 * g2) / fmaxf(eps, ts12);
    s2 = ts12 / fmaxf(eps, t12);
    t2 = t12;
}

struct TwoStream {
    float Rd, Td, Rdir, Tdir, T0;
};

__device__ __forceinline__ TwoStream two_stream(float mu0, float mu0_inv, float tau, float w, float g)
{
    TwoStream o;
    const float g1 = (8.0f - w * (5.0f + 3.0f * g)) * 0.25f;
    const float g2 = 3.0f * (w * (1.0f - g)) * 0.25f;
    const float g3 = (2.0f - 3.0f * mu0 * g) * 0.25f;
    const float g4 = 1.0f - g3;
    const float a1 = g1 * g4 + g2 * g3;
    const float a2 = g1 * g3 + g2 * g4;
    const float k = sqrtf(fmaxf((g1 - g2) * (g1 + g2), 1.0e-12f));
    const float e = expf(-tau * k);
    const float e2 = e * e;
    const float ome2 = -expm1f(-2.0f * tau * k);  // 1 - e2 without cancellation
    const float RT = 1.0f / (k * (1.0f + e2) + g1 * ome2);
    o.Rd = RT * g2 * ome2;
    o.Td = RT * 2.0f * k * e;
    const float x0 = tau * mu0_inv;
    o.T0 = expf(-x0);
    const float omT0 = -expm1f(-x0);
    const float kmu = k * mu0, kg3 = k * g3, kg4 = k * g4;
    const float dd = 1.0f - kmu * kmu;
    const float RT2 = w * RT / (fabsf(dd) >= 1.1920929e-7f ? dd : 1.1920929e-7f);
    float Rdir = RT2 * ((1.0f - kmu) * (a2 + kg3) - (1.0f + kmu) * (a2 - kg3) * e2 -
                        2.0f * (kg3 - a2 * kmu) * e * o.T0);
    float Tdir = -RT2 * ((1.0f + kmu) * (a1 + kg4) * o.T0 - (1.0f - kmu) * (a1 - kg4) * e2 * o.T0 -
                         2.0f * (kg4 + a1 * kmu) * e);
    Rdir = fmaxf(0.0f, fminf(Rdir, omT0));
    Tdir = fmaxf(0.0f, fminf(Tdir, omT0 - Rdir));
    o.Rdir = Rdir;
    o.Tdir = Tdir;
    return o;
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

    const float mu0 = (float)A.coszen[c];
    const float mu0_inv = 1.0f / mu0;
    const int cat = A.alb_cat[b];
    float adir = 0.0f, adif = 0.0f;
    if (cat == 0) { adir = (float)A.alb_nir_dir[c]; adif = (float)A.alb_nir_dif[c]; }
    else if (cat == 1) {
        adir = (float)(0.5 * (A.alb_nir_dir[c] + A.alb_uv_dir[c]));
        adif = (float)(0.5 * (A.alb_nir_dif[c] + A.alb_uv_dif[c]));
    } else if (cat == 2) { adir = (float)A.alb_uv_dir[c]; adif = (float)A.alb_uv_dif[c]; }

    const size_t cl = (size_t)d * nlay;
    auto optics = [&](int l, float &t, float &w, float &gq) {
        const float ta = A.tau_abs[(cl + l) * ngpt + gg], tr = A.tau_ray[(cl + l) * ngpt + gg];
        t = ta + tr;
        if (t > 2.0f * 1.17549435e-38f) w = tr / t; else w = 0.0f;
        gq = 0.0f;
        if (A.aer_tau) {
            const size_t q = (size_t)c + (size_t)ld * (l + (size_t)nlay * b);
            inc2s_f((float)A.aer_tau[q], (float)A.aer_ssa[q], (float)A.aer_g[q], t, w, gq);
        }
        if (cloudy) {
            const unsigned mw = A.mask[(cl + l) * MASK_WORDS + (gg >> 5)];
            const bool m = (mw >> (gg & 31)) & 1u;
            const size_t q = (cl + l) * nb + b;
            inc2s_f(m ? A.ctau[q] : 0.0f, m ? A.cssa[q] : 0.0f, m ? A.cg[q] : 0.0f, t, w, gq);
        }
    };
    float *sF = A.sF + (size_t)d * nlev * ngpt, *sA = A.sA + (size_t)d * nlev * ngpt;
    float *sQ = A.sQ + (size_t)d * nlev * ngpt;
    const float toa = T.toa_norm[gg];

    // ---- sweep 1: direct beam, top-down
    float fdir = toa * mu0;
    sF[(size_t)lev_toa * ngpt + gg] = fdir;
    for (int k = 0; k < nlay; ++k) {
        const int l = top ? k : nlay - 1 - k;
        const int lev_below = top ? l + 1 : l;
        float t, w, gq;
        optics(l, t, w, gq);
        fdir = expf(-t * mu0_inv) * fdir;
        sF[(size_t)lev_below * ngpt + gg] = fdir;
    }
    // ---- sweep 2: adding albedo / source, bottom-up
    float Ab = adif, Qb = fdir * adir;
    sA[(size_t)lev_sfc * ngpt + gg] = Ab;
    sQ[(size_t)lev_sfc * ngpt + gg] = Qb;
    for (int k = 0; k < nlay; ++k) {
        const int l = top ? nlay - 1 - k : k;
        const int lev_above = top ? l : l + 1;
        float t, w, gq;
        optics(l, t, w, gq);
        const TwoStream o = two_stream(mu0, mu0_inv, t, w, gq);
        const float fa = sF[(size_t)lev_above * ngpt + gg];
        const float su = o.Rdir * fa, sd = o.Tdir * fa;
        const float den = 1.0f / (1.0f - o.Rd * Ab);
        const float Aa = o.Rd + o.Td * o.Td * Ab * den;
        const float Qa = su + o.Td * den * (Qb + Ab * sd);
        sA[(size_t)lev_above * ngpt + gg] = Aa;
        sQ[(size_t)lev_above * ngpt + gg] = Qa;
        Ab = Aa; Qb = Qa;
    }
    // ---- sweep 3: fluxes, top-down
    float dnd = 0.0f;                    // diffuse down at TOA
    fdir = toa * mu0;
    {
        const float up = dnd * Ab + Qb;  // Ab, Qb now hold TOA values
        const double su = wave_sum(act ? (double)up : 0.0);
        const double sdn = wave_sum(act ? (double)(dnd + fdir) : 0.0);
        if (lane == 0) { rup[lev_toa] = su; rdn[lev_toa] = sdn; }
    }
    for (int k = 0; k < nlay; ++k) {
        const int l = top ? k : nlay - 1 - k;
        const int lev_below = top ? l + 1 : l;
        float t, w, gq;
        optics(l, t, w, gq);
        const TwoStream o = two_stream(mu0, mu0_inv, t, w, gq);
        const float sd = o.Tdir * fdir;
        const float Abl = sA[(size_t)lev_below * ngpt + gg], Qbl = sQ[(size_t)lev_below * ngpt + gg];
        const float den = 1.0f / (1.0f - o.Rd * Abl);
        dnd = (o.Td * dnd + o.Rd * Qbl + sd) * den;
        const float up = dnd * Abl + Qbl;
        fdir = o.T0 * fdir;
        const float dn = dnd + fdir;
        const double su = wave_sum(act ? (double)up : 0.0);
        const double sdn = wave_sum(act ? (double)dn : 0.0);
        if (lane == 0) { rup[lev_below] = su; rdn[lev_below] = sdn; }
        if (lev_below == A.isfc) {
            // per-band surface sums (8-lane groups share a band) for scmpsw
            double vdn = act ? (double)dn : 0.0, vdir = act ? (double)fdir : 0.0;
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
