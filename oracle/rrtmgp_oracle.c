/*
 * rrtmgp_oracle.c -- FP64 CPU restatement of the CCPP RRTMGP radiation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it.
 *
 * What it restates:
 *  - The CCPP scheme layer, from the reference sources in-tree:
 *      rrtmgp_lw_main_run   physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:104-613
 *      rrtmgp_sw_main_run   physics/Radiation/RRTMGP/rrtmgp_sw_main.F90:87-683
 *      sampled_mask / draw_samples / apply_cloud_mask
 *                           physics/Radiation/RRTMGP/rrtmgp_sampling.F90:39-205
 *      MT19937 (random_setseed_t / random_number_t)
 *                           physics/Radiation/mersenne_twister.f:285-372
 *      rain/snow optics     rrtmgp_lw_main.F90:409-431, rrtmgp_sw_main.F90:436-471
 *                           constants rrtmgp_lw_cloud_optics.F90:63-66,
 *                           rrtmgp_sw_cloud_optics.F90:59-63,369-381
 *  - The rte-rrtmgp kernels the scheme calls (the submodule
 *    physics/Radiation/RRTMGP/rte-rrtmgp is empty in the reference tree;
 *    call sites rrtmgp_lw_main.F90:337,366,508-525,584-599 and
 *    rrtmgp_sw_main.F90:346,396,548,607), from the published BSD-3
 *    rte-rrtmgp algorithm (Pincus, Mlawer & Delamere 2019): get_col_dry,
 *    interpolation, gas_optical_depths_major/minor, compute_Planck_source,
 *    compute_tau_rayleigh + combine_and_reorder_2str, compute_all_from_table,
 *    increment_*, compute_optimal_angles, lw_solver_noscat(_GaussQuad),
 *    sw_two_stream + sw_source_2str + adding, sum_byband,
 *    compute_heating_rate.
 *
 * Parity pins: see the test files; each function below names the test that
 * pins it.  Functions without a pin say "parity unpinned".
 *
 * Array layouts: CCPP fields are Fortran (ncol, nlay...) column-fastest.
 * Oracle intermediates are g-point fastest: (ngpt, nlay, ncol).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <tgmath.h>
#include "../include/ccpprad.h"

/* Working precision of the per-g-point optics and transport (the arrays
 * shaped (ngpt, nlay, ncol) and the solvers' recurrences): double, or float
 * for the FP32 CPU baseline build (-DORC_REAL=float, liboracle32.so).  The
 * interpolation indices and weights, the cloud optics by band, McICA and
 * the band / broadband reductions stay FP64 in both builds.  <tgmath.h>
 * picks exp / sqrt / fmax ... by argument type; R() makes every literal of
 * the working precision, so the double build is the same program as before
 * and the float build computes the per-g-point work in FP32. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
#define R(x) ((real)(x))
#define R_EPS ((real)(sizeof(real) == sizeof(float) ? (double)FLT_EPSILON : DBL_EPSILON))
#define R_MIN ((real)(sizeof(real) == sizeof(float) ? (double)FLT_MIN : DBL_MIN))

/* rte-rrtmgp mo_gas_optics_constants (recalled; parity unpinned) */
#define AVOGAD 6.02214076e23
#define M_DRY 0.028964
#define M_H2O 0.018016
#define GRAV 9.80665
#define CP_DRY 1004.64

#define MAXFLAV 16
#define MAXG 256
#define MAXB 32
#define IH2O 1

#define F2(a, i, j, n1) ((a)[(size_t)(i) + (size_t)(n1) * (size_t)(j)])
#define F3(a, i, j, k, n1, n2) ((a)[(size_t)(i) + (size_t)(n1) * ((size_t)(j) + (size_t)(n2) * (size_t)(k))])
#define F4(a, i, j, k, l, n1, n2, n3) \
    ((a)[(size_t)(i) + (size_t)(n1) * ((size_t)(j) + (size_t)(n2) * ((size_t)(k) + (size_t)(n3) * (size_t)(l)))])

/* ===================================================================== */
/* Gas optics                                                             */
/* ===================================================================== */
typedef struct {
    int jtemp, jpress, itropo; /* 1-based; itropo 1 = lower, 2 = upper */
    double ftemp, fpress;
    double col[CCPPRAD_MAX_GAS + 1]; /* col[0] = dry air */
    int jeta[MAXFLAV][2];
    double col_mix[MAXFLAV][2];
    double fminor[MAXFLAV][2][2]; /* [flav][itemp][ieta] */
} layer_interp;

/* get_col_dry + interpolation for one layer.  Pinned by
 * tests/test_oracle_gas_optics.py (exact-trilinear tables, hydrostatic col_dry,
 * weights sum to 1, brute-force nearest nodes). */
static void interp_layer(const ccpprad_kdist *k, double play, double tlay, double dp,
                         const double *vmr /* [ngas], gas g at vmr[g-1] */, layer_interp *L)
{
    const int ng = k->ngas;
    double vh = vmr[IH2O - 1];
    double fact = 1.0 / (1.0 + vh);
    double m_air = (M_DRY + M_H2O * vh) * fact;
    double col_dry = 10.0 * dp * AVOGAD / (1000.0 * m_air * 100.0 * GRAV) * fact;
    L->col[0] = col_dry;
    for (int g = 1; g <= ng; ++g) L->col[g] = vmr[g - 1] * col_dry;

    double tmin = k->temp_ref[0], tmax = k->temp_ref[k->ntemp - 1];
    double dT = (tmax - tmin) / (double)(k->ntemp - 1);
    int jt = (int)((tlay - (tmin - dT)) / dT);
    if (jt < 1) jt = 1;
    if (jt > k->ntemp - 1) jt = k->ntemp - 1;
    L->jtemp = jt;
    L->ftemp = (tlay - k->temp_ref[jt - 1]) / dT;
    double lp1 = log(k->press_ref[0]);
    double dlp = (lp1 - log(k->press_ref[k->npres - 1])) / (double)(k->npres - 1);
    double locp = 1.0 + (lp1 - log(play)) / dlp;
    int jp = (int)locp;
    if (jp < 1) jp = 1;
    if (jp > k->npres - 1) jp = k->npres - 1;
    L->jpress = jp;
    L->fpress = locp - (double)jp;
    int tropo = log(play) > log(k->press_ref_trop);
    L->itropo = tropo ? 1 : 2;
    for (int f = 0; f < k->nflav; ++f) {
        int g1 = F2(k->flavor, 0, f, 2), g2 = F2(k->flavor, 1, f, 2);
        for (int t = 0; t < 2; ++t) {
            double r = F3(k->vmr_ref, L->itropo - 1, g1, jt - 1 + t, 2, ng + 1) /
                       F3(k->vmr_ref, L->itropo - 1, g2, jt - 1 + t, 2, ng + 1);
            double cm = L->col[g1] + r * L->col[g2];
            double eta = (cm > 2.0 * DBL_MIN) ? L->col[g1] / cm : 0.5;
            double loceta = eta * (double)(k->neta - 1);
            int je = (int)loceta + 1;
            if (je > k->neta - 1) je = k->neta - 1;
            double feta = fmod(loceta, 1.0);
            double ftt = (t == 0) ? (1.0 - L->ftemp) : L->ftemp;
            L->col_mix[f][t] = cm;
            L->jeta[f][t] = je;
            L->fminor[f][t][0] = (1.0 - feta) * ftt;
            L->fminor[f][t][1] = feta * ftt;
        }
    }
}

/* interpolate3D on a (ngpt, neta, npres+1, ntemp) table; jp is jpress+itropo. */
static real interp3d(const ccpprad_kdist *k, const double *tab, const layer_interp *L,
                     int f, int g, const double scale[2])
{
    const int ng = k->ngpt, ne = k->neta, np1 = k->npres + 1;
    int jp = L->jpress + L->itropo; /* 1-based upper row index */
    real fp0 = R(1.0 - L->fpress), fp1 = R(L->fpress);
    real res = R(0.0);
    for (int t = 0; t < 2; ++t) {
        int je = L->jeta[f][t], jt = L->jtemp + t;
        real fe0 = R(L->fminor[f][t][0]), fe1 = R(L->fminor[f][t][1]);
        real s = fp0 * fe0 * R(F4(tab, g, je - 1, jp - 2, jt - 1, ng, ne, np1)) +
                 fp1 * fe0 * R(F4(tab, g, je - 1, jp - 1, jt - 1, ng, ne, np1)) +
                 fp0 * fe1 * R(F4(tab, g, je, jp - 2, jt - 1, ng, ne, np1)) +
                 fp1 * fe1 * R(F4(tab, g, je, jp - 1, jt - 1, ng, ne, np1));
        res += R(scale[t]) * s;
    }
    return res;
}

/* interpolate2D on (ncontrib, neta, ntemp) or (ngpt, neta, ntemp). */
static real interp2d(int n1, int neta, const double *tab, const layer_interp *L, int f, int idx)
{
    real res = R(0.0);
    for (int t = 0; t < 2; ++t) {
        int je = L->jeta[f][t], jt = L->jtemp + t;
        res += R(L->fminor[f][t][0]) * R(F3(tab, idx, je - 1, jt - 1, n1, neta)) +
               R(L->fminor[f][t][1]) * R(F3(tab, idx, je, jt - 1, n1, neta));
    }
    return res;
}

/* Major + minor absorption (and Planck fraction / Rayleigh) for one layer. */
static void gas_optics_layer(const ccpprad_kdist *k, const layer_interp *L, double play,
                             double tlay, real *tau, real *pfrac, real *tau_ray)
{
    const int ng = k->ngpt;
    const int it = L->itropo;
    for (int g = 0; g < ng; ++g) {
        int f = F2(k->gpoint_flavor, it - 1, g, 2) - 1;
        tau[g] = interp3d(k, k->kmajor, L, f, g, L->col_mix[f]);
        if (pfrac) {
            static const double one[2] = {1.0, 1.0};
            pfrac[g] = interp3d(k, k->planck_frac, L, f, g, one);
        }
    }
    const ccpprad_minor *m = (it == 1) ? &k->minor_lower : &k->minor_upper;
    for (int i = 0; i < m->n; ++i) {
        double scaling = L->col[m->idx_minor[i]];
        if (m->scales_with_density[i]) {
            scaling *= 0.01 * play / tlay;
            if (m->idx_scaling[i] > 0) {
                double vmr_fact = 1.0 / L->col[0];
                double dry_fact = 1.0 / (1.0 + L->col[IH2O] * vmr_fact);
                if (m->scale_by_complement[i])
                    scaling *= (1.0 - L->col[m->idx_scaling[i]] * vmr_fact * dry_fact);
                else
                    scaling *= L->col[m->idx_scaling[i]] * vmr_fact * dry_fact;
            }
        }
        int gs = F2(m->limits_gpt, 0, i, 2) - 1, ge = F2(m->limits_gpt, 1, i, 2) - 1;
        int f = F2(k->gpoint_flavor, it - 1, gs, 2) - 1;
        for (int g = gs; g <= ge; ++g)
            tau[g] += R(scaling) * interp2d(m->ncontrib, k->neta, m->kminor, L, f,
                                            m->kminor_start[i] - 1 + (g - gs));
    }
    if (tau_ray) {
        const double *kr = (it == 1) ? k->krayl_lower : k->krayl_upper;
        double cr = L->col[IH2O] + L->col[0];
        for (int g = 0; g < ng; ++g) {
            int f = F2(k->gpoint_flavor, it - 1, g, 2) - 1;
            tau_ray[g] = interp2d(ng, k->neta, kr, L, f, g) * R(cr);
        }
    }
}

/* vmr: (ncol, nlay, ngas).  Outputs (ngpt, nlay, ncol).  pfrac LW only,
 * tau_ray SW only (NULL otherwise). */
void orc_gas_optics(const ccpprad_kdist *k, int ncol, int nlay, const double *p_lay,
                    const double *p_lev, const double *t_lay, const double *vmr, real *tau,
                    real *pfrac, real *tau_ray)
{
    const int ng = k->ngpt;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ncol; ++c) {
        layer_interp L;
        double v[CCPPRAD_MAX_GAS];
        for (int l = 0; l < nlay; ++l) {
            for (int g = 0; g < k->ngas; ++g) v[g] = F3(vmr, c, l, g, ncol, nlay);
            double dp = fabs(F2(p_lev, c, l, ncol) - F2(p_lev, c, l + 1, ncol));
            double pl = F2(p_lay, c, l, ncol), tl = F2(t_lay, c, l, ncol);
            interp_layer(k, pl, tl, dp, v, &L);
            size_t o = (size_t)ng * ((size_t)l + (size_t)nlay * c);
            gas_optics_layer(k, &L, pl, tl, tau + o, pfrac ? pfrac + o : NULL,
                             tau_ray ? tau_ray + o : NULL);
        }
    }
}

/* Expose the per-layer interpolation for tests (one layer). */
void orc_interp_layer(const ccpprad_kdist *k, double play, double tlay, double dp,
                      const double *vmr, int *jtemp, int *jpress, int *itropo, double *ftemp,
                      double *fpress, double *col /*[ngas+1]*/, int *jeta /*[nflav][2]*/,
                      double *col_mix /*[nflav][2]*/, double *fminor /*[nflav][2][2]*/)
{
    layer_interp L;
    interp_layer(k, play, tlay, dp, vmr, &L);
    *jtemp = L.jtemp; *jpress = L.jpress; *itropo = L.itropo;
    *ftemp = L.ftemp; *fpress = L.fpress;
    for (int g = 0; g <= k->ngas; ++g) col[g] = L.col[g];
    for (int f = 0; f < k->nflav; ++f)
        for (int t = 0; t < 2; ++t) {
            jeta[2 * f + t] = L.jeta[f][t];
            col_mix[2 * f + t] = L.col_mix[f][t];
            fminor[4 * f + 2 * t] = L.fminor[f][t][0];
            fminor[4 * f + 2 * t + 1] = L.fminor[f][t][1];
        }
}

/* interpolate1D on the band Planck table. */
static double planck1d(const ccpprad_kdist *k, int b, double T)
{
    double tmin = k->temp_ref[0], tmax = k->temp_ref[k->ntemp - 1];
    double delta = (tmax - tmin) / (double)(k->nplanck - 1);
    double v0 = (T - tmin) / delta;
    double frac = v0 - (double)(int)v0;
    int idx = (int)v0 + 1;
    if (idx < 1) idx = 1;
    if (idx > k->nplanck - 1) idx = k->nplanck - 1;
    const double *tab = k->totplnk + (size_t)k->nplanck * b;
    return tab[idx - 1] + frac * (tab[idx] - tab[idx - 1]);
}

double orc_planck_band(const ccpprad_kdist *k, int b, double T) { return planck1d(k, b, T); }

/* compute_Planck_source: pfrac (ngpt,nlay,ncol) -> sources.  Pinned by the
 * band-sum invariant sum_{g in b} lay_src = B_b(T) and the Jacobian slope. */
void orc_lw_sources(const ccpprad_kdist *k, int ncol, int nlay, int top_at_1,
                    const double *t_lay, const double *t_lev, const double *tsfc,
                    const real *pfrac, real *lay_src, real *lev_inc, real *lev_dec,
                    real *sfc_src, real *sfc_jac)
{
    const int ng = k->ngpt, nb = k->nband;
    const int sfc_lay = top_at_1 ? nlay - 1 : 0;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ncol; ++c) {
        double pf0[MAXB], pf1[MAXB];
        for (int b = 0; b < nb; ++b) {
            pf0[b] = planck1d(k, b, tsfc[c]);
            pf1[b] = planck1d(k, b, tsfc[c] + 1.0);
        }
        for (int b = 0; b < nb; ++b)
            for (int g = F2(k->band_lims_gpt, 0, b, 2) - 1; g < F2(k->band_lims_gpt, 1, b, 2); ++g) {
                real p = pfrac[g + (size_t)ng * (sfc_lay + (size_t)nlay * c)];
                sfc_src[g + (size_t)ng * c] = p * R(pf0[b]);
                sfc_jac[g + (size_t)ng * c] = p * R(pf1[b] - pf0[b]);
            }
        double plo[MAXB], phi[MAXB], play_[MAXB];
        for (int b = 0; b < nb; ++b) plo[b] = planck1d(k, b, F2(t_lev, c, 0, ncol));
        for (int l = 0; l < nlay; ++l) {
            for (int b = 0; b < nb; ++b) {
                phi[b] = planck1d(k, b, F2(t_lev, c, l + 1, ncol));
                play_[b] = planck1d(k, b, F2(t_lay, c, l, ncol));
            }
            for (int b = 0; b < nb; ++b)
                for (int g = F2(k->band_lims_gpt, 0, b, 2) - 1; g < F2(k->band_lims_gpt, 1, b, 2); ++g) {
                    size_t o = g + (size_t)ng * (l + (size_t)nlay * c);
                    lay_src[o] = pfrac[o] * R(play_[b]);
                    lev_inc[o] = pfrac[o] * R(phi[b]);
                    lev_dec[o] = pfrac[o] * R(plo[b]);
                }
            for (int b = 0; b < nb; ++b) plo[b] = phi[b];
        }
    }
}

/* ===================================================================== */
/* LW solver                                                              */
/* ===================================================================== */
/* Gauss-Jacobi(0,1) nodes for int_0^1 f(mu) mu dmu as secants D = 1/mu, with
 * weights w (sum w = 1/2), ordered by increasing D; n = 1 is the diffusivity
 * convention (1.66, 0.5) of rte_lw.  Nodes: roots of the monic Jacobi
 * polynomials p_k on [-1,1] (alpha=0, beta=1; a_k = 1/((2k+1)(2k+3)),
 * b_k = k(k+1)/(2k+1)^2) by bisection; weights by Christoffel numbers;
 * mu = (x+1)/2, w = w_x/4.  Pinned by tests/test_oracle_lw.py against
 * scipy.special.roots_jacobi and the closed form for n = 2. */
static double jac_p(int n, double x, double *pk /* [n] values p_0..p_{n-1} */)
{
    double pm = 0.0, pc = 1.0;
    for (int k = 0; k < n; ++k) {
        if (pk) pk[k] = pc;
        double ak = 1.0 / ((2.0 * k + 1.0) * (2.0 * k + 3.0));
        double bk = (k == 0) ? 0.0 : (double)k * (k + 1.0) / ((2.0 * k + 1.0) * (2.0 * k + 1.0));
        double pn = (x - ak) * pc - bk * pm;
        pm = pc; pc = pn;
    }
    return pc;
}

void orc_gauss(int n, double *D, double *w)
{
    if (n <= 1) { D[0] = 1.66; w[0] = 0.5; return; }
    double xr[16];
    int nr = 0;
    const int NS = 20000;
    double xa = -1.0, fa = jac_p(n, xa, NULL);
    for (int s = 1; s <= NS && nr < n; ++s) {
        double xb = -1.0 + 2.0 * s / NS, fb = jac_p(n, xb, NULL);
        if ((fa < 0) != (fb < 0)) {
            double lo = xa, hi = xb, flo = fa;
            for (int it = 0; it < 200; ++it) {
                double mid = 0.5 * (lo + hi), fm = jac_p(n, mid, NULL);
                if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else hi = mid;
                if (hi - lo < 1e-17) break;
            }
            xr[nr++] = 0.5 * (lo + hi);
        }
        xa = xb; fa = fb;
    }
    /* norms h_k of monic p_k for weight (1+x) on [-1,1]: h_0 = 2, h_k = h_{k-1} b_k */
    double h[16];
    h[0] = 2.0;
    for (int k = 1; k < n; ++k) h[k] = h[k - 1] * (double)k * (k + 1.0) / ((2.0 * k + 1.0) * (2.0 * k + 1.0));
    double mu[16], wx[16];
    for (int i = 0; i < nr; ++i) {
        double pk[16], s = 0.0;
        jac_p(n, xr[i], pk);
        for (int k = 0; k < n; ++k) s += pk[k] * pk[k] / h[k];
        wx[i] = 1.0 / s;
        mu[i] = 0.5 * (xr[i] + 1.0);
    }
    /* increasing D = decreasing mu: roots were found in increasing x */
    for (int i = 0; i < nr; ++i) {
        D[i] = 1.0 / mu[nr - 1 - i];
        w[i] = 0.25 * wx[nr - 1 - i];
    }
}

/* lw_source_noscat + lw_transport_noscat for one (column, g-point, angle).
 * Adds 2*pi*w*I into up/dn (nlev) and the Jacobian into jac.
 * tau, lay, inc, dec are strided by ngpt over layers. */
static void lw_one(int nlay, int ngpt, int top_at_1, real D, real w, const real *tau,
                   const real *lay, const real *inc, const real *dec, real emis,
                   real sfc, real sfcj, real *up, real *dn, real *jac, real *tr,
                   real *su, real *sd, real *Iu, real *Id)
{
    const real tau_thresh = sqrt(R_EPS);
    for (int l = 0; l < nlay; ++l) {
        real t = tau[(size_t)l * ngpt] * D;
        real trans = exp(-t);
        real fact = (t > tau_thresh) ? (R(1.0) - trans) / t - trans : t * (R(0.5) - t / R(3.0));
        real ls = lay[(size_t)l * ngpt];
        real lup = top_at_1 ? dec[(size_t)l * ngpt] : inc[(size_t)l * ngpt];
        real ldn = top_at_1 ? inc[(size_t)l * ngpt] : dec[(size_t)l * ngpt];
        tr[l] = trans;
        sd[l] = (R(1.0) - trans) * ldn + R(2.0) * fact * (ls - ldn);
        su[l] = (R(1.0) - trans) * lup + R(2.0) * fact * (ls - lup);
    }
    real alb = R(1.0) - emis, ssrc = emis * sfc;
    if (top_at_1) {
        Id[0] = R(0.0);
        for (int v = 1; v <= nlay; ++v) Id[v] = tr[v - 1] * Id[v - 1] + sd[v - 1];
        Iu[nlay] = Id[nlay] * alb + ssrc;
        for (int v = nlay - 1; v >= 0; --v) Iu[v] = tr[v] * Iu[v + 1] + su[v];
    } else {
        Id[nlay] = R(0.0);
        for (int v = nlay - 1; v >= 0; --v) Id[v] = tr[v] * Id[v + 1] + sd[v];
        Iu[0] = Id[0] * alb + ssrc;
        for (int v = 1; v <= nlay; ++v) Iu[v] = tr[v - 1] * Iu[v - 1] + su[v - 1];
    }
    real f = R(2.0 * M_PI) * w;
    for (int v = 0; v <= nlay; ++v) { up[v] += f * Iu[v]; dn[v] += f * Id[v]; }
    if (jac) {
        real J = emis * sfcj;
        if (top_at_1) {
            Iu[nlay] = J;
            for (int v = nlay - 1; v >= 0; --v) Iu[v] = tr[v] * Iu[v + 1];
        } else {
            Iu[0] = J;
            for (int v = 1; v <= nlay; ++v) Iu[v] = tr[v - 1] * Iu[v - 1];
        }
        for (int v = 0; v <= nlay; ++v) jac[v] += f * Iu[v];
    }
}

/* lw_solver_noscat with n Gauss angles, or per-(g,col) secants Ds when Ds
 * != NULL (rte_lw with lw_Ds, weight gauss_wts(1,1) = 0.5).  Band/broadband
 * reduction (sum_byband then the CCPP sum over bands) into flux_up/flux_dn
 * (ncol, nlev) and optional flux_jac.  emis_b (nband, ncol). */
void orc_lw_solve(const ccpprad_kdist *k, int ncol, int nlay, int top_at_1, int nang,
                  const real *Ds, const real *tau, const real *lay_src,
                  const real *lev_inc, const real *lev_dec, const double *emis_b,
                  const real *sfc_src, const real *sfc_jac, double *flux_up,
                  double *flux_dn, double *flux_jac)
{
    const int ng = k->ngpt, nb = k->nband, nlev = nlay + 1;
    double Dq[8], wq[8];
    orc_gauss(nang, Dq, wq);
#pragma omp parallel
    {
        real *gup = malloc(sizeof(real) * nlev * ng), *gdn = malloc(sizeof(real) * nlev * ng);
        real *gj = malloc(sizeof(real) * nlev * ng);
        real *tr = malloc(sizeof(real) * nlay), *su = malloc(sizeof(real) * nlay);
        real *sd = malloc(sizeof(real) * nlay);
        real *Iu = malloc(sizeof(real) * nlev), *Id = malloc(sizeof(real) * nlev);
#pragma omp for schedule(static)
        for (int c = 0; c < ncol; ++c) {
            memset(gup, 0, sizeof(real) * nlev * ng);
            memset(gdn, 0, sizeof(real) * nlev * ng);
            memset(gj, 0, sizeof(real) * nlev * ng);
            for (int b = 0; b < nb; ++b) {
                real emis = R(F2(emis_b, b, c, nb));
                for (int g = F2(k->band_lims_gpt, 0, b, 2) - 1; g < F2(k->band_lims_gpt, 1, b, 2); ++g) {
                    size_t o = g + (size_t)ng * (size_t)nlay * c;
                    real s = sfc_src[g + (size_t)ng * c], sj = sfc_jac[g + (size_t)ng * c];
                    real *u = gup + (size_t)g * nlev, *d = gdn + (size_t)g * nlev, *j = gj + (size_t)g * nlev;
                    if (Ds) {
                        lw_one(nlay, ng, top_at_1, Ds[g + (size_t)ng * c], R(0.5), tau + o, lay_src + o,
                               lev_inc + o, lev_dec + o, emis, s, sj, u, d, flux_jac ? j : NULL, tr, su, sd, Iu, Id);
                    } else {
                        for (int a = 0; a < nang; ++a)
                            lw_one(nlay, ng, top_at_1, R(Dq[a]), R(wq[a]), tau + o, lay_src + o, lev_inc + o,
                                   lev_dec + o, emis, s, sj, u, d, flux_jac ? j : NULL, tr, su, sd, Iu, Id);
                    }
                }
            }
            for (int v = 0; v < nlev; ++v) {
                double su_ = 0.0, sd_ = 0.0, sj_ = 0.0;
                for (int b = 0; b < nb; ++b) {
                    int g0 = F2(k->band_lims_gpt, 0, b, 2) - 1, g1 = F2(k->band_lims_gpt, 1, b, 2);
                    double bu = (double)gup[(size_t)g0 * nlev + v], bd = (double)gdn[(size_t)g0 * nlev + v];
                    double bj = (double)gj[(size_t)g0 * nlev + v];
                    for (int g = g0 + 1; g < g1; ++g) {
                        bu += (double)gup[(size_t)g * nlev + v];
                        bd += (double)gdn[(size_t)g * nlev + v];
                        bj += (double)gj[(size_t)g * nlev + v];
                    }
                    su_ += bu; sd_ += bd; sj_ += bj;
                }
                F2(flux_up, c, v, ncol) = su_;
                F2(flux_dn, c, v, ncol) = sd_;
                if (flux_jac) F2(flux_jac, c, v, ncol) = sj_;
            }
        }
        free(gup); free(gdn); free(gj); free(tr); free(su); free(sd); free(Iu); free(Id);
    }
}

/* ===================================================================== */
/* GFS_rrtmgp_pre_run's preparation of the radiation inputs               */
/* (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_pre.F90:217-274,     */
/* 349-371) with cmp_tlev (physics/Radiation/radiation_tools.F90:16-84).   */
/* Fortran (ncol, nlev[+1]) arrays; nlev = layers.  qo3 may be NULL (then  */
/* vmr_o3 is not written).  Returns top_at_1.  cmp_tlev is pinned against  */
/* the reference's own Fortran (oracle/_ref); the rest by invariants.      */
/* ===================================================================== */
int orc_prepare(int ncol, int nlev, double pmin, double pmax, double tmin, double tmax,
                const double *prsi, const double *prsl, const double *tgrs, const double *qh2o,
                const double *qo3, const double *tsfc, double *p_lay, double *p_lev, double *t_lay,
                double *t_lev, double *q_lay, double *vmr_h2o, double *vmr_o3)
{
    const double amd = 28.9644, amw = 18.0154, amo3 = 47.9982, con_epsqs = 1.0e-10;
    const double amdw = amd / amw, amdo3 = amd / amo3, eps = DBL_EPSILON;
    const int top_at_1 = prsi[0] < prsi[(size_t)ncol * (nlev - 1)];
    const size_t nl = (size_t)ncol * nlev, nv = (size_t)ncol * (nlev + 1);
    for (size_t i = 0; i < nl; ++i) {
        q_lay[i] = qh2o[i] < 1.0e-6 ? 1.0e-6 : qh2o[i];
        p_lay[i] = prsl[i];
        t_lay[i] = tgrs[i];
        if (t_lay[i] <= tmin) t_lay[i] = tmin + eps;
        if (p_lay[i] <= pmin) p_lay[i] = pmin + eps;
        if (t_lay[i] >= tmax) t_lay[i] = tmax - eps;
        if (p_lay[i] >= pmax) p_lay[i] = pmax - eps;
    }
    for (size_t i = 0; i < nv; ++i) p_lev[i] = prsi[i];
    /* cmp_tlev(nCol, nLev, minP, p_lay, t_lay, p_lev, tsfc, t_lev) */
    double *a = malloc(sizeof(double) * nl), *b = malloc(sizeof(double) * nl);
    for (int c = 0; c < ncol; ++c) {
        for (int k = 0; k < nlev; ++k) {
            F2(a, c, k, ncol) = log(F2(p_lay, c, k, ncol));
            F2(b, c, k, ncol) = log(F2(p_lev, c, k, ncol));
        }
        if (top_at_1) {
            F2(b, c, 0, ncol) = log(fmax(pmin, F2(p_lev, c, 0, ncol)));
            F2(t_lev, c, 0, ncol) = F2(t_lay, c, 0, ncol);
            for (int k = 1; k < nlev; ++k)
                F2(t_lev, c, k, ncol) = F2(t_lay, c, k, ncol) + (F2(t_lay, c, k - 1, ncol) - F2(t_lay, c, k, ncol)) *
                                        (F2(b, c, k, ncol) - F2(a, c, k, ncol)) /
                                        (F2(a, c, k - 1, ncol) - F2(a, c, k, ncol));
            F2(t_lev, c, nlev, ncol) = tsfc[c];
        } else {
            F2(b, c, nlev - 1, ncol) = log(fmax(pmin, F2(p_lev, c, nlev - 1, ncol)));
            F2(t_lev, c, 0, ncol) = tsfc[c];
            for (int k = 0; k < nlev - 1; ++k)
                F2(t_lev, c, k + 1, ncol) = F2(t_lay, c, k, ncol) + (F2(t_lay, c, k + 1, ncol) - F2(t_lay, c, k, ncol)) *
                                            (F2(b, c, k + 1, ncol) - F2(a, c, k, ncol)) /
                                            (F2(a, c, k + 1, ncol) - F2(a, c, k, ncol));
            F2(t_lev, c, nlev, ncol) = F2(t_lay, c, nlev - 1, ncol);
        }
    }
    free(a);
    free(b);
    for (size_t i = 0; i < nv; ++i) {
        if (t_lev[i] <= tmin) t_lev[i] = tmin + eps;
        if (t_lev[i] >= tmax) t_lev[i] = tmax - eps;
    }
    for (size_t i = 0; i < nl; ++i) {
        vmr_h2o[i] = q_lay[i] != 1.0 ? (q_lay[i] / (1.0 - q_lay[i])) * amdw : 0.0;
        if (qo3) {
            const double o = fmax(con_epsqs, qo3[i]);
            vmr_o3[i] = o > 0.0 ? o * amdo3 : 0.0;
        }
    }
    return top_at_1;
}

/* GFS_rrtmgp_post_run's flux bookkeeping (GFS_rrtmgp_post.F90:173-389, the
 * heating rates aside) over host arrays, statement by statement as the
 * reference's loops.  Pinned by hand cases (tests/test_oracle_prepare.py):
 * it is data movement and time-weighted sums, with no arithmetic of its own
 * beyond them. */
/* GFS_rrtmgp_pre_run's virtual temperature and layer thicknesses
 * (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_pre.F90:289,294-341), in
 * the reference's loop structure: per column, deltaZ, the boundary heights
 * from the surface, the centre heights, then the thicknesses between
 * centres.  No FMA contraction, so the statements round as written. */
__attribute__((optimize("fp-contract=off")))
void orc_layer_geometry(int ncol, int nlev, double con_rd, double con_g, double con_fvirt,
                        const double *p_lev, const double *p_lay, const double *t_lay,
                        const double *q_lay, double *tv_lay, double *dz, double *dzc, double *dp)
{
#define L2(p, i, k) (p)[(size_t)(i) + (size_t)ncol * ((k) - 1)]
    const double rdog = con_rd / con_g;
    const int top_at_1 = p_lev[0] < p_lev[(size_t)ncol * (nlev - 1)];
    double *hgtb = (double *)malloc(sizeof(double) * (nlev + 2));
    double *hgtc = (double *)malloc(sizeof(double) * (nlev + 1));
    for (int k = 1; k <= nlev; ++k)
        for (int i = 0; i < ncol; ++i) {
            L2(tv_lay, i, k) = L2(t_lay, i, k) * (1.0 + con_fvirt * L2(q_lay, i, k));
            L2(dp, i, k) = fabs(L2(p_lev, i, k + 1) - L2(p_lev, i, k));
        }
    for (int i = 0; i < ncol; ++i) {
        if (top_at_1) {
            for (int k = 1; k <= nlev; ++k)
                L2(dz, i, k) = rdog * fabs(log(L2(p_lev, i, k + 1)) - log(L2(p_lev, i, k))) * L2(tv_lay, i, k);
            hgtb[nlev + 1] = 0.0;
            for (int k = nlev; k >= 1; --k) hgtb[k] = hgtb[k + 1] + L2(dz, i, k);
            for (int k = nlev; k >= 1; --k) {
                const double pfac = fabs(log(L2(p_lev, i, k + 1)) - log(L2(p_lay, i, k))) /
                                    fabs(log(L2(p_lev, i, k + 1)) - log(L2(p_lev, i, k)));
                hgtc[k] = hgtb[k + 1] + pfac * (hgtb[k] - hgtb[k + 1]);
            }
            for (int k = nlev - 1; k >= 1; --k) L2(dzc, i, k) = hgtc[k] - hgtc[k + 1];
            L2(dzc, i, nlev) = hgtc[nlev] - hgtb[nlev + 1];
        } else {
            for (int k = nlev; k >= 1; --k)
                L2(dz, i, k) = rdog * fabs(log(L2(p_lev, i, k)) - log(L2(p_lev, i, k + 1))) * L2(tv_lay, i, k);
            hgtb[1] = 0.0;
            for (int k = 1; k <= nlev; ++k) hgtb[k + 1] = hgtb[k] + L2(dz, i, k);
            for (int k = 1; k <= nlev; ++k) {
                const double pfac = fabs(log(L2(p_lev, i, k)) - log(L2(p_lay, i, k))) /
                                    fabs(log(L2(p_lev, i, k)) - log(L2(p_lev, i, k + 1)));
                hgtc[k] = hgtb[k] + pfac * (hgtb[k + 1] - hgtb[k]);
            }
            for (int k = 2; k <= nlev; ++k) L2(dzc, i, k) = hgtc[k] - hgtc[k - 1];
            L2(dzc, i, 1) = hgtc[1] - hgtb[1];
        }
    }
    free(hgtb);
    free(hgtc);
#undef L2
}

/* GFS_rrtmgp_cloud_overlap_run (GFS_rrtmgp_cloud_overlap.F90:97-137) with
 * cmp_dcorr_lgth_hogan (radiation_cloud_overlap.F90:28-50),
 * cmp_dcorr_lgth_oreopoulos (:58-95) and get_alpha_exper (:100-143).  The
 * unsuffixed literals take the precision literal_r8 selects. */
__attribute__((optimize("fp-contract=off")))
void orc_cloud_overlap(const ccpprad_cloud_overlap_args *a)
{
    const int nc = a->nCol, nl = a->nLev;
    const int r8 = a->literal_r8 != 0;
    if (!(a->doSWrad || a->doLWrad)) return;
    for (int i = 0; i < nc; ++i) a->de_lgth[i] = 0.0;
    if (a->idcor == a->idcor_hogan) {
        const double min_dcorr = r8 ? 0.6 : (double)0.6f;
        const double c1 = r8 ? 2.78 : (double)2.78f, c2 = r8 ? 4.6 : (double)4.6f;
        for (int i = 0; i < nc; ++i) {
            const double v = c1 - c2 * fabs(a->lat[i] / a->con_pi);
            a->de_lgth[i] = min_dcorr > v ? min_dcorr : v;   /* max(min_dcorr, v) */
        }
    }
    if (a->idcor == a->idcor_oreopoulos) {
        const double am1 = 1.4315, am2 = 2.1219, am4 = -25.584, amr = 7.0;
        const double f = 180.0 / a->con_pi;
        for (int i = 0; i < nc; ++i) {
            const double latd = a->lat[i] * f;
            double am3;
            if (a->julian > 181.0) am3 = -4.0 * amr * (a->julian - 272.0) / a->yearlen;
            else am3 = 4.0 * amr * (a->julian - 91.0) / a->yearlen;
            const double x = latd - am3;
            a->de_lgth[i] = am1 + am2 * exp(-(x * x) / (am4 * am4));
        }
    }
    if (a->idcor == a->idcor_con)
        for (int i = 0; i < nc; ++i) a->de_lgth[i] = a->dcorr_con;
    const double km = r8 ? 0.001 : (double)0.001f;
#define OV(p, i, k) (p)[(size_t)(i) + (size_t)nc * ((k) - 1)]
    /* get_alpha_exper, for the clouds and then the convective clouds */
    for (int pass = 0; pass < 2; ++pass) {
        const int iovr = pass == 0 ? a->iovr : a->iovr_convcld;
        const double *cf = pass == 0 ? a->cld_frac : a->cld_cnv_frac;
        double *out = pass == 0 ? a->cloud_overlap_param : a->cnv_cloud_overlap_param;
        if (pass == 1 && !(a->imfdeepcnv == a->imfdeepcnv_samf || a->imfdeepcnv == a->imfdeepcnv_gf)) break;
        if (iovr == a->iovr_dcorr || iovr == a->iovr_exp || iovr == a->iovr_exprand) {
            for (int i = 0; i < nc; ++i) {
                OV(out, i, 1) = 0.0;
                for (int k = 2; k <= nl; ++k) OV(out, i, k) = exp(-(OV(a->deltaZc, i, k) * km) / a->de_lgth[i]);
            }
            if (iovr == a->iovr_exprand)
                for (int k = 2; k <= nl; ++k)
                    for (int i = 0; i < nc; ++i)
                        if (OV(cf, i, k) == 0.0 && OV(cf, i, k - 1) > 0.0) OV(out, i, k) = 0.0;
        } else {
            for (size_t j = 0; j < (size_t)nc * nl; ++j) out[j] = 0.0;
        }
    }
    for (size_t j = 0; j < (size_t)nc * nl; ++j) a->precip_overlap_param[j] = a->cloud_overlap_param[j];
#undef OV
}

void orc_post(const ccpprad_post_args *a)
{
    const int nc = a->nCol, nl = a->nLev;
#define FX(i, j) a->fluxr[(size_t)(i) + (size_t)nc * ((j) - 1)]
#define C2(p, i, k) (p)[(size_t)(i) + (size_t)nc * ((k) - 1)]
    if (!(a->doLWrad || a->doSWrad)) return;
    if (a->doLWrad) {
        for (int i = 0; i < nc; ++i) {
            a->topflw[2 * i + 0] = C2(a->fluxlwUP_allsky, i, a->iTOA);
            a->topflw[2 * i + 1] = C2(a->fluxlwUP_clrsky, i, a->iTOA);
            a->sfcflw[4 * i + 0] = C2(a->fluxlwUP_allsky, i, a->iSFC);
            a->sfcflw[4 * i + 1] = C2(a->fluxlwUP_clrsky, i, a->iSFC);
            a->sfcflw[4 * i + 2] = C2(a->fluxlwDOWN_allsky, i, a->iSFC);
            a->sfcflw[4 * i + 3] = C2(a->fluxlwDOWN_clrsky, i, a->iSFC);
            a->tsflw[i] = a->tsfa[i];
            a->sfcdlw[i] = a->sfcflw[4 * i + 2];
            a->sfculw[i] = a->sfcflw[4 * i + 0];
        }
        for (size_t q = 0; q < (size_t)nc * nl; ++q) a->htrlwu[q] = a->htrlw[q];
        if (a->save_diag) {
            for (int i = 0; i < nc; ++i) {
                FX(i, 1) = FX(i, 1) + a->fhlwr * C2(a->fluxlwUP_allsky, i, a->iTOA);
                FX(i, 19) = FX(i, 19) + a->fhlwr * C2(a->fluxlwDOWN_allsky, i, a->iSFC);
                FX(i, 20) = FX(i, 20) + a->fhlwr * C2(a->fluxlwUP_allsky, i, a->iSFC);
                FX(i, 28) = FX(i, 28) + a->fhlwr * C2(a->fluxlwUP_clrsky, i, a->iTOA);
                FX(i, 30) = FX(i, 30) + a->fhlwr * C2(a->fluxlwDOWN_clrsky, i, a->iSFC);
                FX(i, 33) = FX(i, 33) + a->fhlwr * C2(a->fluxlwUP_clrsky, i, a->iSFC);
            }
            for (int j = 1; j <= 3; ++j)
                for (int i = 0; i < nc; ++i) {
                    double tem0d = a->raddt * C2(a->cldsa, i, j);
                    int itop = C2(a->mtopa, i, j), ibtc = C2(a->mbota, i, j);
                    double tem2 = 0.0;
                    for (int k = ibtc; k <= itop; ++k) tem2 = tem2 + C2(a->cldtaulw, i, k);
                    FX(i, 46 - j) = FX(i, 46 - j) + tem0d * (1.0 - exp(-tem2));
                }
        }
    }
    if (a->doSWrad) {
        for (int i = 0; i < nc; ++i) {
            double *tf = a->topfsw + 3 * (size_t)i, *sf = a->sfcfsw + 4 * (size_t)i;
            const double *cm = a->scmpsw + 6 * (size_t)i;
            if (a->nDay > 0) {
                tf[0] = C2(a->fluxswUP_allsky, i, a->iTOA);
                tf[2] = C2(a->fluxswUP_clrsky, i, a->iTOA);
                tf[1] = C2(a->fluxswDOWN_allsky, i, a->iTOA);
                sf[0] = C2(a->fluxswUP_allsky, i, a->iSFC);
                sf[2] = C2(a->fluxswUP_clrsky, i, a->iSFC);
                sf[1] = C2(a->fluxswDOWN_allsky, i, a->iSFC);
                sf[3] = C2(a->fluxswDOWN_clrsky, i, a->iSFC);
                a->nirbmdi[i] = cm[2];
                a->nirdfdi[i] = cm[3];
                a->visbmdi[i] = cm[4];
                a->visdfdi[i] = cm[5];
                a->nirbmui[i] = cm[2] * a->sfc_alb_nir_dir[i];
                a->nirdfui[i] = cm[3] * a->sfc_alb_nir_dif[i];
                a->visbmui[i] = cm[4] * a->sfc_alb_uvvis_dir[i];
                a->visdfui[i] = cm[5] * a->sfc_alb_uvvis_dif[i];
            } else {
                for (int k = 0; k < 4; ++k) sf[k] = 0.0;
                for (int k = 0; k < 3; ++k) tf[k] = 0.0;
                a->nirbmdi[i] = a->nirdfdi[i] = a->visbmdi[i] = a->visdfdi[i] = 0.0;
                a->nirbmui[i] = a->nirdfui[i] = a->visbmui[i] = a->visdfui[i] = 0.0;
            }
        }
        for (int i = 0; i < nc; ++i) {
            a->sfcnsw[i] = a->sfcfsw[4 * i + 1] - a->sfcfsw[4 * i + 0];
            a->sfcdsw[i] = a->sfcfsw[4 * i + 1];
        }
        if (a->save_diag) {
            for (int i = 0; i < nc; ++i) {
                for (int j = 1; j <= 6; ++j) FX(i, 33 + j) = C2(a->aerodp, i, j);
                if (a->coszen[i] > 0.0) {
                    const double *tf = a->topfsw + 3 * (size_t)i, *sf = a->sfcfsw + 4 * (size_t)i;
                    const double *cm = a->scmpsw + 6 * (size_t)i;
                    double tem0d = a->fhswr * a->coszdg[i] / a->coszen[i];
                    FX(i, 2) = FX(i, 2) + tf[0] * tem0d;
                    FX(i, 3) = FX(i, 3) + sf[0] * tem0d;
                    FX(i, 4) = FX(i, 4) + sf[1] * tem0d;
                    FX(i, 21) = FX(i, 21) + cm[0] * tem0d;
                    FX(i, 22) = FX(i, 22) + cm[1] * tem0d;
                    FX(i, 23) = FX(i, 23) + tf[1] * tem0d;
                    FX(i, 24) = FX(i, 24) + a->visbmdi[i] * tem0d;
                    FX(i, 25) = FX(i, 25) + a->visdfdi[i] * tem0d;
                    FX(i, 26) = FX(i, 26) + a->nirbmdi[i] * tem0d;
                    FX(i, 27) = FX(i, 27) + a->nirdfdi[i] * tem0d;
                    FX(i, 29) = FX(i, 29) + tf[2] * tem0d;
                    FX(i, 31) = FX(i, 31) + sf[2] * tem0d;
                    FX(i, 32) = FX(i, 32) + sf[3] * tem0d;
                }
            }
            for (int i = 0; i < nc; ++i) {
                FX(i, 17) = FX(i, 17) + a->raddt * C2(a->cldsa, i, 4);
                FX(i, 18) = FX(i, 18) + a->raddt * C2(a->cldsa, i, 5);
            }
            for (int j = 1; j <= 3; ++j)
                for (int i = 0; i < nc; ++i) {
                    double tem0d = a->raddt * C2(a->cldsa, i, j);
                    int itop = C2(a->mtopa, i, j), ibtc = C2(a->mbota, i, j);
                    FX(i, 8 - j) = FX(i, 8 - j) + tem0d;
                    FX(i, 11 - j) = FX(i, 11 - j) + tem0d * C2(a->p_lev, i, itop);
                    FX(i, 14 - j) = FX(i, 14 - j) + tem0d * C2(a->p_lev, i, ibtc);
                    FX(i, 17 - j) = FX(i, 17 - j) + tem0d * C2(a->p_lev, i, itop);
                    double tem1 = 0.0;
                    for (int k = ibtc; k <= itop; ++k) tem1 = tem1 + C2(a->cldtausw, i, k);
                    FX(i, 43 - j) = FX(i, 43 - j) + tem0d * tem1;
                }
        }
    }
#undef FX
#undef C2
}

/* compute_optimal_angles: Ds(g,c) = fit1(b) * exp(-sum_l tau) + fit2(b). */
void orc_optimal_angles(const ccpprad_kdist *k, int ncol, int nlay, const real *tau, real *Ds)
{
    const int ng = k->ngpt;
    for (int c = 0; c < ncol; ++c)
        for (int b = 0; b < k->nband; ++b)
            for (int g = F2(k->band_lims_gpt, 0, b, 2) - 1; g < F2(k->band_lims_gpt, 1, b, 2); ++g) {
                real s = R(0.0);
                for (int l = 0; l < nlay; ++l) s += tau[g + (size_t)ng * (l + (size_t)nlay * c)];
                Ds[g + (size_t)ng * c] = R(F2(k->optimal_angle_fit, 0, b, 2)) * exp(-s) +
                                         R(F2(k->optimal_angle_fit, 1, b, 2));
            }
}

/* ===================================================================== */
/* Heating rate (compute_heating_rate), K/s, FP64                         */
/* ===================================================================== */
void orc_heating_rate(int ncol, int nlay, const double *up, const double *dn,
                      const double *p_lev, double *htr)
{
    for (int l = 0; l < nlay; ++l)
        for (int c = 0; c < ncol; ++c)
            F2(htr, c, l, ncol) = (F2(up, c, l + 1, ncol) - F2(up, c, l, ncol) -
                                   F2(dn, c, l + 1, ncol) + F2(dn, c, l, ncol)) * GRAV /
                                  (CP_DRY * (F2(p_lev, c, l + 1, ncol) - F2(p_lev, c, l, ncol)));
}

/* ===================================================================== */
/* MT19937 as in physics/Radiation/mersenne_twister.f:285-372             */
/* Pinned bit-exactly by tests/test_oracle_mcica.py against the reference */
/* Fortran file compiled with flang (oracle/_ref) and the committed        */
/* golden vectors tests/golden/mt19937_draws.json.                        */
/* ===================================================================== */
typedef struct { int mti; uint32_t mt[624]; } orc_mt;

void orc_mt_seed(orc_mt *s, int inseed)
{
    int ii = inseed;
    if (ii == 0) ii = 4357;
    s->mti = 624;
    s->mt[0] = (uint32_t)ii;
    for (int i = 1; i < 624; ++i) s->mt[i] = 69069u * s->mt[i - 1];
}

void orc_mt_draw(orc_mt *s, int n, double *out)
{
    const uint32_t UM = 0x80000000u, LM = 0x7fffffffu, MATA = 0x9908b0dfu;
    for (int j = 0; j < n; ++j) {
        if (s->mti >= 624) {
            for (int kk = 0; kk < 623; ++kk) {
                uint32_t y = (s->mt[kk] & UM) | (s->mt[kk + 1] & LM);
                s->mt[kk] = s->mt[(kk + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
            }
            uint32_t y = (s->mt[623] & UM) | (s->mt[0] & LM);
            s->mt[623] = s->mt[396] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
            s->mti = 0;
        }
        uint32_t y = s->mt[s->mti];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        int32_t ys = (int32_t)y;
        if (ys < 0)
            out[j] = ((double)ys + 4294967296.0) / (4294967296.0 - 1.0);
        else
            out[j] = (double)(float)ys / (4294967296.0 - 1.0);
        s->mti++;
    }
}

/* Convenience for tests: n draws after seeding. */
void orc_mt_draws(int seed, int n, double *out)
{
    orc_mt s;
    orc_mt_seed(&s, seed);
    orc_mt_draw(&s, n, out);
}

/* sampled_mask (rrtmgp_sampling.F90:84-182) for one column.
 * randoms/randoms2 (ngpt, nlay); cf (nlay) column-strided by cstride;
 * overlap (nlay-1) or NULL; mask (nlay, ngpt) as bytes [l + nlay*g]. */
static void sampled_mask_col(int ngpt, int nlay, const double *randoms, const double *cf,
                             size_t cstride, const double *overlap, size_t ostride,
                             const double *randoms2, unsigned char *mask)
{
    unsigned char cml[1024];
    double lr[MAXG];
    int any = 0, fst = -1, lst = -1;
    for (int l = 0; l < nlay; ++l) {
        cml[l] = cf[l * cstride] > 0.0;
        if (cml[l]) { any = 1; if (fst < 0) fst = l; lst = l; }
    }
    memset(mask, 0, (size_t)nlay * ngpt);
    if (!any) return;
    for (int g = 0; g < ngpt; ++g) lr[g] = randoms[g + (size_t)ngpt * fst];
    for (int g = 0; g < ngpt; ++g) mask[fst + (size_t)nlay * g] = lr[g] > (1.0 - cf[fst * cstride]);
    for (int l = fst + 1; l <= lst; ++l) {
        double cfl = cf[l * cstride];
        if (!overlap) {
            if (cml[l]) {
                if (!cml[l - 1]) for (int g = 0; g < ngpt; ++g) lr[g] = randoms[g + (size_t)ngpt * l];
                for (int g = 0; g < ngpt; ++g) mask[l + (size_t)nlay * g] = lr[g] > (1.0 - cfl);
            } else {
                for (int g = 0; g < ngpt; ++g) mask[l + (size_t)nlay * g] = 0;
            }
        } else {
            if (cml[l]) {
                if (cml[l - 1]) {
                    double rho = overlap[(l - 1) * ostride];
                    for (int g = 0; g < ngpt; ++g)
                        lr[g] = rho * (lr[g] - 0.5) +
                                sqrt(1.0 - rho * rho) * (randoms[g + (size_t)ngpt * l] - 0.5) + 0.5;
                } else {
                    for (int g = 0; g < ngpt; ++g) lr[g] = randoms[g + (size_t)ngpt * l];
                }
                for (int g = 0; g < ngpt; ++g) mask[l + (size_t)nlay * g] = lr[g] > (1.0 - cfl);
            }
            if (randoms2) {
                for (int g = 0; g < ngpt; ++g) {
                    if (randoms2[g + (size_t)ngpt * l] <= overlap[(l - 1) * ostride])
                        mask[l + (size_t)nlay * g] = randoms[g + (size_t)ngpt * (l - 1)] > (1.0 - cfl);
                    else
                        mask[l + (size_t)nlay * g] = randoms[g + (size_t)ngpt * l] > (1.0 - cfl);
                }
            }
        }
    }
}

/* Test entry: full McICA mask for ncol columns the way the scheme layer
 * builds it (rrtmgp_lw_main.F90:441-489): seeds, MT draws, sampled_mask.
 * seeds (ncol) are the ipseed values; cf (ncol,nlay); overlap (ncol,nlay-1
 * with leading dim ncol) or NULL; mask out (nlay, ngpt, ncol) bytes. */
void orc_mcica_mask(int ncol, int nlay, int ngpt, const int *seeds, const double *cf,
                    const double *overlap, int iovr, int iovr_max, int iovr_maxrand,
                    int iovr_rand, int iovr_dcorr, int iovr_exp, int iovr_exprand,
                    unsigned char *mask)
{
#pragma omp parallel
    {
        double *r3 = malloc(sizeof(double) * ngpt * nlay), *r2 = malloc(sizeof(double) * ngpt * nlay);
#pragma omp for schedule(static)
        for (int c = 0; c < ncol; ++c) {
            orc_mt s;
            orc_mt_seed(&s, seeds[c]);
            if (iovr == iovr_max) {
                orc_mt_draw(&s, ngpt, r3);
                for (int l = 1; l < nlay; ++l) memcpy(r3 + (size_t)ngpt * l, r3, sizeof(double) * ngpt);
            } else {
                for (int l = 0; l < nlay; ++l) orc_mt_draw(&s, ngpt, r3 + (size_t)ngpt * l);
            }
            unsigned char *m = mask + (size_t)nlay * ngpt * c;
            if (iovr == iovr_maxrand || iovr == iovr_rand || iovr == iovr_max) {
                sampled_mask_col(ngpt, nlay, r3, cf + c, ncol, NULL, 0, NULL, m);
            } else if (iovr == iovr_dcorr) {
                orc_mt_seed(&s, seeds[c]);
                orc_mt_draw(&s, ngpt * nlay, r2);
                sampled_mask_col(ngpt, nlay, r3, cf + c, ncol, overlap + c, ncol, r2, m);
            } else if (iovr == iovr_exp || iovr == iovr_exprand) {
                sampled_mask_col(ngpt, nlay, r3, cf + c, ncol, overlap + c, ncol, NULL, m);
            } else {
                memset(m, 0, (size_t)nlay * ngpt);
            }
        }
        free(r3); free(r2);
    }
}

/* ===================================================================== */
/* Cloud optics (compute_all_from_table + combine), FP64                  */
/* outputs (ncol, nlay, nband) 2-stream tau/ssa/g                         */
/* ===================================================================== */
static void lut_interp(const double *ext, const double *ssa, const double *asy, int nsz,
                       double lwr, double upr, double path, double re, double *t,
                       double *ts, double *tsg)
{
    double step = (upr - lwr) / (double)(nsz - 1);
    int idx = (int)floor((re - lwr) / step) + 1;
    if (idx > nsz - 1) idx = nsz - 1;
    double fint = (re - lwr) / step - (double)(idx - 1);
    double tt = path * (ext[idx - 1] + fint * (ext[idx] - ext[idx - 1]));
    double s = tt * (ssa[idx - 1] + fint * (ssa[idx] - ssa[idx - 1]));
    *tsg = s * (asy[idx - 1] + fint * (asy[idx] - asy[idx - 1]));
    *ts = s;
    *t = tt;
}

/* rte-rrtmgp's pade_eval: numerator degree m, denominator degree n with
 * constant term 1, coefficient k of band b and size regime r at
 * c[b + nb (r + 3 k)] (the Fortran (nband, nsizereg, ncoeff) array).
 * PARITY UNPINNED: the rte-rrtmgp source is absent from the reference tree;
 * this follows the published algorithm (mo_cloud_optics compute_all_from_pade). */
static double pade_eval(const double *c, int nb, int b, int r, int m, int n, double re)
{
#define PC(k) c[(size_t)b + (size_t)nb * ((size_t)r + 3 * (size_t)(k))]
    double denom = PC(n + m);
    for (int i = n - 1 + m; i >= 1 + m; --i) denom = PC(i) + re * denom;
    denom = 1.0 + re * denom;
    double numer = PC(m);
    for (int i = m - 1; i >= 0; --i) numer = PC(i) + re * numer;
#undef PC
    return numer / denom;
}

/* The size regime (0-based) of a radius: min(floor((re - b2) / b3) + 2, 3),
 * rte's expression for three regimes (four bounds). */
static int pade_regime(const double *bnd, double re)
{
    int ir = (int)floor((re - bnd[1]) / bnd[2]) + 2;
    if (ir > 3) ir = 3;
    if (ir < 1) ir = 1;
    return ir - 1;
}

/* compute_all_from_pade for one phase and band: t = path ext,
 * ts = t (1 - max(0, co-albedo)), tsg = ts g. */
static void pade_phase(const double *ce, const double *cs, const double *ca, const double *be,
                       const double *bs, const double *ba, int nb, int b, double path, double re,
                       double *t, double *ts, double *tsg)
{
    const double tt = path * pade_eval(ce, nb, b, pade_regime(be, re), 2, 3, re);
    const double s = tt * (1.0 - fmax(0.0, pade_eval(cs, nb, b, pade_regime(bs, re), 2, 2, re)));
    *tsg = s * pade_eval(ca, nb, b, pade_regime(ba, re), 2, 2, re);
    *ts = s;
    *t = tt;
}

/* Returns nonzero if an effective radius of a masked element is out of range. */
int orc_cloud_optics(const ccpprad_cloud_lut *lut, int ncol, int nlay, const double *lwp,
                     const double *iwp, const double *reliq, const double *reice,
                     double *tau, double *ssa, double *g)
{
    const int nb = lut->nband;
    int icergh = lut->icergh > 0 ? lut->icergh : 1;
    for (int l = 0; l < nlay; ++l)
        for (int c = 0; c < ncol; ++c) {
            double cl = F2(lwp, c, l, ncol), ci = F2(iwp, c, l, ncol);
            double rl = F2(reliq, c, l, ncol), ri = F2(reice, c, l, ncol);
            if (cl > 0.0 && (rl < lut->radliq_lwr || rl > lut->radliq_upr)) return 1;
            if (ci > 0.0 && (ri < lut->radice_lwr || ri > lut->radice_upr)) return 2;
        }
    for (int b = 0; b < nb; ++b)
        for (int l = 0; l < nlay; ++l)
            for (int c = 0; c < ncol; ++c) {
                double lt = 0, lts = 0, ltsg = 0, it = 0, its = 0, itsg = 0;
                double cl = F2(lwp, c, l, ncol), ci = F2(iwp, c, l, ncol);
                if (lut->pade) {
                    const size_t io = (size_t)nb * 3 * 6 * (icergh - 1), is = (size_t)nb * 3 * 5 * (icergh - 1);
                    if (cl > 0.0)
                        pade_phase(lut->pade_extliq, lut->pade_ssaliq, lut->pade_asyliq, lut->pade_sizreg_extliq,
                                   lut->pade_sizreg_ssaliq, lut->pade_sizreg_asyliq, nb, b, cl,
                                   F2(reliq, c, l, ncol), &lt, &lts, &ltsg);
                    if (ci > 0.0)
                        pade_phase(lut->pade_extice + io, lut->pade_ssaice + is, lut->pade_asyice + is,
                                   lut->pade_sizreg_extice, lut->pade_sizreg_ssaice, lut->pade_sizreg_asyice, nb,
                                   b, ci, F2(reice, c, l, ncol), &it, &its, &itsg);
                } else {
                if (cl > 0.0)
                    lut_interp(lut->extliq + (size_t)lut->nsize_liq * b, lut->ssaliq + (size_t)lut->nsize_liq * b,
                               lut->asyliq + (size_t)lut->nsize_liq * b, lut->nsize_liq, lut->radliq_lwr,
                               lut->radliq_upr, cl, F2(reliq, c, l, ncol), &lt, &lts, &ltsg);
                if (ci > 0.0) {
                    size_t o = (size_t)lut->nsize_ice * (b + (size_t)nb * (icergh - 1));
                    lut_interp(lut->extice + o, lut->ssaice + o, lut->asyice + o, lut->nsize_ice,
                               lut->radice_lwr, lut->radice_upr, ci, F2(reice, c, l, ncol), &it, &its, &itsg);
                }
                }
                double t = lt + it, ts = lts + its, tsg = ltsg + itsg;
                F3(tau, c, l, b, ncol, nlay) = t;
                F3(ssa, c, l, b, ncol, nlay) = ts / fmax(DBL_EPSILON, t);
                F3(g, c, l, b, ncol, nlay) = tsg / fmax(DBL_EPSILON, ts);
            }
    return 0;
}

/* increment_2stream_by_2stream: (t2,s2,g2) += (t1,s1,g1), eps = 3*tiny. */
static inline void inc2s(double t1, double s1, double g1, double *t2, double *s2, double *g2)
{
    const double eps = 3.0 * DBL_MIN;
    double t12 = t1 + *t2;
    double ts12 = t1 * s1 + *t2 * *s2;
    *g2 = (t1 * s1 * g1 + *t2 * *s2 * *g2) / fmax(eps, ts12);
    *s2 = ts12 / fmax(eps, t12);
    *t2 = t12;
}

/* inc2s on the per-g-point optics, in the working precision. */
static inline void inc2s_r(real t1, real s1, real g1, real *t2, real *s2, real *g2)
{
    const real eps = R(3.0) * R_MIN;
    real t12 = t1 + *t2;
    real ts12 = t1 * s1 + *t2 * *s2;
    *g2 = (t1 * s1 * g1 + *t2 * *s2 * *g2) / fmax(eps, ts12);
    *s2 = ts12 / fmax(eps, t12);
    *t2 = t12;
}

/* ===================================================================== */
/* adding (Shonk & Hogan 2008, rte's adding()), shared by the SW and the  */
/* LW two-stream solvers.  Per array layer l: diffuse reflectance Rd,     */
/* transmittance Td and the layer sources sup (upward, leaving its top)   */
/* and sdn (downward, leaving its bottom); surface albedo alb_sfc and     */
/* source src_sfc.  No diffuse flux enters at the top.  up/dn at the      */
/* nlay+1 levels; alb/src (nlev) and den (nlay) are scratch.              */
/* ===================================================================== */
static void adding(int nlay, int top_at_1, real alb_sfc, real src_sfc, const real *Rd,
                   const real *Td, const real *sup, const real *sdn, real *up, real *dn,
                   real *alb, real *src, real *den)
{
    const int nlev = nlay + 1, vs = top_at_1 ? nlay : 0;
    alb[vs] = alb_sfc;
    src[vs] = src_sfc;
    if (top_at_1) {
        for (int l = nlay - 1; l >= 0; --l) {
            den[l] = R(1.0) / (R(1.0) - Rd[l] * alb[l + 1]);
            alb[l] = Rd[l] + Td[l] * Td[l] * alb[l + 1] * den[l];
            src[l] = sup[l] + Td[l] * den[l] * (src[l + 1] + alb[l + 1] * sdn[l]);
        }
        dn[0] = R(0.0);
        up[0] = dn[0] * alb[0] + src[0];
        for (int v = 1; v < nlev; ++v) {
            dn[v] = (Td[v - 1] * dn[v - 1] + Rd[v - 1] * src[v] + sdn[v - 1]) * den[v - 1];
            up[v] = dn[v] * alb[v] + src[v];
        }
    } else {
        for (int l = 0; l < nlay; ++l) {
            den[l] = R(1.0) / (R(1.0) - Rd[l] * alb[l]);
            alb[l + 1] = Rd[l] + Td[l] * Td[l] * alb[l] * den[l];
            src[l + 1] = sup[l] + Td[l] * den[l] * (src[l] + alb[l] * sdn[l]);
        }
        dn[nlay] = R(0.0);
        up[nlay] = dn[nlay] * alb[nlay] + src[nlay];
        for (int v = nlay - 1; v >= 0; --v) {
            dn[v] = (Td[v] * dn[v + 1] + Rd[v] * src[v] + sdn[v]) * den[v];
            up[v] = dn[v] * alb[v] + src[v];
        }
    }
}

/* ===================================================================== */
/* LW two-stream solver with scattering: rte_lw with two-stream optical   */
/* properties (the doGP_lwscat branch, rrtmgp_lw_main.F90:552-575) runs   */
/* lw_solver_2stream = lw_combine_sources + lw_two_stream (Fu et al. 1997 */
/* coefficients, diffusivity secant 1.66) + lw_source_2str (Toon et al.  */
/* 1989 linear-in-tau source) + adding, per g-point.                      */
/* ===================================================================== */
/* One g-point, level sources given: tau/ssa/g (nlay) and levs (nlev) by
 * array index, stride 1; sfc the surface Planck source (pfrac * B(T_sfc));
 * up/dn (nlev) out.  ws: scratch of 7 nlev doubles.  Pinned by
 * tests/test_oracle_lw.py against the two-stream equations integrated by
 * matrix exponentials, layer by layer, with the boundary conditions solved
 * as one linear system (no adding), and against lw_solver_noscat at one
 * angle (1.66, 0.5) when ssa = 0. */
static void lw2s_core(int nlay, int top_at_1, const real *tau, const real *ssa, const real *gg,
                      const real *levs, real emis, real sfc, real *up, real *dn, real *ws)
{
    const real D = R(1.66);
    const int nlev = nlay + 1;
    real *Rd = ws, *Td = ws + nlev, *sup = ws + 2 * nlev, *sdn = ws + 3 * nlev;
    real *alb = ws + 4 * nlev, *src = ws + 5 * nlev, *den = ws + 6 * nlev;
    for (int l = 0; l < nlay; ++l) {
        const real t = tau[l], w = ssa[l], g = gg[l];
        /* lw_two_stream */
        const real g1 = D * (R(1.0) - R(0.5) * w * (R(1.0) + g));
        const real g2 = D * R(0.5) * w * (R(1.0) - g);
        const real kk = sqrt(fmax((g1 - g2) * (g1 + g2), R(1.0e-12)));
        const real e = exp(-t * kk), e2 = e * e;
        const real RT = R(1.0) / (kk * (R(1.0) + e2) + g1 * (R(1.0) - e2));
        Rd[l] = RT * g2 * (R(1.0) - e2);
        Td[l] = RT * R(2.0) * kk * e;
        /* lw_source_2str */
        if (t > R(1.0e-8)) {
            const real st = top_at_1 ? levs[l] : levs[l + 1];
            const real sb = top_at_1 ? levs[l + 1] : levs[l];
            const real Z = (sb - st) / (t * (g1 + g2));
            const real zut = Z + st, zub = Z + sb, zdt = -Z + st, zdb = -Z + sb;
            sup[l] = R(M_PI) * (zut - Rd[l] * zdt - Td[l] * zub);
            sdn[l] = R(M_PI) * (zdb - Rd[l] * zub - Td[l] * zdt);
        } else {
            sup[l] = R(0.0);
            sdn[l] = R(0.0);
        }
    }
    adding(nlay, top_at_1, R(1.0) - emis, R(M_PI) * emis * sfc, Rd, Td, sup, sdn, up, dn, alb, src, den);
}

void orc_lw2s_gpt(int nlay, int top_at_1, const real *tau, const real *ssa, const real *g,
                  const real *levs, real emis, real sfc, real *up, real *dn)
{
    real *ws = malloc(sizeof(real) * 7 * (size_t)(nlay + 1));
    lw2s_core(nlay, top_at_1, tau, ssa, g, levs, emis, sfc, up, dn, ws);
    free(ws);
}

/* lw_combine_sources: the level source at array level v is
 * sqrt(lev_dec(v) * lev_inc(v-1)) between two layers, lev_dec(0) at the
 * first level and lev_inc(nlay-1) at the last (inc/dec strided by ngpt). */
static void lw_combine_sources(int nlay, int ngpt, const real *inc, const real *dec, real *levs)
{
    levs[0] = dec[0];
    for (int v = 1; v < nlay; ++v) levs[v] = sqrt(dec[(size_t)v * ngpt] * inc[(size_t)(v - 1) * ngpt]);
    levs[nlay] = inc[(size_t)(nlay - 1) * ngpt];
}

/* lw_solver_2stream over ncol columns with the band / broadband reduction of
 * orc_lw_solve.  tau/ssa/g/lev_inc/lev_dec (ngpt, nlay, ncol), emis_b
 * (nband, ncol), sfc_src (ngpt, ncol); flux_up/flux_dn (ncol, nlev). */
void orc_lw_solve_2str(const ccpprad_kdist *k, int ncol, int nlay, int top_at_1, const real *tau,
                       const real *ssa, const real *g, const real *lev_inc, const real *lev_dec,
                       const double *emis_b, const real *sfc_src, double *flux_up, double *flux_dn)
{
    const int ng = k->ngpt, nb = k->nband, nlev = nlay + 1;
#pragma omp parallel
    {
        real *gup = malloc(sizeof(real) * nlev * ng), *gdn = malloc(sizeof(real) * nlev * ng);
        real *ws = malloc(sizeof(real) * 7 * nlev), *levs = malloc(sizeof(real) * nlev);
        real *t1 = malloc(sizeof(real) * nlay), *s1 = malloc(sizeof(real) * nlay);
        real *g1 = malloc(sizeof(real) * nlay);
#pragma omp for schedule(static)
        for (int c = 0; c < ncol; ++c) {
            for (int b = 0; b < nb; ++b) {
                real emis = R(F2(emis_b, b, c, nb));
                for (int gp = F2(k->band_lims_gpt, 0, b, 2) - 1; gp < F2(k->band_lims_gpt, 1, b, 2); ++gp) {
                    size_t o = gp + (size_t)ng * (size_t)nlay * c;
                    for (int l = 0; l < nlay; ++l) {
                        t1[l] = tau[o + (size_t)l * ng];
                        s1[l] = ssa[o + (size_t)l * ng];
                        g1[l] = g[o + (size_t)l * ng];
                    }
                    lw_combine_sources(nlay, ng, lev_inc + o, lev_dec + o, levs);
                    lw2s_core(nlay, top_at_1, t1, s1, g1, levs, emis, sfc_src[gp + (size_t)ng * c],
                              gup + (size_t)gp * nlev, gdn + (size_t)gp * nlev, ws);
                }
            }
            for (int v = 0; v < nlev; ++v) {
                double su_ = 0.0, sd_ = 0.0;
                for (int b = 0; b < nb; ++b) {
                    int g0 = F2(k->band_lims_gpt, 0, b, 2) - 1, gl = F2(k->band_lims_gpt, 1, b, 2);
                    double bu = (double)gup[(size_t)g0 * nlev + v], bd = (double)gdn[(size_t)g0 * nlev + v];
                    for (int gp = g0 + 1; gp < gl; ++gp) {
                        bu += (double)gup[(size_t)gp * nlev + v];
                        bd += (double)gdn[(size_t)gp * nlev + v];
                    }
                    su_ += bu; sd_ += bd;
                }
                F2(flux_up, c, v, ncol) = su_;
                F2(flux_dn, c, v, ncol) = sd_;
            }
        }
        free(gup); free(gdn); free(ws); free(levs); free(t1); free(s1); free(g1);
    }
}

/* ===================================================================== */
/* SW solver (sw_two_stream + sw_source_2str + adding), one column, one g */
/* ===================================================================== */
static void sw_one(int nlay, int ngpt, int top_at_1, real mu0, real toa, real adir,
                   real adif, const real *tau, const real *ssa, const real *gg,
                   real *up, real *dn, real *dir, real *Rd, real *Td, real *sup,
                   real *sdn, real *alb, real *src, real *den)
{
    /* two-stream coefficients (Meador & Weaver 1980, Zdunkowski PIFM) */
    real Rdir_l, Tdir_l, T0;
    real fdir_top = toa * mu0;
    /* surface-first (top_at_1 false): layer l spans levels l (bottom) .. l+1 (top) */
    int nlev = nlay + 1;
    real *fd = dir; /* direct flux at levels */
    if (top_at_1) fd[0] = fdir_top; else fd[nlay] = fdir_top;
    for (int ii = 0; ii < nlay; ++ii) {
        int l = top_at_1 ? ii : nlay - 1 - ii;
        real t = tau[(size_t)l * ngpt], w = ssa[(size_t)l * ngpt], g = gg[(size_t)l * ngpt];
        real mu0_inv = R(1.0) / mu0;
        real g1 = (R(8.0) - w * (R(5.0) + R(3.0) * g)) * R(0.25);
        real g2 = R(3.0) * (w * (R(1.0) - g)) * R(0.25);
        real g3 = (R(2.0) - R(3.0) * mu0 * g) * R(0.25);
        real g4 = R(1.0) - g3;
        real a1 = g1 * g4 + g2 * g3;
        real a2 = g1 * g3 + g2 * g4;
        real kk = sqrt(fmax((g1 - g2) * (g1 + g2), R(1.0e-12)));
        real e = exp(-t * kk);
        real e2 = e * e;
        real RT = R(1.0) / (kk * (R(1.0) + e2) + g1 * (R(1.0) - e2));
        Rd[l] = RT * g2 * (R(1.0) - e2);
        Td[l] = RT * R(2.0) * kk * e;
        T0 = exp(-t * mu0_inv);
        real kmu = kk * mu0, kg3 = kk * g3, kg4 = kk * g4;
        real d = R(1.0) - kmu * kmu;
        real RT2 = w * RT / (fabs(d) >= R_EPS ? d : R_EPS);
        Rdir_l = RT2 * ((R(1.0) - kmu) * (a2 + kg3) - (R(1.0) + kmu) * (a2 - kg3) * e2 -
                        R(2.0) * (kg3 - a2 * kmu) * e * T0);
        Tdir_l = -RT2 * ((R(1.0) + kmu) * (a1 + kg4) * T0 - (R(1.0) - kmu) * (a1 - kg4) * e2 * T0 -
                         R(2.0) * (kg4 + a1 * kmu) * e);
        Rdir_l = fmax(R(0.0), fmin(Rdir_l, R(1.0) - T0));
        Tdir_l = fmax(R(0.0), fmin(Tdir_l, R(1.0) - T0 - Rdir_l));
        /* sw_source_2str */
        int vin = top_at_1 ? l : l + 1, vout = top_at_1 ? l + 1 : l;
        sup[l] = Rdir_l * fd[vin];
        sdn[l] = Tdir_l * fd[vin];
        fd[vout] = T0 * fd[vin];
    }
    int vs = top_at_1 ? nlay : 0;
    adding(nlay, top_at_1, adif, fd[vs] * adir, Rd, Td, sup, sdn, up, dn, alb, src, den);
    for (int v = 0; v < nlev; ++v) dn[v] += fd[v];
}

/* rte_sw for ncol columns.  tau/ssa/g (ngpt, nlay, ncol); toa (ngpt, ncol)
 * already normalised; alb_dir/dif (nband, ncol).  Band fluxes out
 * (nband, nlev, ncol) for up, dn, dn_dir. */
void orc_sw_solve(const ccpprad_kdist *k, int ncol, int nlay, int top_at_1, const double *mu0,
                  const real *tau, const real *ssa, const real *g, const real *toa,
                  const double *alb_dir, const double *alb_dif, double *bup, double *bdn,
                  double *bdir)
{
    const int ng = k->ngpt, nb = k->nband, nlev = nlay + 1;
#pragma omp parallel
    {
        real *w = malloc(sizeof(real) * 11 * (nlev + 1));
        real *up = w, *dn = w + nlev + 1, *dir = w + 2 * (nlev + 1), *Rd = w + 3 * (nlev + 1);
        real *Td = w + 4 * (nlev + 1), *sup = w + 5 * (nlev + 1), *sdn = w + 6 * (nlev + 1);
        real *alb = w + 7 * (nlev + 1), *src = w + 8 * (nlev + 1), *den = w + 9 * (nlev + 1);
#pragma omp for schedule(static)
        for (int c = 0; c < ncol; ++c) {
            for (size_t i = 0; i < (size_t)nb * nlev; ++i) {
                bup[i + (size_t)nb * nlev * c] = 0.0;
                bdn[i + (size_t)nb * nlev * c] = 0.0;
                bdir[i + (size_t)nb * nlev * c] = 0.0;
            }
            for (int b = 0; b < nb; ++b) {
                int g0 = F2(k->band_lims_gpt, 0, b, 2) - 1, g1 = F2(k->band_lims_gpt, 1, b, 2);
                for (int gp = g0; gp < g1; ++gp) {
                    size_t o = gp + (size_t)ng * (size_t)nlay * c;
                    sw_one(nlay, ng, top_at_1, R(mu0[c]), toa[gp + (size_t)ng * c],
                           R(F2(alb_dir, b, c, nb)), R(F2(alb_dif, b, c, nb)), tau + o, ssa + o, g + o,
                           up, dn, dir, Rd, Td, sup, sdn, alb, src, den);
                    for (int v = 0; v < nlev; ++v) {
                        size_t q = b + (size_t)nb * (v + (size_t)nlev * c);
                        if (gp == g0) { bup[q] = (double)up[v]; bdn[q] = (double)dn[v]; bdir[q] = (double)dir[v]; }
                        else { bup[q] += (double)up[v]; bdn[q] += (double)dn[v]; bdir[q] += (double)dir[v]; }
                    }
                }
            }
        }
        free(w);
    }
}

/* ===================================================================== */
/* Scheme layer: rrtmgp_lw_main_run                                       */
/* ===================================================================== */
static const double EPS_PRE = 1.0e-6, ONEMINUS = 1.0 - 1.0e-6, FTINY = 1.0e-12;

static void gather_vmr(int ncol_all, int nlay, const int *cols, int nc, const double *const vm[6],
                       double *vmr /*(nc, nlay, 6)*/)
{
    for (int g = 0; g < 6; ++g)
        for (int l = 0; l < nlay; ++l)
            for (int i = 0; i < nc; ++i)
                F3(vmr, i, l, g, nc, nlay) = F2(vm[g], cols[i], l, ncol_all);
}

static void gather2(int ncol_all, int nlay, const int *cols, int nc, const double *src, double *dst)
{
    for (int l = 0; l < nlay; ++l)
        for (int i = 0; i < nc; ++i) F2(dst, i, l, nc) = F2(src, cols[i], l, ncol_all);
}

/* Returns 0, or nonzero with errmsg.  Host arrays.  Processes blocks of
 * rrtmgp_phys_blksz columns exactly as the reference loop does. */
int orc_lw_main_run(const ccpprad_kdist *k, const ccpprad_cloud_lut *lut, const ccpprad_lw_args *a,
                    char *errmsg, int errlen)
{
    if (!a->doLWrad) return 0;
    const int nC = a->nCol, nL = a->nLay, nV = nL + 1, bs = a->rrtmgp_phys_blksz;
    const int ng = k->ngpt, nb = k->nband;
    if (a->doGP_lwscat && a->use_LW_jacobian) {
        snprintf(errmsg, errlen, "doGP_lwscat: rte_lw gives no surface-temperature Jacobian with two-stream optics");
        return 1;
    }
    if (bs <= 0 || nC % bs || bs > 4096) { snprintf(errmsg, errlen, "nCol must be a multiple of rrtmgp_phys_blksz (<= 4096)"); return 1; }
    const double *vm[6] = {a->vmr_h2o, a->vmr_co2, a->vmr_o3, a->vmr_n2o, a->vmr_ch4, a->vmr_o2};
    size_t cl = (size_t)bs * nL, cg = cl * ng;
    double *vmr = malloc(sizeof(double) * cl * 6), *pl = malloc(sizeof(double) * cl);
    double *tl = malloc(sizeof(double) * cl), *pv = malloc(sizeof(double) * bs * nV);
    double *tv = malloc(sizeof(double) * bs * nV);
    real *tau = malloc(sizeof(real) * cg), *pf = malloc(sizeof(real) * cg);
    real *lay = malloc(sizeof(real) * cg), *inc = malloc(sizeof(real) * cg);
    real *dec = malloc(sizeof(real) * cg), *ss = malloc(sizeof(real) * bs * ng);
    real *sj = malloc(sizeof(real) * bs * ng);
    double *emis = malloc(sizeof(double) * bs * nb);
    double *ctau = malloc(sizeof(double) * cl * nb), *cssa = malloc(sizeof(double) * cl * nb);
    double *cg_ = malloc(sizeof(double) * cl * nb), *tmp = malloc(sizeof(double) * cl * nb * 3);
    real *Ds = malloc(sizeof(real) * bs * ng), *tau_all = malloc(sizeof(real) * cg);
    double *fu = malloc(sizeof(double) * bs * nV), *fd = malloc(sizeof(double) * bs * nV);
    double *fj = malloc(sizeof(double) * bs * nV), *cf = malloc(sizeof(double) * cl);
    double *ov = malloc(sizeof(double) * cl);
    unsigned char *mask = malloc((size_t)cg);
    int *cols = malloc(sizeof(int) * bs), *seeds = malloc(sizeof(int) * bs);
    double *l4 = malloc(sizeof(double) * cl * 4);
    int rc = 0;
    for (int c0 = 0; c0 < nC; c0 += bs) {
        for (int i = 0; i < bs; ++i) cols[i] = c0 + i;
        gather_vmr(nC, nL, cols, bs, vm, vmr);
        gather2(nC, nL, cols, bs, a->p_lay, pl);
        gather2(nC, nL, cols, bs, a->t_lay, tl);
        gather2(nC, nV, cols, bs, a->p_lev, pv);
        gather2(nC, nV, cols, bs, a->t_lev, tv);
        gather2(nC, nL, cols, bs, a->cld_frac, cf);
        for (int i = 0; i < bs; ++i) {
            double s = a->semis[c0 + i];
            for (int b = 0; b < nb; ++b) F2(emis, b, i, nb) = (s > EPS_PRE && s <= 1.0) ? s : 1.0;
        }
        /* range checks of gas_optics */
        for (size_t i = 0; i < cl; ++i) {
            if (pl[i] < k->press_ref[k->npres - 1] || pl[i] > k->press_ref[0] ||
                tl[i] < k->temp_ref[0] || tl[i] > k->temp_ref[k->ntemp - 1]) {
                snprintf(errmsg, errlen, "gas_optics(): p_lay or t_lay outside the table range");
                rc = 1; goto done;
            }
        }
        orc_gas_optics(k, bs, nL, pl, pv, tl, vmr, tau, pf, NULL);
        orc_lw_sources(k, bs, nL, a->top_at_1, tl, tv, a->tsfg + c0, pf, lay, inc, dec, ss, sj);
        /* clear/cloudy indicator */
        int anycld = 0;
        double zcf1[4096];
        for (int i = 0; i < bs; ++i) {
            double z0 = 1.0;
            for (int l = 0; l < nL; ++l) z0 = fmin(z0, 1.0 - F2(cf, i, l, bs));
            if (z0 <= FTINY) z0 = 0.0;
            if (z0 > ONEMINUS) z0 = 1.0;
            zcf1[i] = 1.0 - z0;
            if (zcf1[i] > EPS_PRE) anycld = 1;
        }
        memset(ctau, 0, sizeof(double) * cl * nb);
        memset(cssa, 0, sizeof(double) * cl * nb);
        memset(cg_, 0, sizeof(double) * cl * nb);
        if (anycld) {
            double *lw_ = l4, *iw_ = l4 + cl, *rl_ = l4 + 2 * cl, *ri_ = l4 + 3 * cl;
            gather2(nC, nL, cols, bs, a->cld_lwp, lw_);
            gather2(nC, nL, cols, bs, a->cld_iwp, iw_);
            gather2(nC, nL, cols, bs, a->cld_reliq, rl_);
            gather2(nC, nL, cols, bs, a->cld_reice, ri_);
            if (orc_cloud_optics(lut, bs, nL, lw_, iw_, rl_, ri_, ctau, cssa, cg_)) {
                snprintf(errmsg, errlen, "cloud_optics(): effective radius outside LUT range");
                rc = 1; goto done;
            }
            const double *cnv[2][4] = {{a->cld_cnv_lwp, a->cld_cnv_iwp, a->cld_cnv_reliq, a->cld_cnv_reice},
                                       {a->cld_pbl_lwp, a->cld_pbl_iwp, a->cld_pbl_reliq, a->cld_pbl_reice}};
            int on[2] = {a->doGP_sgs_cnv, a->doGP_sgs_pbl};
            for (int s = 0; s < 2; ++s) {
                if (!on[s]) continue;
                gather2(nC, nL, cols, bs, cnv[s][0], lw_);
                gather2(nC, nL, cols, bs, cnv[s][1], iw_);
                gather2(nC, nL, cols, bs, cnv[s][2], rl_);
                gather2(nC, nL, cols, bs, cnv[s][3], ri_);
                double *t2 = tmp, *s2 = tmp + cl * nb, *g2 = tmp + 2 * cl * nb;
                if (orc_cloud_optics(lut, bs, nL, lw_, iw_, rl_, ri_, t2, s2, g2)) {
                    snprintf(errmsg, errlen, "cloud_optics(): effective radius outside LUT range");
                    rc = 1; goto done;
                }
                for (size_t i = 0; i < cl * nb; ++i) inc2s(t2[i], s2[i], g2[i], &ctau[i], &cssa[i], &cg_[i]);
            }
        }
        /* precipitation optics: rain + snow tau added to all bands (ssa=g=0) */
        {
            const double absrain = (double)0.33e-3f, abssnow0 = 1.5;
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l) {
                    double tp = 0.0;
                    int set = 0;
                    int c = c0 + i;
                    if (F2(a->cld_frac, c, l, nC) > EPS_PRE) {
                        double tr = absrain * F2(a->cld_rwp, c, l, nC);
                        double ts = 0.0;
                        if (F2(a->cld_swp, c, l, nC) > 0.0 && F2(a->cld_resnow, c, l, nC) > 10.0)
                            ts = abssnow0 * (double)1.05756f * F2(a->cld_swp, c, l, nC) / F2(a->cld_resnow, c, l, nC);
                        tp = tr + ts;
                        set = 1;
                    }
                    for (int b = 0; b < nb; ++b) {
                        size_t q = i + (size_t)bs * (l + (size_t)nL * b);
                        inc2s(set ? tp : 0.0, 0.0, 0.0, &ctau[q], &cssa[q], &cg_[q]);
                    }
                }
        }
        /* McICA */
        memset(mask, 0, (size_t)cg);
        if (anycld) {
            for (int i = 0; i < bs; ++i)
                seeds[i] = (a->isubc_lw == 1) ? ng + c0 + i + 1 : (a->isubc_lw == 2 ? a->icseed_lw[c0 + i] : 0);
            if (a->cloud_overlap_param) gather2(nC, nL, cols, bs, a->cloud_overlap_param, ov);
            orc_mcica_mask(bs, nL, ng, seeds, cf, a->cloud_overlap_param ? ov : NULL, a->iovr, a->iovr_max,
                           a->iovr_maxrand, a->iovr_rand, a->iovr_dcorr, a->iovr_exp, a->iovr_exprand, mask);
        }
        /* aerosol increment into clear sky (1scl += 1scl by band) */
        if (a->aerlw_tau) {
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l)
                    for (int b = 0; b < nb; ++b) {
                        double at = F3(a->aerlw_tau, c0 + i, l, b, nC, nL);
                        for (int gp = F2(k->band_lims_gpt, 0, b, 2) - 1; gp < F2(k->band_lims_gpt, 1, b, 2); ++gp)
                            tau[gp + (size_t)ng * (l + (size_t)nL * i)] += R(at);
                    }
        }
        /* clear sky */
        if (a->doLWclrsky) {
            if (a->nGauss_angles > 1) {
                orc_lw_solve(k, bs, nL, a->top_at_1, a->nGauss_angles, NULL, tau, lay, inc, dec, emis, ss, sj, fu, fd, NULL);
            } else {
                orc_optimal_angles(k, bs, nL, tau, Ds);
                orc_lw_solve(k, bs, nL, a->top_at_1, 1, Ds, tau, lay, inc, dec, emis, ss, sj, fu, fd, NULL);
            }
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    F2(a->fluxlwUP_clrsky, c0 + i, v, nC) = F2(fu, i, v, bs);
                    F2(a->fluxlwDOWN_clrsky, c0 + i, v, nC) = F2(fd, i, v, bs);
                }
        } else {
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    F2(a->fluxlwUP_clrsky, c0 + i, v, nC) = 0.0;
                    F2(a->fluxlwDOWN_clrsky, c0 + i, v, nC) = 0.0;
                }
        }
        if (a->doGP_lwscat) {
            /* all sky with scattering: the clear sky (1scl) incremented into
             * the McICA-sampled clouds (2str), increment_2stream_by_1scalar:
             * tau = tau_c + tau, ssa = tau_c ssa_c / tau (0 unless tau > eps),
             * g = g_c; then lw_solver_2stream */
            const real eps = R(3.0) * R_MIN;
            real *ssa_all = lay, *g_all = pf;   /* pfrac and the layer source are no longer needed */
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l)
                    for (int b = 0; b < nb; ++b) {
                        size_t q = i + (size_t)bs * (l + (size_t)nL * b);
                        for (int gp = F2(k->band_lims_gpt, 0, b, 2) - 1; gp < F2(k->band_lims_gpt, 1, b, 2); ++gp) {
                            size_t o = gp + (size_t)ng * (l + (size_t)nL * i);
                            int m = mask[l + (size_t)nL * gp + (size_t)nL * ng * i];
                            real t1 = m ? R(ctau[q]) : R(0.0), s1 = m ? R(cssa[q]) : R(0.0), g1 = m ? R(cg_[q]) : R(0.0);
                            real t12 = t1 + tau[o];
                            ssa_all[o] = t12 > eps ? t1 * s1 / t12 : R(0.0);
                            g_all[o] = g1;
                            tau_all[o] = t12;
                        }
                    }
            orc_lw_solve_2str(k, bs, nL, a->top_at_1, tau_all, ssa_all, g_all, inc, dec, emis, ss, fu, fd);
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    int c = c0 + i;
                    F2(a->fluxlwUP_allsky, c, v, nC) = F2(fu, i, v, bs);
                    F2(a->fluxlwDOWN_allsky, c, v, nC) = F2(fd, i, v, bs);
                    if (a->fluxlwUP_radtime) F2(a->fluxlwUP_radtime, c, v, nC) = F2(fu, i, v, bs);
                    if (a->fluxlwDOWN_radtime) F2(a->fluxlwDOWN_radtime, c, v, nC) = F2(fd, i, v, bs);
                }
            continue;
        }
        /* all sky: clouds (by g-point via McICA) incremented into clear sky
         * as absorption only: tau += tau_c * (1 - ssa_c) */
        for (int i = 0; i < bs; ++i)
            for (int l = 0; l < nL; ++l)
                for (int b = 0; b < nb; ++b) {
                    size_t q = i + (size_t)bs * (l + (size_t)nL * b);
                    for (int gp = F2(k->band_lims_gpt, 0, b, 2) - 1; gp < F2(k->band_lims_gpt, 1, b, 2); ++gp) {
                        size_t o = gp + (size_t)ng * (l + (size_t)nL * i);
                        int m = mask[l + (size_t)nL * gp + (size_t)nL * ng * i];
                        real t1 = m ? R(ctau[q]) : R(0.0), s1 = m ? R(cssa[q]) : R(0.0);
                        tau_all[o] = tau[o] + t1 * (R(1.0) - s1);
                    }
                }
        orc_lw_solve(k, bs, nL, a->top_at_1, a->nGauss_angles, NULL, tau_all, lay, inc, dec, emis, ss, sj, fu, fd,
                     a->use_LW_jacobian ? fj : NULL);
        for (int v = 0; v < nV; ++v)
            for (int i = 0; i < bs; ++i) {
                int c = c0 + i;
                F2(a->fluxlwUP_allsky, c, v, nC) = F2(fu, i, v, bs);
                F2(a->fluxlwDOWN_allsky, c, v, nC) = F2(fd, i, v, bs);
                if (a->fluxlwUP_radtime) F2(a->fluxlwUP_radtime, c, v, nC) = F2(fu, i, v, bs);
                if (a->fluxlwDOWN_radtime) F2(a->fluxlwDOWN_radtime, c, v, nC) = F2(fd, i, v, bs);
                if (a->use_LW_jacobian && a->fluxlwUP_jac) F2(a->fluxlwUP_jac, c, v, nC) = F2(fj, i, v, bs);
            }
    }
    if (a->htrlw) orc_heating_rate(nC, nL, a->fluxlwUP_allsky, a->fluxlwDOWN_allsky, a->p_lev, a->htrlw);
    if (a->htrlwc && a->doLWclrsky)
        orc_heating_rate(nC, nL, a->fluxlwUP_clrsky, a->fluxlwDOWN_clrsky, a->p_lev, a->htrlwc);
done:
    free(vmr); free(pl); free(tl); free(pv); free(tv); free(tau); free(pf); free(lay); free(inc);
    free(dec); free(ss); free(sj); free(emis); free(ctau); free(cssa); free(cg_); free(tmp); free(Ds);
    free(tau_all); free(fu); free(fd); free(fj); free(cf); free(ov); free(mask); free(cols); free(seeds);
    free(l4);
    return rc;
}

/* ===================================================================== */
/* Scheme layer: rrtmgp_sw_main_run                                       */
/* ===================================================================== */
int orc_sw_main_run(const ccpprad_kdist *k, const ccpprad_cloud_lut *lut, const ccpprad_sw_args *a,
                    char *errmsg, int errlen)
{
    if (!a->doSWrad) return 0;
    const int nC = a->nCol, nL = a->nLay, nV = nL + 1, bs = a->rrtmgp_phys_blksz;
    const int ng = k->ngpt, nb = k->nband;
    if (a->nDay <= 0) {
        for (size_t i = 0; i < (size_t)nC * nV; ++i) {
            a->fluxswUP_allsky[i] = 0; a->fluxswDOWN_allsky[i] = 0;
            a->fluxswUP_clrsky[i] = 0; a->fluxswDOWN_clrsky[i] = 0;
        }
        for (size_t i = 0; i < (size_t)6 * nC; ++i) a->scmpsw[i] = 0.0;
        return 0;
    }
    if (bs <= 0 || a->nDay % bs || bs > 4096) { snprintf(errmsg, errlen, "nDay must be a multiple of rrtmgp_phys_blksz (<= 4096)"); return 1; }
    const double *vm[6] = {a->vmr_h2o, a->vmr_co2, a->vmr_o3, a->vmr_n2o, a->vmr_ch4, a->vmr_o2};
    size_t cl = (size_t)bs * nL, cg = cl * ng, bvl = (size_t)nb * nV * bs;
    double *vmr = malloc(sizeof(double) * cl * 6), *pl = malloc(sizeof(double) * cl);
    double *tl = malloc(sizeof(double) * cl), *pv = malloc(sizeof(double) * bs * nV);
    double *cf = malloc(sizeof(double) * cl), *ov = malloc(sizeof(double) * cl);
    real *tau = malloc(sizeof(real) * cg), *ray = malloc(sizeof(real) * cg);
    real *ssa = malloc(sizeof(real) * cg), *gg = malloc(sizeof(real) * cg);
    real *toa = malloc(sizeof(real) * bs * ng);
    double *mu0 = malloc(sizeof(double) * bs);
    double *adir = malloc(sizeof(double) * bs * nb), *adif = malloc(sizeof(double) * bs * nb);
    double *ctau = malloc(sizeof(double) * cl * nb), *cssa = malloc(sizeof(double) * cl * nb);
    double *cgg = malloc(sizeof(double) * cl * nb), *tmp = malloc(sizeof(double) * cl * nb * 3);
    double *ptau = malloc(sizeof(double) * cl * nb), *pssa = malloc(sizeof(double) * cl * nb);
    double *pg = malloc(sizeof(double) * cl * nb), *l4 = malloc(sizeof(double) * cl * 4);
    double *cup = calloc(bvl, sizeof(double)), *cdn = calloc(bvl, sizeof(double));
    double *cdir = calloc(bvl, sizeof(double)), *aup = calloc(bvl, sizeof(double));
    double *adn = calloc(bvl, sizeof(double)), *adr = calloc(bvl, sizeof(double));
    unsigned char *mask = malloc((size_t)cg);
    int *cols = malloc(sizeof(int) * bs), *seeds = malloc(sizeof(int) * bs);
    double zcf1[4096];
    int rc = 0;
    /* NIR / UV-VIS split (rrtmgp_sw_main.F90:211-213,368-387) */
    int ibd = 0, ibd_uv = 0;
    for (int b = 0; b < nb; ++b) {
        if (F2(k->band_lims_wvn, 0, b, 2) == 12850.0) ibd = b + 1;
        if (F2(k->band_lims_wvn, 0, b, 2) == 29000.0) ibd_uv = b + 1;
    }
    if (!ibd || !ibd_uv) { snprintf(errmsg, errlen, "SW band limits lack the 12850 / 29000 cm-1 bands"); rc = 1; goto done; }
    for (int c0 = 0; c0 < a->nDay; c0 += bs) {
        for (int i = 0; i < bs; ++i) cols[i] = a->idx[c0 + i] - 1;
        gather2(nC, nL, cols, bs, a->cld_frac, cf);
        int cloudy = 0;
        for (int i = 0; i < bs; ++i) {
            double z0 = 1.0;
            for (int l = 0; l < nL; ++l) z0 = fmin(z0, 1.0 - F2(cf, i, l, bs));
            if (z0 <= FTINY) z0 = 0.0;
            if (z0 > ONEMINUS) z0 = 1.0;
            zcf1[i] = 1.0 - z0;
            if (zcf1[i] > EPS_PRE) cloudy = 1;
        }
        memset(ctau, 0, sizeof(double) * cl * nb); memset(cssa, 0, sizeof(double) * cl * nb);
        memset(cgg, 0, sizeof(double) * cl * nb); memset(ptau, 0, sizeof(double) * cl * nb);
        memset(pssa, 0, sizeof(double) * cl * nb); memset(pg, 0, sizeof(double) * cl * nb);
        /* scmpsw components per block column: [6][bs] */
        double scl[6 * 4096], sal[6 * 4096];
        memset(scl, 0, sizeof(double) * 6 * bs); memset(sal, 0, sizeof(double) * 6 * bs);
        /* cldtausw wiped every block (rrtmgp_sw_main.F90:308) */
        for (size_t i = 0; i < (size_t)nC * nL; ++i) a->cldtausw[i] = 0.0;
        memset(cup, 0, sizeof(double) * bvl); memset(cdn, 0, sizeof(double) * bvl);
        memset(cdir, 0, sizeof(double) * bvl); memset(aup, 0, sizeof(double) * bvl);
        memset(adn, 0, sizeof(double) * bvl); memset(adr, 0, sizeof(double) * bvl);
        gather_vmr(nC, nL, cols, bs, vm, vmr);
        gather2(nC, nL, cols, bs, a->p_lay, pl);
        gather2(nC, nL, cols, bs, a->t_lay, tl);
        gather2(nC, nV, cols, bs, a->p_lev, pv);
        for (size_t i = 0; i < cl; ++i) {
            if (pl[i] < k->press_ref[k->npres - 1] || pl[i] > k->press_ref[0] ||
                tl[i] < k->temp_ref[0] || tl[i] > k->temp_ref[k->ntemp - 1]) {
                snprintf(errmsg, errlen, "gas_optics(): p_lay or t_lay outside the table range");
                rc = 1; goto done;
            }
        }
        orc_gas_optics(k, bs, nL, pl, pv, tl, vmr, tau, NULL, ray);
        /* combine_and_reorder_2str */
        for (size_t i = 0; i < cg; ++i) {
            real t = tau[i] + ray[i];
            if (t > R(2.0) * R_MIN) { tau[i] = t; ssa[i] = ray[i] / t; }
            else { tau[i] = t; ssa[i] = R(0.0); }
            gg[i] = R(0.0);
        }
        for (int i = 0; i < bs; ++i) {
            double s = 0.0;
            for (int gp = 0; gp < ng; ++gp) s += k->solar_src[gp];
            for (int gp = 0; gp < ng; ++gp) toa[gp + (size_t)ng * i] = R(k->solar_src[gp] * a->solcon / s);
            mu0[i] = a->coszen[cols[i]];
            for (int b = 0; b < nb; ++b) {
                double lo = F2(k->band_lims_wvn, 0, b, 2);
                int c = cols[i];
                if (lo < 12850.0) { F2(adir, b, i, nb) = a->sfc_alb_nir_dir[c]; F2(adif, b, i, nb) = a->sfc_alb_nir_dif[c]; }
                if (lo == 12850.0) {
                    F2(adir, b, i, nb) = 0.5 * (a->sfc_alb_nir_dir[c] + a->sfc_alb_uvvis_dir[c]);
                    F2(adif, b, i, nb) = 0.5 * (a->sfc_alb_nir_dif[c] + a->sfc_alb_uvvis_dif[c]);
                }
                if (lo >= 16000.0) { F2(adir, b, i, nb) = a->sfc_alb_uvvis_dir[c]; F2(adif, b, i, nb) = a->sfc_alb_uvvis_dif[c]; }
            }
        }
        memset(mask, 0, (size_t)cg);
        if (cloudy) {
            double *lw_ = l4, *iw_ = l4 + cl, *rl_ = l4 + 2 * cl, *ri_ = l4 + 3 * cl;
            gather2(nC, nL, cols, bs, a->cld_lwp, lw_);
            gather2(nC, nL, cols, bs, a->cld_iwp, iw_);
            gather2(nC, nL, cols, bs, a->cld_reliq, rl_);
            gather2(nC, nL, cols, bs, a->cld_reice, ri_);
            if (orc_cloud_optics(lut, bs, nL, lw_, iw_, rl_, ri_, ctau, cssa, cgg)) {
                snprintf(errmsg, errlen, "cloud_optics(): effective radius outside LUT range");
                rc = 1; goto done;
            }
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l)
                    F2(a->cldtausw, cols[i], l, nC) = F3(ctau, i, l, 10, bs, nL);
            const double *cnv[2][4] = {{a->cld_cnv_lwp, a->cld_cnv_iwp, a->cld_cnv_reliq, a->cld_cnv_reice},
                                       {a->cld_pbl_lwp, a->cld_pbl_iwp, a->cld_pbl_reliq, a->cld_pbl_reice}};
            int on[2] = {a->doGP_sgs_cnv, a->doGP_sgs_pbl};
            for (int s = 0; s < 2; ++s) {
                if (!on[s]) continue;
                gather2(nC, nL, cols, bs, cnv[s][0], lw_);
                gather2(nC, nL, cols, bs, cnv[s][1], iw_);
                gather2(nC, nL, cols, bs, cnv[s][2], rl_);
                gather2(nC, nL, cols, bs, cnv[s][3], ri_);
                double *t2 = tmp, *s2 = tmp + cl * nb, *g2 = tmp + 2 * cl * nb;
                if (orc_cloud_optics(lut, bs, nL, lw_, iw_, rl_, ri_, t2, s2, g2)) {
                    snprintf(errmsg, errlen, "cloud_optics(): effective radius outside LUT range");
                    rc = 1; goto done;
                }
                for (size_t i = 0; i < cl * nb; ++i) inc2s(t2[i], s2[i], g2[i], &ctau[i], &cssa[i], &cgg[i]);
            }
            /* precipitation (rain + snow) optics, delta-scaled (Fu) */
            static const float b0r[14] = {0.496f, 0.466f, 0.437f, 0.416f, 0.391f, 0.374f, 0.352f,
                                          0.183f, 0.048f, 0.012f, 0.000f, 0.000f, 0.000f, 0.000f};
            static const float b0s[14] = {0.460f, 0.460f, 0.460f, 0.460f, 0.460f, 0.460f, 0.460f,
                                          0.460f, 0.000f, 0.000f, 0.000f, 0.000f, 0.000f, 0.000f};
            static const float b1s[14] = {0.000f, 0.000f, 0.000f, 0.000f, 0.000f, 0.000f, 0.000f,
                                          0.000f, 1.62e-5f, 1.62e-5f, 0.000f, 0.000f, 0.000f, 0.000f};
            static const float c0r[14] = {0.980f, 0.975f, 0.965f, 0.960f, 0.955f, 0.952f, 0.950f,
                                          0.944f, 0.894f, 0.884f, 0.883f, 0.883f, 0.883f, 0.883f};
            static const float c0s[14] = {0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f, 0.970f,
                                          0.970f, 0.970f, 0.970f, 0.700f, 0.700f, 0.700f, 0.700f};
            const double a0r = (double)3.07e-3f, a0s = 0.0, a1s = 1.5;
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l) {
                    int c = cols[i];
                    if (!(F2(a->cld_frac, c, l, nC) > FTINY)) continue;
                    double rwp = F2(a->cld_rwp, c, l, nC), swp = F2(a->cld_swp, c, l, nC);
                    double res = F2(a->cld_resnow, c, l, nC);
                    double tr = rwp * a0r, ts = 0.0;
                    if (swp > 0.0 && res > 10.0) ts = swp * (double)1.09087f * (a0s + a1s / ((double)1.0315f * res));
                    for (int b = 0; b < nb && b < 14; ++b) {
                        double sr = tr * (1.0 - (double)b0r[b]);
                        double ar = sr * (double)c0r[b];
                        double sn = ts * (1.0 - ((double)b0s[b] + (double)b1s[b] * (double)1.0315f * res));
                        double an = sn * (double)c0s[b];
                        double tp = fmax(1.0e-12, tr + ts);
                        double sp = fmax(1.0e-12, sr + sn);
                        double ap = fmax(1.0e-12, ar + an);
                        double asyw = ap / fmax(1.0e-12, sp);
                        double ssaw = fmin(1.0 - (double)0.000001f, sp / tp);
                        double za1 = asyw * asyw, za2 = ssaw * za1;
                        size_t q = i + (size_t)bs * (l + (size_t)nL * b);
                        ptau[q] = (1.0 - za2) * tp;
                        pssa[q] = (ssaw - za2) / (1.0 - za2);
                        pg[q] = asyw / (1.0 + asyw);
                    }
                }
            for (size_t i = 0; i < cl * nb; ++i) inc2s(ptau[i], pssa[i], pg[i], &ctau[i], &cssa[i], &cgg[i]);
            for (int i = 0; i < bs; ++i)
                seeds[i] = (a->isubc_sw == 1) ? ng + cols[i] + 1 : (a->isubc_sw == 2 ? a->icseed_sw[cols[i]] : 0);
            if (a->cloud_overlap_param) gather2(nC, nL, cols, bs, a->cloud_overlap_param, ov);
            orc_mcica_mask(bs, nL, ng, seeds, cf, a->cloud_overlap_param ? ov : NULL, a->iovr, a->iovr_max,
                           a->iovr_maxrand, a->iovr_rand, a->iovr_dcorr, a->iovr_exp, a->iovr_exprand, mask);
        }
        /* aerosol increment (2str by band into 2str by g-point) */
        if (a->aersw_tau) {
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l)
                    for (int b = 0; b < nb; ++b) {
                        int c = cols[i];
                        double at = F3(a->aersw_tau, c, l, b, nC, nL), as = F3(a->aersw_ssa, c, l, b, nC, nL);
                        double ag = F3(a->aersw_g, c, l, b, nC, nL);
                        for (int gp = F2(k->band_lims_gpt, 0, b, 2) - 1; gp < F2(k->band_lims_gpt, 1, b, 2); ++gp) {
                            size_t o = gp + (size_t)ng * (l + (size_t)nL * i);
                            inc2s_r(R(at), R(as), R(ag), &tau[o], &ssa[o], &gg[o]);
                        }
                    }
        }
        int clear = !cloudy;
        if (clear || a->doSWclrsky) {
            orc_sw_solve(k, bs, nL, a->top_at_1, mu0, tau, ssa, gg, toa, adir, adif, cup, cdn, cdir);
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    double su = 0, sd = 0;
                    for (int b = 0; b < nb; ++b) {
                        su += cup[b + (size_t)nb * (v + (size_t)nV * i)];
                        sd += cdn[b + (size_t)nb * (v + (size_t)nV * i)];
                    }
                    F2(a->fluxswUP_clrsky, cols[i], v, nC) = su;
                    F2(a->fluxswDOWN_clrsky, cols[i], v, nC) = sd;
                }
            int vs = a->iSFC - 1;
            for (int i = 0; i < bs; ++i)
                for (int b = 0; b < nb; ++b) {
                    double fdir = cdn[b + (size_t)nb * (vs + (size_t)nV * i)], fdif = 0.0;
                    double *s = scl + 6 * i;
                    if (b + 1 < ibd) { s[2] += fdir; s[3] += fdif; }
                    if (b + 1 == ibd) { s[2] += fdir * 0.5; s[3] += fdif * 0.5; s[4] += fdir * 0.5; s[5] += fdif * 0.5; }
                    if (b + 1 > ibd) { s[4] += fdir; s[5] += fdif; }
                    s[0] = cdn[(ibd_uv - 1) + (size_t)nb * (vs + (size_t)nV * i)];
                }
        } else {
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    F2(a->fluxswUP_clrsky, cols[i], v, nC) = 0.0;
                    F2(a->fluxswDOWN_clrsky, cols[i], v, nC) = 0.0;
                }
            for (size_t i = 0; i < (size_t)6 * nC; ++i) a->scmpsw[i] = 0.0;
        }
        int vs = a->iSFC - 1;
        if (cloudy) {
            for (int i = 0; i < bs; ++i)
                for (int l = 0; l < nL; ++l)
                    for (int b = 0; b < nb; ++b) {
                        size_t q = i + (size_t)bs * (l + (size_t)nL * b);
                        for (int gp = F2(k->band_lims_gpt, 0, b, 2) - 1; gp < F2(k->band_lims_gpt, 1, b, 2); ++gp) {
                            size_t o = gp + (size_t)ng * (l + (size_t)nL * i);
                            int m = mask[l + (size_t)nL * gp + (size_t)nL * ng * i];
                            inc2s_r(m ? R(ctau[q]) : R(0.0), m ? R(cssa[q]) : R(0.0), m ? R(cgg[q]) : R(0.0),
                                    &tau[o], &ssa[o], &gg[o]);
                        }
                    }
            orc_sw_solve(k, bs, nL, a->top_at_1, mu0, tau, ssa, gg, toa, adir, adif, aup, adn, adr);
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    double su = 0, sd = 0;
                    for (int b = 0; b < nb; ++b) {
                        su += aup[b + (size_t)nb * (v + (size_t)nV * i)];
                        sd += adn[b + (size_t)nb * (v + (size_t)nV * i)];
                    }
                    F2(a->fluxswUP_allsky, cols[i], v, nC) = su;
                    F2(a->fluxswDOWN_allsky, cols[i], v, nC) = sd;
                }
            for (int i = 0; i < bs; ++i) {
                for (int b = 0; b < nb; ++b) {
                    size_t q = b + (size_t)nb * (vs + (size_t)nV * i);
                    double fdir = adr[q], fdif = adn[q] - adr[q];
                    double *s = sal + 6 * i;
                    if (b + 1 < ibd) { s[2] += fdir; s[3] += fdif; }
                    if (b + 1 == ibd) { s[2] += fdir * 0.5; s[3] += fdif * 0.5; s[4] += fdir * 0.5; s[5] += fdif * 0.5; }
                    if (b + 1 > ibd) { s[4] += fdir; s[5] += fdif; }
                    s[0] = adn[(ibd_uv - 1) + (size_t)nb * (vs + (size_t)nV * i)];
                }
                double *o = a->scmpsw + 6 * (size_t)cols[i];
                double *src = (zcf1[i] > EPS_PRE) ? sal + 6 * i : scl + 6 * i;
                o[2] = src[2]; o[3] = src[3]; o[4] = src[4]; o[5] = src[5];
                o[0] = (zcf1[i] > EPS_PRE) ? adn[(ibd_uv - 1) + (size_t)nb * (vs + (size_t)nV * i)]
                                           : cdn[(ibd_uv - 1) + (size_t)nb * (vs + (size_t)nV * i)];
                o[1] = cdn[(ibd_uv - 1) + (size_t)nb * (vs + (size_t)nV * i)];
            }
        } else {
            for (int v = 0; v < nV; ++v)
                for (int i = 0; i < bs; ++i) {
                    F2(a->fluxswUP_allsky, cols[i], v, nC) = F2(a->fluxswUP_clrsky, cols[i], v, nC);
                    F2(a->fluxswDOWN_allsky, cols[i], v, nC) = F2(a->fluxswDOWN_clrsky, cols[i], v, nC);
                }
            for (int i = 0; i < bs; ++i) {
                double *o = a->scmpsw + 6 * (size_t)cols[i];
                o[2] = scl[6 * i + 2]; o[3] = scl[6 * i + 3]; o[4] = scl[6 * i + 4]; o[5] = scl[6 * i + 5];
                o[0] = cdn[(ibd_uv - 1) + (size_t)nb * (vs + (size_t)nV * i)];
                o[1] = o[0];
            }
        }
    }
    if (a->htrsw) {
        /* GFS_rrtmgp_post_run: heating rate on daylit columns, 0 elsewhere */
        double *h = malloc(sizeof(double) * (size_t)nC * nL);
        orc_heating_rate(nC, nL, a->fluxswUP_allsky, a->fluxswDOWN_allsky, a->p_lev, h);
        for (size_t i = 0; i < (size_t)nC * nL; ++i) a->htrsw[i] = 0.0;
        for (int d = 0; d < a->nDay; ++d)
            for (int l = 0; l < nL; ++l) F2(a->htrsw, a->idx[d] - 1, l, nC) = F2(h, a->idx[d] - 1, l, nC);
        if (a->htrswc && a->doSWclrsky) {
            orc_heating_rate(nC, nL, a->fluxswUP_clrsky, a->fluxswDOWN_clrsky, a->p_lev, h);
            for (size_t i = 0; i < (size_t)nC * nL; ++i) a->htrswc[i] = 0.0;
            for (int d = 0; d < a->nDay; ++d)
                for (int l = 0; l < nL; ++l) F2(a->htrswc, a->idx[d] - 1, l, nC) = F2(h, a->idx[d] - 1, l, nC);
        }
        free(h);
    }
done:
    free(vmr); free(pl); free(tl); free(pv); free(cf); free(ov); free(tau); free(ray); free(ssa); free(gg);
    free(toa); free(mu0); free(adir); free(adif); free(ctau); free(cssa); free(cgg); free(tmp); free(ptau);
    free(pssa); free(pg); free(l4); free(cup); free(cdn); free(cdir); free(aup); free(adn); free(adr);
    free(mask); free(cols); free(seeds);
    return rc;
}

int orc_nthreads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
