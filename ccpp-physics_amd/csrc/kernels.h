// Launch interfaces of the gfx950 kernels (argument blocks passed by value).
#pragma once
#include "ccpprad_dev.h"

struct ccpprad_mlp;

namespace ccpprad {

struct VmrPtrs {
    const double *v[MAX_GAS];  // gas g (1-based) at v[g-1]; each (ncol_ld, nlay)
};

// k-distribution gas optics of ncols columns (column col0 + d of the CCPP
// arrays, or col_idx[col0 + d] - 1 with a gather map): tau and aux (LW
// Planck fraction, SW Rayleigh depth), [d][l][g] FP32.  row_mask / col_mask
// (both given or both null, LW only): the emulator's fallback, only rows
// with row_mask[d*nlay+l] != 0 of columns with col_mask[d] != 0 are written.
// err: bit 0 is set for a layer outside the table's p/T range.
void launch_gas_optics(const TablesView &T, int ncols, int nlay, int ncol_ld, const int *col_idx, int col0,
                       const double *p_lay, const double *p_lev, const double *t_lay, const VmrPtrs &vmr, int *err,
                       float *tau, float *aux, hipStream_t st, const unsigned char *row_mask = nullptr,
                       const unsigned char *col_mask = nullptr);
// LDS per block of the gas-optics kernel
size_t gas_gather_smem(int is_lw, int nflav, int nminor0, int nminor1, int ngpt, int neta);

struct LwSolveArgs {
    int nlay, ncol_ld, col0, top_at_1;
    const float *tau;      // [d][l][g]
    const float *pfrac;    // [d][l][g]
    const unsigned *mask;  // [d][l][MASK_WORDS] or null (clear sky)
    const float *ctau;     // [l][d][nband] cloud absorption depth (d = column of the launch) or null
    const float3 *copt;    // [l][d][nband] cloud (tau, co-albedo, g) for the scattering solver, or null
    const double *aer_tau; // Fortran (ncol_ld, nlay, nband) or null
    const float *aer1;     // the chunk's aerosol depths packed by launch_lw_aer_pack, [d][l][nband] FP32, or null
    const double *t_lay, *t_lev, *tsfc, *semis, *p_lev;
    float D[MAX_ANG], W[MAX_ANG];
    double *flux_up, *flux_dn, *flux_up2, *flux_dn2, *flux_jac, *htr;  // Fortran (ncol_ld, nlev)
};
void launch_lw_solve(const TablesView &T, const LwSolveArgs &A, int ncols, int nang, bool clouds, bool jac,
                     bool optang, hipStream_t st);
size_t lw_solve_smem(int nlay, int nband);
// Packs the aerosol depths of the chunk's columns (A.col0 ..) into
// out[d][l][nband] in FP32, the form the LW solvers read.
void launch_lw_aer_pack(const LwSolveArgs &A, int ncols, int nband, float *out, hipStream_t st);
// the two-stream solver with scattering (doGP_lwscat): all-sky fluxes, no Jacobian
void launch_lw_solve_2str(const TablesView &T, const LwSolveArgs &A, int ncols, bool clouds, hipStream_t st);
size_t lw_solve_2str_smem(int nlay, int nband);
void launch_heating_rate(int ncol, int nlay, const double *up, const double *dn, const double *p_lev,
                         double *htr, hipStream_t st);
// SW heating rate over the daylit columns, 0 elsewhere (day: nCol flags)
void launch_mark_days(int nday, const int *idx, unsigned char *day, hipStream_t st);
void launch_heating_rate_day(int ncol, int nlay, const unsigned char *day, const double *up, const double *dn,
                             const double *p_lev, double *htr, hipStream_t st);

// GFS_rrtmgp_pre's preparation of the radiation inputs (prep.hip)
struct PrepArgs {
    int ncol, nlev;                                   // nlev = layers
    const double *prsi, *prsl, *tgrs, *qh2o, *qo3, *tsfc;
    double *p_lay, *p_lev, *t_lay, *t_lev, *q_lay, *vmr_h2o, *vmr_o3;   // q_lay, qo3/vmr_o3 optional
    int *top_at_1;                                    // device int, optional
    double press_min, press_max, temp_min, temp_max;  // the k-distribution's range
};
void launch_prepare(const PrepArgs &A, hipStream_t st);

// GFS_rrtmgp_pre's virtual temperature and layer thicknesses (prep.hip)
struct GeomArgs {
    int ncol, nlev;                                   // nlev = layers
    double rdog, fvirt;                               // con_rd / con_g, con_fvirt
    const double *p_lev, *p_lay, *t_lay, *q_lay;
    double *tv_lay, *dz, *dzc, *dp;                   // tv_lay, dz, dp optional
};
void launch_layer_geometry(const GeomArgs &A, hipStream_t st);

struct CloudArgs {
    int nlay, ncol_ld, col0, blksz, nband, is_lw;
    const int *col_idx;            // 1-based gather map (SW daylit) or null
    const unsigned char *blkflag;  // per scheme block: any cloudy column
    const double *cld_frac, *lwp, *iwp, *reliq, *reice;
    const double *cnv[4], *pbl[4];  // lwp, iwp, reliq, reice (null = off)
    const double *rwp, *swp, *resnow;
    float *ctau;              // LW without scattering: [l][d][b] absorption depth (d = column of the launch)
    float3 *copt;             // SW, and LW with scattering: [l][d][b] (tau, co-albedo 1 - ssa, g)
    double *cldtau;           // SW: (ncol_ld, nlay) band-11 cloud tau, or null
    int cldtau_blk;           // the one scheme block whose columns write cldtau
    int *err;
};
void launch_cloud_optics(const LutView &L, const CloudArgs &A, int ncols, hipStream_t st);
size_t cloud_lut_floats(const LutView &L);   // LUT floats the cloud kernel stages in LDS

struct McicaArgs {
    int nlay, ngpt, ncol_ld, col0, blksz, isubc;
    int iovr, iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr, iovr_exp, iovr_exprand;
    const int *col_idx;
    const unsigned char *blkflag;
    const double *cld_frac, *overlap;
    const int *icseed;
    const uint32_t *mtpow;  // 69069^i mod 2^32, i = 0..623
    unsigned *mask;         // [d][l][MASK_WORDS]
};
void launch_mcica(const McicaArgs &A, int ncols, hipStream_t st);
void launch_block_flags(int ncols, int nlay, int ncol_ld, const int *col_idx, int blksz, const double *cld_frac,
                        unsigned char *blkflag, unsigned char *colflag, hipStream_t st);

struct SwSolveArgs {
    int nlay, ncol_ld, col0, top_at_1, blksz, isfc, write_clr_zero, role;  // role 0 all sky, 1 clear sky
    const int *col_idx;
    const float *tau_abs, *tau_ray;        // [d][l][g]
    const unsigned *mask;                  // [d][l][MASK_WORDS] or null
    const float3 *copt;                    // [l][d][nband] (tau, co-albedo 1 - ssa, g) or null
    const unsigned char *blkflag;
    const double *aer_tau, *aer_ssa, *aer_g;  // (ncol_ld, nlay, nband) or null
    // the chunk's aerosol depths packed by launch_sw_aer_pack, [d][l][nband]:
    // (tau ssa, tau (1 - ssa), tau ssa g) in FP32, the extinction being the
    // sum of the first two; null without aerosols
    const float3 *aer4;
    const double *coszen, *alb_nir_dir, *alb_nir_dif, *alb_uv_dir, *alb_uv_dif, *p_lev;
    const int *alb_cat;                    // [nband] 0 NIR, 1 average, 2 UV-VIS, 3 unset
    double *up_all, *dn_all, *up_clr, *dn_clr;  // (ncol_ld, nlev)
    double *sb;                            // [d][3][nband]: dn, dn_dir (all sky), dn (clear sky)
    double *htr, *htr_clr;                 // (ncol_ld, nlay) or null
};
void launch_sw_solve(const TablesView &T, const SwSolveArgs &A, int ncols, bool clouds, hipStream_t st);
// Packs the aerosol optics of the chunk's columns (A.col0 .., A.col_idx) into
// out[d][l][nband], the form the solver reads.
void launch_sw_aer_pack(const SwSolveArgs &A, int ncols, int nband, float3 *out, hipStream_t st);
size_t sw_solve_smem(int nlay, int nband);

struct ScmpArgs {
    int ndays, blksz, nband, ibd, ibd_uv, do_clr;  // ibd, ibd_uv 1-based
    const int *col_idx;
    const unsigned char *blkflag, *colflag;
    const double *sb;        // [d][3][nband]
    const int *last_wipe;    // block index of the last reset, -1 if none
    double *scmpsw;          // [nCol][6]
};
void launch_scmpsw(const ScmpArgs &A, int ncol, unsigned char *daymap, int *last_wipe, hipStream_t st);

// Learned gas optics: fused MLP inference (mlp.hip)
constexpr int MLP_NIN = 7;
struct MlpInferArgs {
    int nlay, ncol_ld, col0, ncols;
    const double *p_lay, *p_lev, *t_lay;
    const double *vmr[5];               // h2o, o3, co2, ch4, n2o
    const unsigned char *image;         // device copy of the LDS image
    float xmean[MLP_NIN], xistd[MLP_NIN], lo[MLP_NIN], hi[MLP_NIN];
    const float *ymean, *ystd;          // [256]
    float *tau, *pfrac;                 // [d][l][g]
    unsigned char *oob;                 // [d][l] 1 = out of hull
    unsigned char *colflag;             // [d] any row out of hull
};
void launch_mlp_infer(const struct ::ccpprad_mlp *m, MlpInferArgs &a, hipStream_t st);

void launch_fill_f64(double *p, size_t n, double v, hipStream_t st);
void launch_fill_i32(int *p, size_t n, int v, hipStream_t st);
void launch_fill_u8(unsigned char *p, size_t n, hipStream_t st);
void launch_toa_norm(const double *solar, int ngpt, double solcon, double *out64, float *out32, hipStream_t st);

}  // namespace ccpprad
