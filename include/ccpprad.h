/*
 * ccpprad.h -- C-ABI of the MI355X RRTMGP radiation hot path.
 *
 * The entry points replace the CCPP RRTMGP scheme calls of ccpp-physics:
 *   ccpprad_lw_init / ccpprad_lw_run   <- rrtmgp_lw_main_init / _run
 *        (physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:43-91, :104-613)
 *   ccpprad_sw_init / ccpprad_sw_run   <- rrtmgp_sw_main_init / _run
 *        (physics/Radiation/RRTMGP/rrtmgp_sw_main.F90:34-79, :87-683)
 *   ccpprad_heating_rate               <- compute_heating_rate as called by
 *        GFS_rrtmgp_post_run (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_post.F90:159-171,251-266)
 *   ccpprad_post / ccpprad_post_dev    <- GFS_rrtmgp_post_run
 *        (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_post.F90:35-389): heating
 *        rates and flux bookkeeping (host pointers) / the bookkeeping (device)
 *   ccpprad_prepare_inputs_dev         <- the radiation-input preparation of
 *        GFS_rrtmgp_pre_run (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_pre.F90:217-274,349-371)
 *   ccpprad_layer_geometry_dev         <- the virtual temperature and layer thicknesses of
 *        GFS_rrtmgp_pre_run (GFS_rrtmgp_pre.F90:289,294-341)
 *   ccpprad_cloud_overlap(_dev)        <- GFS_rrtmgp_cloud_overlap_run
 *        (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_cloud_overlap.F90:31-138)
 *   ccpprad_*_gas_optics_dev           <- ty_gas_optics_rrtmgp%gas_optics
 *        (call sites rrtmgp_lw_main.F90:337-345, rrtmgp_sw_main.F90:346-353)
 *
 * Conventions
 *  - Every CCPP field is FP64, Fortran column-major, column index fastest:
 *    p_lay(nCol,nLay), p_lev(nCol,nLay+1), aer*(nCol,nLay,nBand).
 *  - Logical flags are int (0 = .false.).  Integer index tables are 1-based.
 *  - Every entry returns 0 on success or a nonzero status; the message is
 *    written into errmsg (NUL-terminated, at most errlen bytes).  The Fortran
 *    scheme modules map a nonzero status to errflg = 1.
 *  - *_run take HOST pointers and copy in/out; *_run_dev take DEVICE pointers
 *    (inputs already in HBM) and an explicit hipStream_t (passed as void*).
 *  - The library holds no mutable global state in the run path: one context
 *    per scheme, and calls on distinct streams may overlap.
 */
#ifndef CCPPRAD_H
#define CCPPRAD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCPPRAD_MAX_GAS 6

/* Minor-absorber description of one atmospheric regime.  Mirrors the
 * minor_* arrays loaded at rrtmgp_lw_gas_optics.F90:29-39,70-74,330-341. */
typedef struct ccpprad_minor {
    int n;                           /* number of minor intervals            */
    int ncontrib;                    /* total contributors (sum of widths)   */
    const int *limits_gpt;           /* (2, n)  first/last g-point, 1-based  */
    const int *scales_with_density;  /* (n)                                  */
    const int *idx_scaling;          /* (n) gas index of scaling gas, 0=none */
    const int *scale_by_complement;  /* (n)                                  */
    const int *idx_minor;            /* (n) gas index of the minor absorber  */
    const int *kminor_start;         /* (n) 1-based start in contributors    */
    const double *kminor;            /* (ncontrib, neta, ntemp)  cm2/molec   */
} ccpprad_minor;

/* A k-distribution: the arguments of ty_gas_optics_rrtmgp%load
 * (rrtmgp_lw_gas_optics.F90:439-448, rrtmgp_sw_gas_optics.F90:487-496).
 * Gas indices: 0 = dry air, 1..ngas in the order h2o co2 o3 n2o ch4 o2. */
typedef struct ccpprad_kdist {
    int is_lw;
    int ngas, nband, ngpt, ntemp, npres, neta, nflav, nplanck;
    const int *band_lims_gpt;        /* (2, nband)                          */
    const double *band_lims_wvn;     /* (2, nband) cm-1                     */
    const int *flavor;               /* (2, nflav) gas indices              */
    const int *gpoint_flavor;        /* (2, ngpt) flavor (lower, upper)     */
    const double *press_ref;         /* (npres) Pa, decreasing              */
    const double *temp_ref;          /* (ntemp) K, uniform                  */
    double press_ref_trop;           /* Pa                                  */
    const double *vmr_ref;           /* (2, 0:ngas, ntemp)                  */
    const double *kmajor;            /* (ngpt, neta, npres+1, ntemp)        */
    ccpprad_minor minor_lower, minor_upper;
    /* LW only (NULL for SW) */
    const double *planck_frac;       /* (ngpt, neta, npres+1, ntemp)        */
    const double *totplnk;           /* (nplanck, nband)                    */
    const double *optimal_angle_fit; /* (2, nband)                          */
    /* SW only (NULL for LW) */
    const double *krayl_lower;       /* (ngpt, neta, ntemp)                 */
    const double *krayl_upper;       /* (ngpt, neta, ntemp)                 */
    const double *solar_src;         /* (ngpt)                              */
} ccpprad_kdist;

/* Lookup-table cloud optics: the LUT arguments of ty_cloud_optics%load
 * (rrtmgp_lw_cloud_optics.F90:184-191,362-367; SW :352-366). */
typedef struct ccpprad_cloud_lut {
    int nband, nsize_liq, nsize_ice, nrghice, icergh;
    double radliq_lwr, radliq_upr, radice_lwr, radice_upr;
    const double *extliq, *ssaliq, *asyliq;   /* (nsize_liq, nband)          */
    const double *extice, *ssaice, *asyice;   /* (nsize_ice, nband, nrghice) */
    /* Pade form (the Pade arguments of ty_cloud_optics%load,
     * rrtmgp_lw_cloud_optics.F90:192-205,345-350), used when pade != 0 (the
     * LUT arrays and sizes above are then unused; radliq/radice bounds are
     * the extinction size bounds).  Coefficients (nband, 3 size regimes,
     * ncoeff): extinction 6 (numerator degree 2, denominator 3), co-albedo
     * and asymmetry 5 (2, 2); ice (.., nrghice).  Size-regime bounds (4). */
    int pade;
    const double *pade_extliq, *pade_ssaliq, *pade_asyliq;
    const double *pade_extice, *pade_ssaice, *pade_asyice;
    const double *pade_sizreg_extliq, *pade_sizreg_ssaliq, *pade_sizreg_asyliq;
    const double *pade_sizreg_extice, *pade_sizreg_ssaice, *pade_sizreg_asyice;
} ccpprad_cloud_lut;

/* Argument list of rrtmgp_lw_main_run (rrtmgp_lw_main.F90:104-194), in order.
 * active_gases_array is bound at init.  htrlw/htrlwc are optional extra
 * outputs (NULL = not computed): the K/s heating rates GFS_rrtmgp_post_run
 * derives from the all-sky / clear-sky fluxes. */
typedef struct ccpprad_lw_args {
    int doLWrad, doLWclrsky, top_at_1, doGP_lwscat, use_LW_jacobian;
    int doGP_sgs_cnv, doGP_sgs_pbl;
    int nCol, nLay, nGases, rrtmgp_phys_blksz, nGauss_angles;
    const int *icseed_lw;                                   /* (nCol) */
    int iovr, iovr_convcld, iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr;
    int iovr_exp, iovr_exprand, isubc_lw;
    const double *semis, *tsfg;                             /* (nCol) */
    const double *p_lay, *p_lev, *t_lay, *t_lev;
    const double *vmr_o2, *vmr_h2o, *vmr_o3, *vmr_ch4, *vmr_n2o, *vmr_co2;
    const double *cld_frac, *cld_lwp, *cld_reliq, *cld_iwp, *cld_reice;
    const double *cld_swp, *cld_resnow, *cld_rwp, *cld_rerain, *precip_frac;
    const double *cld_cnv_lwp, *cld_cnv_reliq, *cld_cnv_iwp, *cld_cnv_reice;
    const double *cld_pbl_lwp, *cld_pbl_reliq, *cld_pbl_iwp, *cld_pbl_reice;
    const double *cloud_overlap_param;                      /* (nCol,nLay) */
    const double *aerlw_tau, *aerlw_ssa, *aerlw_g;          /* (nCol,nLay,nBand) or NULL */
    double *fluxlwUP_allsky, *fluxlwDOWN_allsky;            /* (nCol,nLay+1) */
    double *fluxlwUP_clrsky, *fluxlwDOWN_clrsky;
    double *fluxlwUP_jac, *fluxlwUP_radtime, *fluxlwDOWN_radtime;
    double *htrlw, *htrlwc;                                 /* (nCol,nLay) or NULL */
} ccpprad_lw_args;

/* Argument list of rrtmgp_sw_main_run (rrtmgp_sw_main.F90:87-189), in order.
 * scmpsw is cmpfsw_type(nCol) = 6 doubles per column in the order
 * uvbfc, uvbf0, nirbm, nirdf, visbm, visdf (Radiation/RRTMG/radsw_param.f:111-118).
 * idx holds the 1-based daylit column indices (nDay of them). */
typedef struct ccpprad_sw_args {
    int doSWrad, doSWclrsky, top_at_1, doGP_sgs_cnv, doGP_sgs_pbl;
    int nCol, nDay, nLay, nGases, rrtmgp_phys_blksz;
    const int *idx;                                         /* (nDay) */
    const int *icseed_sw;                                   /* (nCol) */
    int iovr, iovr_convcld, iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr;
    int iovr_exp, iovr_exprand, isubc_sw, iSFC;
    const double *sfc_alb_nir_dir, *sfc_alb_nir_dif;        /* (nCol) */
    const double *sfc_alb_uvvis_dir, *sfc_alb_uvvis_dif, *coszen;
    const double *p_lay, *p_lev, *t_lay, *t_lev;
    const double *vmr_o2, *vmr_h2o, *vmr_o3, *vmr_ch4, *vmr_n2o, *vmr_co2;
    const double *cld_frac, *cld_lwp, *cld_reliq, *cld_iwp, *cld_reice;
    const double *cld_swp, *cld_resnow, *cld_rwp, *cld_rerain, *precip_frac;
    const double *cld_cnv_lwp, *cld_cnv_reliq, *cld_cnv_iwp, *cld_cnv_reice;
    const double *cld_pbl_lwp, *cld_pbl_reliq, *cld_pbl_iwp, *cld_pbl_reice;
    const double *cloud_overlap_param;
    const double *aersw_tau, *aersw_ssa, *aersw_g;          /* (nCol,nLay,nBand) or NULL */
    double solcon;
    double *scmpsw;                                         /* (6, nCol) */
    double *fluxswUP_allsky, *fluxswDOWN_allsky;
    double *fluxswUP_clrsky, *fluxswDOWN_clrsky;
    double *cldtausw;                                       /* (nCol,nLay) */
    double *htrsw, *htrswc;                                 /* (nCol,nLay) or NULL */
} ccpprad_sw_args;

typedef struct ccpprad_ctx ccpprad_ctx;

/* ---- scheme entry points (host pointers) ---------------------------- */

/* rrtmgp_lw_main_init: device-resident tables for one GPU.  gas_names lists
 * the ngases active gas names (CCPP active_gases_array); every name must be
 * one of the k-distribution's gases.  nrghice is the ice roughness category
 * used when lut->icergh is 0. */
int ccpprad_lw_init(const ccpprad_kdist *kd, const ccpprad_cloud_lut *lut,
                    const char *const *gas_names, int ngases, int nLay,
                    int rrtmgp_phys_blksz, ccpprad_ctx **ctx, char *errmsg, int errlen);
/* rrtmgp_sw_main_init */
int ccpprad_sw_init(const ccpprad_kdist *kd, const ccpprad_cloud_lut *lut,
                    const char *const *gas_names, int ngases, int nLay,
                    int rrtmgp_phys_blksz, ccpprad_ctx **ctx, char *errmsg, int errlen);
/* Init from table files: the reference's netCDF k-distribution and cloud
 * files (rrtmgp_lw_gas_optics.F90:130-327, rrtmgp_sw_gas_optics.F90:288-376,
 * rrtmgp_lw_cloud_optics.F90:130-290), read natively in the netCDF classic
 * formats (CDF-1, CDF-2, CDF-5; netCDF-4/HDF5 files are refused with a
 * message) and reduced to the active gases as ty_gas_optics_rrtmgp%load
 * does, or files in the package's binary table format; each file's format is
 * taken from its signature.  A netCDF cloud file gives the lookup tables
 * when it has them, else its Pade coefficients.  nrghice is the namelist's
 * ice-roughness category: nonzero selects that category of the cloud file
 * (an error beyond the file's count), as the reference's nonzero nrghice
 * overrides the file's (rrtmgp_lw_cloud_optics.F90:171-173,377,
 * rrtmgp_sw_cloud_optics.F90:160-162,366); 0 keeps the file's choice (the
 * last category of a netCDF file). */
int ccpprad_lw_init_files(const char *gas_file, const char *cloud_file, int nrghice,
                          const char *const *gas_names, int ngases, int nLay,
                          int rrtmgp_phys_blksz, ccpprad_ctx **ctx, char *errmsg, int errlen);
int ccpprad_sw_init_files(const char *gas_file, const char *cloud_file, int nrghice,
                          const char *const *gas_names, int ngases, int nLay,
                          int rrtmgp_phys_blksz, ccpprad_ctx **ctx, char *errmsg, int errlen);
/* The same with doGP_cldoptics_PADE (rrtmgp_lw_cloud_optics.F90:192-205,
 * 261-288,345-350): the netCDF cloud file's Pade coefficients, also where the
 * file holds lookup tables too. */
int ccpprad_lw_init_files_pade(const char *gas_file, const char *cloud_file, int nrghice,
                               const char *const *gas_names, int ngases, int nLay,
                               int rrtmgp_phys_blksz, ccpprad_ctx **ctx, char *errmsg, int errlen);
int ccpprad_sw_init_files_pade(const char *gas_file, const char *cloud_file, int nrghice,
                               const char *const *gas_names, int ngases, int nLay,
                               int rrtmgp_phys_blksz, ccpprad_ctx **ctx, char *errmsg, int errlen);

/* rrtmgp_lw_main_run / rrtmgp_sw_main_run with host arrays. */
int ccpprad_lw_run(ccpprad_ctx *ctx, const ccpprad_lw_args *a, char *errmsg, int errlen);
int ccpprad_sw_run(ccpprad_ctx *ctx, const ccpprad_sw_args *a, char *errmsg, int errlen);

/* Same, every array pointer in the args struct is a DEVICE pointer.  No
 * host<->device copies, no synchronisation; errors in the inputs that need
 * a device-side check (table range) are reported only when validate != 0,
 * which synchronises the stream. */
int ccpprad_lw_run_dev(ccpprad_ctx *ctx, const ccpprad_lw_args *a, void *stream,
                       int validate, char *errmsg, int errlen);
int ccpprad_sw_run_dev(ccpprad_ctx *ctx, const ccpprad_sw_args *a, void *stream,
                       int validate, char *errmsg, int errlen);

/* Free a context (tables, work buffers).  The reference has no scheme-level
 * finalize; this plays the role of GFS_rrtmgp_setup_finalize
 * (GFS_rrtmgp_setup.F90:260-273) for the device state. */
void ccpprad_finalize(ccpprad_ctx *ctx);

/* Columns per launch (default 65,536, at most what keeps the solvers'
 * 32-bit cloud-record offsets in range): the work buffers of a stream hold
 * one launch, so this bounds their size.  No counterpart in the reference
 * (its block size rrtmgp_phys_blksz is the host's, and stays that). */
int ccpprad_set_launch_columns(ccpprad_ctx *ctx, int ncol);
int ccpprad_launch_columns(const ccpprad_ctx *ctx);

/* ---- component entry points ----------------------------------------- */

/* compute_heating_rate (mo_heating_rates, called at
 * GFS_rrtmgp_post.F90:11,159-171,249-266): FP64 on the device.
 * htr(nCol,nLay) [K/s] = (F_up(k+1) - F_up(k) - F_dn(k+1) + F_dn(k)) g / (cp (p(k+1) - p(k))),
 * fluxes and p_lev (nCol, nLay+1).  Host pointers: copies in, runs, copies
 * out, synchronous (the Fortran mo_heating_rates module calls this). */
int ccpprad_heating_rate(int nCol, int nLay, const double *flux_up, const double *flux_dn,
                         const double *p_lev, double *htr, char *errmsg, int errlen);
/* The same with device pointers, ordered on stream. */
int ccpprad_heating_rate_dev(int nCol, int nLay, const double *flux_up,
                             const double *flux_dn, const double *p_lev, double *htr,
                             void *stream, char *errmsg, int errlen);

/* The McICA cloud mask of a scheme call on its own: random_setseed /
 * random_number per column and sampled_mask (rrtmgp_lw_main.F90:353-362,
 * 441-493; rrtmgp_sw_main.F90:263-273,479-523; rrtmgp_sampling.F90:84-182)
 * for the flags, seeds and overlap of the args struct (only the fields the
 * sampling reads are used: dimensions, blksz, isubc, icseed, the iovr
 * values, cld_frac, cloud_overlap_param and, for SW, nDay and idx).  Blocks
 * with no cloudy column get no bits, as the reference skips them.  Device
 * pointers; mask (MASK words = 4, nLay, n) uint32 with n = nCol (LW) or
 * nDay (SW, daylit order): bit (g-1) mod 32 of word (g-1)/32 is g-point g. */
int ccpprad_lw_mcica_mask_dev(ccpprad_ctx *ctx, const ccpprad_lw_args *a, unsigned *mask,
                              void *stream, char *errmsg, int errlen);
int ccpprad_sw_mcica_mask_dev(ccpprad_ctx *ctx, const ccpprad_sw_args *a, unsigned *mask,
                              void *stream, char *errmsg, int errlen);

/* ccpprad_lw_mcica_mask_dev without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The per-block cloud flags live in a
 * library-owned device array of 2^20 bytes per GPU, held until the process
 * ends, so nCol / rrtmgp_phys_blksz may not exceed 1,048,576 (larger calls
 * are rejected with a message).  The caller still passes the inputs and
 * the mask.  Calls that overlap on one GPU share that array and race; a
 * context must have been initialised on the current device first.  Same
 * result as ccpprad_lw_mcica_mask_dev. */
int ccpprad_lw_mcica_mask_noalloc_dev(ccpprad_ctx *ctx, const ccpprad_lw_args *a, unsigned *mask,
                                      void *stream, char *errmsg, int errlen);

/* Argument list of GFS_rrtmgp_post_run
 * (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_post.F90:35-389) less
 * its heating rates (those come fused out of the _run entries, or from
 * ccpprad_heating_rate_dev).  Device pointers, FP64, Fortran layouts:
 * level fluxes (nCol, nLev+1); the flux types are arrays of derived types,
 * i.e. per column their fields in declaration order
 * (Radiation/RRTMG/radlw_param.f:78-92, radsw_param.f:79-118):
 *   topflw (2, nCol) upfxc, upfx0         sfcflw (4, nCol) upfxc, upfx0, dnfxc, dnfx0
 *   topfsw (3, nCol) upfxc, dnfxc, upfx0  sfcfsw (4, nCol) upfxc, dnfxc, upfx0, dnfx0
 *   scmpsw (6, nCol) uvbfc, uvbf0, nirbm, nirdf, visbm, visdf
 * iSFC / iTOA, mtopa / mbota are 1-based level indices; cldsa (nCol, 5);
 * aerodp (nCol, naerodp >= 6); fluxr (nCol, nfxr >= 45) is accumulated when
 * save_diag.  The cloud and aerosol diagnostics may be NULL when save_diag
 * is 0. */
typedef struct ccpprad_post_args {
    int nCol, nLev, nDay, iSFC, iTOA;
    int doLWrad, doSWrad, save_diag, naerodp, nfxr;
    double fhlwr, fhswr, raddt;
    const double *fluxlwUP_allsky, *fluxlwDOWN_allsky, *fluxlwUP_clrsky, *fluxlwDOWN_clrsky;
    const double *fluxswUP_allsky, *fluxswDOWN_allsky, *fluxswUP_clrsky, *fluxswDOWN_clrsky;
    const double *scmpsw;
    const double *sfc_alb_nir_dir, *sfc_alb_nir_dif, *sfc_alb_uvvis_dir, *sfc_alb_uvvis_dif;
    const double *tsfa, *coszen, *coszdg, *htrlw, *p_lev;
    const double *cldsa, *cldtaulw, *cldtausw, *aerodp;
    const int *mtopa, *mbota;
    double *topflw, *sfcflw, *sfcdlw, *sfculw, *tsflw, *htrlwu;
    double *topfsw, *sfcfsw;
    double *nirbmdi, *nirdfdi, *visbmdi, *visdfdi, *nirbmui, *nirdfui, *visbmui, *visdfui;
    double *sfcnsw, *sfcdsw, *fluxr;
} ccpprad_post_args;

/* GFS_rrtmgp_post_run's flux bookkeeping on the device (see the struct). */
int ccpprad_post_dev(const ccpprad_post_args *args, void *stream, char *errmsg, int errlen);

/* The heating-rate half of GFS_rrtmgp_post_run (GFS_rrtmgp_post.F90:152-171,
 * 242-266, 298-316) for ccpprad_post: LW all-sky always and clear-sky when
 * do_lw_clrsky_hr; SW zeroed, then the daylit rows idxday(1:nDay) (1-based,
 * distinct, in 1..nCol); htrswc likewise when do_sw_clrsky_hr.  (nCol, nLev)
 * FP64 in K/s, written only where the reference writes them. */
typedef struct ccpprad_post_hr_args {
    int do_lw_clrsky_hr, do_sw_clrsky_hr;
    const int *idxday;
    double *htrlw, *htrlwc, *htrsw, *htrswc;
} ccpprad_post_hr_args;

/* GFS_rrtmgp_post_run on HOST pointers: the heating rates (FP64 on the GPU,
 * from the level fluxes of args) and then the bookkeeping of
 * ccpprad_post_dev, whose htrlw input is the freshly computed LW heating
 * rate (args->htrlw is ignored).  Every output of args is intent(inout):
 * it is copied in and back, so what the reference leaves alone keeps its
 * value.  A private stream; returns after the copies back. */
int ccpprad_post(const ccpprad_post_args *args, const ccpprad_post_hr_args *hr, char *errmsg, int errlen);

/* The radiation-input preparation of GFS_rrtmgp_pre_run
 * (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_pre.F90:217-274,349-371),
 * FP64 on the device: q_lay = max(qgrs_h2o, 1e-6); p_lev = prsi; p_lay, t_lay
 * = prsl, tgrs bounded into the k-distribution's range of the LW context
 * (min + epsilon, max - epsilon); t_lev by cmp_tlev (radiation_tools.F90:16-84)
 * bounded the same way; vmr_h2o = q/(1-q) amd/amw; vmr_o3 = max(1e-10,
 * qgrs_o3) amd/amo3 when qgrs_o3 is given (else vmr_o3 is not written: the
 * host's ozone climatology).  nLev = layers; fields (nCol, nLev[+1]).
 * top_at_1 (a device int, may be NULL) receives prsi(1,1) < prsi(1,nLev).
 * q_lay may be NULL.  Device pointers; ordered on stream. */
int ccpprad_prepare_inputs_dev(const ccpprad_ctx *ctx, int nCol, int nLev, const double *prsi,
                               const double *prsl, const double *tgrs, const double *qgrs_h2o,
                               const double *qgrs_o3, const double *tsfc, double *p_lay,
                               double *p_lev, double *t_lay, double *t_lev, double *q_lay,
                               double *vmr_h2o, double *vmr_o3, int *top_at_1, void *stream,
                               char *errmsg, int errlen);

/* The virtual temperature and layer thicknesses of GFS_rrtmgp_pre_run
 * (GFS_rrtmgp_pre.F90:289,294-341), FP64 on the device, from the prepared
 * fields (bounded t_lay, floored q_lay, p_lev = prsi):
 *   tv_lay  = t_lay (1 + con_fvirt q_lay)
 *   deltaP  = |p_lev(k+1) - p_lev(k)|
 *   deltaZ  = con_rd/con_g |ln p_lev(k+1) - ln p_lev(k)| tv_lay       (m)
 *   deltaZc = height of the layer centre above the centre of the layer
 *             below it (above the surface for the lowest layer); the centre
 *             lies at the fraction |ln p_lev(bottom) - ln p_lay| /
 *             |ln p_lev(bottom) - ln p_lev(top)| of the layer's thickness.
 * The vertical order comes from the first column (p_lev(1,1) < p_lev(1,nLev)
 * means the top is at index 1), as in the reference.  nLev = layers; fields
 * (nCol, nLev), p_lev (nCol, nLev+1).  tv_lay, deltaZ and deltaP may be NULL
 * (deltaZc is always written).  Device pointers; ordered on stream. */
int ccpprad_layer_geometry_dev(int nCol, int nLev, double con_rd, double con_g, double con_fvirt,
                               const double *p_lev, const double *p_lay, const double *t_lay,
                               const double *q_lay, double *tv_lay, double *deltaZ,
                               double *deltaZc, double *deltaP, void *stream, char *errmsg,
                               int errlen);

/* GFS_rrtmgp_cloud_overlap_run (GFS_rrtmgp_cloud_overlap.F90:31-138) with
 * cmp_dcorr_lgth_hogan / _oreopoulos and get_alpha_exper
 * (physics/Radiation/radiation_cloud_overlap.F90:28-145), FP64 on the device.
 * Nothing is written when neither doSWrad nor doLWrad.  de_lgth (nCol) is 0
 * unless idcor selects Hogan et al. (2010), Oreopoulos et al. (2012) or the
 * constant dcorr_con; cloud_overlap_param (nCol, nLev) is
 * exp(-deltaZc 0.001 / de_lgth) from the second layer on (0 in layer 1) for
 * iovr in {iovr_dcorr, iovr_exp, iovr_exprand}, with exponential-random
 * setting 0 where a clear layer follows a cloudy one, else 0;
 * cnv_cloud_overlap_param is computed the same way from cld_cnv_frac and
 * iovr_convcld, and written only when imfdeepcnv is imfdeepcnv_samf or
 * imfdeepcnv_gf; precip_overlap_param = cloud_overlap_param.
 * The reference's unsuffixed constants (0.6, 2.78 and 4.6 of the Hogan fit,
 * the 0.001 m -> km factor) take the precision of its build's default real:
 * literal_r8 = 1 for physics compiled with default-real-8 (UFS and the SCM),
 * 0 for the compiler's default single-precision literals.  The reference's
 * p_lev, p_lay, tv_lay, con_g, con_rd, con_epsq and top_at_1 arguments are
 * unused by it and omitted.  lat in radians; julian day; yearlen in days. */
typedef struct ccpprad_cloud_overlap_args {
    int nCol, nLev, yearlen, doSWrad, doLWrad, literal_r8;
    int idcor, idcor_con, idcor_hogan, idcor_oreopoulos;
    int iovr, iovr_convcld, iovr_dcorr, iovr_exp, iovr_exprand;
    int imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf;
    double julian, con_pi, dcorr_con;
    const double *lat, *deltaZc, *cld_frac, *cld_cnv_frac;
    double *de_lgth, *cloud_overlap_param, *cnv_cloud_overlap_param, *precip_overlap_param;
} ccpprad_cloud_overlap_args;

int ccpprad_cloud_overlap_dev(const ccpprad_cloud_overlap_args *args, void *stream, char *errmsg,
                              int errlen);
/* The same on HOST pointers (GFS_rrtmgp_cloud_overlap_run's own argument
 * kinds): the outputs are copied in and back, so the ones the reference
 * leaves unwritten keep their values.  A private stream; returns after the
 * copies back. */
int ccpprad_cloud_overlap(const ccpprad_cloud_overlap_args *args, char *errmsg, int errlen);

/* ty_gas_optics_rrtmgp%gas_optics, LW (rrtmgp_lw_main.F90:337-345).  Device
 * pointers.  vmr is (nCol, nLay, ngas) in the k-distribution's gas order.
 * Outputs g-point innermost, FP32: tau(ngpt,nLay,nCol),
 * pfrac(ngpt,nLay,nCol) (Planck fraction; lay_src = pfrac * B_band(T)).
 * With an emulator attached (ccpprad_lw_set_gas_optics_mlp) this is the
 * gas-optics stage of ccpprad_lw_run_dev: the MLP, and the table path on
 * the layers outside its training hull (nLay must then equal init's). */
int ccpprad_lw_gas_optics_dev(ccpprad_ctx *ctx, int nCol, int nLay,
                              const double *p_lay, const double *p_lev,
                              const double *t_lay, const double *vmr,
                              float *tau, float *pfrac, void *stream,
                              char *errmsg, int errlen);
/* SW variant (rrtmgp_sw_main.F90:346-353): absorption and Rayleigh depths. */
int ccpprad_sw_gas_optics_dev(ccpprad_ctx *ctx, int nCol, int nLay,
                              const double *p_lay, const double *p_lev,
                              const double *t_lay, const double *vmr,
                              float *tau_abs, float *tau_ray, void *stream,
                              char *errmsg, int errlen);

/* ---- learned LW gas optics (no counterpart in the reference) -------------
 * A 7 -> 128 -> 128 -> 128 -> 256 softplus MLP on BF16 MFMA with FP32
 * accumulation that stands in for the k-distribution lookup of the LW
 * gas_optics call (rrtmgp_lw_main.F90:337-345).  Inputs per layer: T,
 * ln p, ln vmr of h2o o3 co2 ch4 n2o (standardised with xmean/xstd); outputs
 * 128 ln(tau/col_dry) and 128 Planck fractions (standardised with
 * ymean/ystd).  Flat parameter order: W1(7x128, input-major), b1, W2(128x128),
 * b2, W3(128x128), b3, W4(128x256), b4; 67072 floats.  Layers whose inputs
 * leave the [lo, hi] training hull fall back to the table path. */
typedef struct ccpprad_mlp ccpprad_mlp;
int ccpprad_mlp_create(ccpprad_mlp **m, char *errmsg, int errlen);
void ccpprad_mlp_destroy(ccpprad_mlp *m);
int ccpprad_mlp_nparams(void);
int ccpprad_mlp_set_params(ccpprad_mlp *m, const float *host_params, char *errmsg, int errlen);
int ccpprad_mlp_get_params(ccpprad_mlp *m, float *host_params, float *host_grads);
int ccpprad_mlp_set_norm(ccpprad_mlp *m, const float *xmean, const float *xstd, const float *lo,
                         const float *hi, const float *ymean, const float *ystd, char *errmsg, int errlen);
/* Training on standardised batches (device pointers): X rows x 7 FP32 (fed
 * to the first GEMM as BF16 hi + lo parts), T rows x 256 FP32; loss_out
 * (device, 1 float) receives the batch MSE.  Adam (0.9, 0.999, 1e-8) on
 * FP32 master weights with BF16 copies. */
int ccpprad_mlp_forward_dev(ccpprad_mlp *m, const float *X, int rows, float *Y, void *stream,
                            char *errmsg, int errlen);
int ccpprad_mlp_grads_dev(ccpprad_mlp *m, const float *X, const float *T, int rows, float *loss_out,
                          void *stream, char *errmsg, int errlen);
int ccpprad_mlp_train_step_dev(ccpprad_mlp *m, const float *X, const float *T, int rows, float lr,
                               float *loss_out, void *stream, char *errmsg, int errlen);
/* Rebuild the fused-inference weight image after training or set_params. */
int ccpprad_mlp_finalize_training(ccpprad_mlp *m, char *errmsg, int errlen);
/* Fused inference over (nCol, nLay) CCPP-layout inputs -> tau, pfrac
 * (ngpt, nLay, nCol) FP32; oob (nLay, nCol) and colflag (nCol) mark rows
 * outside the training hull. */
int ccpprad_mlp_infer_dev(ccpprad_mlp *m, int nCol, int nLay, const double *p_lay, const double *p_lev,
                          const double *t_lay, const double *vmr_h2o, const double *vmr_o3,
                          const double *vmr_co2, const double *vmr_ch4, const double *vmr_n2o, float *tau,
                          float *pfrac, unsigned char *oob, unsigned char *colflag, void *stream,
                          char *errmsg, int errlen);
/* Select the LW gas optics of a context: the MLP (m != NULL) with table
 * fallback outside the hull, or the k-distribution table (m == NULL). */
int ccpprad_lw_set_gas_optics_mlp(ccpprad_ctx *ctx, ccpprad_mlp *m);

/* ---- observability (no counterpart in the reference, which has no timers) */

/* Per-kernel device timing with HIP events recorded on the launch stream.
 * Kernel ids: 0 gas optics, 1 LW solver, 2 SW solver, 3 cloud optics,
 * 4 McICA, 5 MLP inference, 6 gas optics of the emulator's fallback rows,
 * 7 packing of the aerosol optics.
 * enable != 0 starts recording (and clears). */
int ccpprad_timing_enable(ccpprad_ctx *ctx, int enable);
/* Synchronises the recorded events; total milliseconds and launch count. */
int ccpprad_timing_query(ccpprad_ctx *ctx, int kernel_id, double *total_ms, int *count);

/* Number of g-points / bands of a context. */
int ccpprad_ngpt(const ccpprad_ctx *ctx);
int ccpprad_nband(const ccpprad_ctx *ctx);

/* Library version string. */
const char *ccpprad_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CCPPRAD_H */
