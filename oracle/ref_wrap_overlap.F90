! Test-harness wrapper (this repository's code, not the reference's): calls
! the reference's own GFS_rrtmgp_cloud_overlap_run, compiled unmodified from
! physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_cloud_overlap.F90 and
! physics/Radiation/radiation_cloud_overlap.F90, through bind(C).  The
! arguments that routine does not use (p_lev, p_lay, tv_lay, con_g, con_rd,
! con_epsq, top_at_1) get placeholders.
module ccpprad_ref_wrap_overlap
  use iso_c_binding
  use machine, only: kind_phys
  use GFS_rrtmgp_cloud_overlap, only: GFS_rrtmgp_cloud_overlap_run
  implicit none
contains
  subroutine ref_cloud_overlap(ncol, nlev, yearlen, julian, con_pi, dcorr_con, idcor, idcor_con, &
       idcor_hogan, idcor_oreopoulos, iovr, iovr_convcld, iovr_dcorr, iovr_exp, iovr_exprand,     &
       imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf, lat, deltaZc, cld_frac, cld_cnv_frac,          &
       de_lgth, cloud_overlap_param, cnv_cloud_overlap_param, precip_overlap_param)              &
       bind(C, name="ref_cloud_overlap")
    integer(c_int), value :: ncol, nlev, yearlen, idcor, idcor_con, idcor_hogan, idcor_oreopoulos
    integer(c_int), value :: iovr, iovr_convcld, iovr_dcorr, iovr_exp, iovr_exprand
    integer(c_int), value :: imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf
    real(c_double), value :: julian, con_pi, dcorr_con
    real(c_double), intent(in) :: lat(ncol), deltaZc(ncol, nlev), cld_frac(ncol, nlev), cld_cnv_frac(ncol, nlev)
    real(c_double), intent(inout) :: de_lgth(ncol), cloud_overlap_param(ncol, nlev)
    real(c_double), intent(inout) :: cnv_cloud_overlap_param(ncol, nlev), precip_overlap_param(ncol, nlev)
    real(kind_phys) :: p_lev(ncol, nlev + 1), p_lay(ncol, nlev), tv_lay(ncol, nlev)
    character(len=512) :: errmsg
    integer :: errflg
    p_lev = 0.0_kind_phys
    p_lay = 0.0_kind_phys
    tv_lay = 0.0_kind_phys
    call GFS_rrtmgp_cloud_overlap_run(int(ncol), int(nlev), int(yearlen), .true., .true., julian, lat, &
         p_lev, p_lay, tv_lay, deltaZc, con_pi, 9.80665_kind_phys, 287.05_kind_phys, 1.0e-12_kind_phys, &
         dcorr_con, int(idcor), int(iovr), int(iovr_dcorr), int(iovr_exp), int(iovr_exprand),           &
         int(idcor_con), int(idcor_hogan), int(idcor_oreopoulos), cld_frac, cld_cnv_frac,               &
         int(iovr_convcld), .true., int(imfdeepcnv), int(imfdeepcnv_gf), int(imfdeepcnv_samf),         &
         de_lgth, cloud_overlap_param, cnv_cloud_overlap_param, precip_overlap_param, errmsg, errflg)
  end subroutine
end module
