! CCPP scheme GFS_rrtmgp_cloud_overlap on the MI355X: same module name and
! GFS_rrtmgp_cloud_overlap_run argument list as
! physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_cloud_overlap.F90:31-138.
! The decorrelation length and the overlap parameters are formed in FP64 on
! the GPU (ccpprad_cloud_overlap of include/ccpprad.h).  The reference's
! unsuffixed literals (the 0.001 m -> km factor, the Hogan fit's constants)
! take the precision of this file's default REAL, so a build with the
! physics' default-real-8 flags gets the same constants as the reference
! built with them.  p_lev, p_lay, tv_lay, con_g, con_rd, con_epsq and
! top_at_1 are unused, as in the reference.  Errors set errflg = 1.
module GFS_rrtmgp_cloud_overlap
  use iso_c_binding
  use ccpprad_c, only: ccpprad_cloud_overlap_args_t, ccpprad_cloud_overlap, f_msg, lgc
  implicit none
  integer, parameter, private :: kind_phys = c_double

  public GFS_rrtmgp_cloud_overlap_run
contains

  subroutine GFS_rrtmgp_cloud_overlap_run(nCol, nLev, yearlen, doSWrad, doLWrad,         &
       julian, lat, p_lev, p_lay, tv_lay, deltaZc, con_pi, con_g, con_rd, con_epsq,      &
       dcorr_con, idcor, iovr, iovr_dcorr, iovr_exp, iovr_exprand, idcor_con,            &
       idcor_hogan, idcor_oreopoulos, cld_frac, cld_cnv_frac, iovr_convcld, top_at_1,    &
       imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf, de_lgth, cloud_overlap_param,         &
       cnv_cloud_overlap_param, precip_overlap_param, errmsg, errflg)
    integer, intent(in) :: nCol, nLev, yearlen, imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf, &
         iovr, iovr_convcld, iovr_dcorr, iovr_exp, iovr_exprand, idcor, idcor_con,         &
         idcor_hogan, idcor_oreopoulos
    logical, intent(in) :: top_at_1, doSWrad, doLWrad
    real(kind_phys), intent(in) :: julian, con_pi, con_g, con_rd, con_epsq, dcorr_con
    real(kind_phys), dimension(:), intent(in), contiguous, target :: lat
    real(kind_phys), dimension(:,:), intent(in), contiguous, target :: tv_lay, p_lay,     &
         cld_frac, cld_cnv_frac
    real(kind_phys), dimension(:,:), intent(in), contiguous, target :: p_lev, deltaZc
    real(kind_phys), dimension(:), intent(out), contiguous, target :: de_lgth
    real(kind_phys), dimension(:,:), intent(out), contiguous, target :: cloud_overlap_param, &
         cnv_cloud_overlap_param, precip_overlap_param
    character(len=*), intent(out) :: errmsg
    integer, intent(out) :: errflg

    type(ccpprad_cloud_overlap_args_t) :: a
    character(kind=c_char), dimension(512) :: buf
    integer :: rc

    errmsg = ''
    errflg = 0
    if (.not. (doSWrad .or. doLWrad)) return
    if (size(lat) < nCol .or. size(de_lgth) < nCol .or.                                   &
         any([size(deltaZc), size(cld_frac), size(cld_cnv_frac), size(cloud_overlap_param), &
              size(cnv_cloud_overlap_param), size(precip_overlap_param)] < nCol * nLev)) then
       errmsg = 'GFS_rrtmgp_cloud_overlap_run: arrays smaller than (nCol, nLev)'
       errflg = 1
       return
    endif
    a%nCol = nCol; a%nLev = nLev; a%yearlen = yearlen
    a%doSWrad = lgc(doSWrad); a%doLWrad = lgc(doLWrad)
    a%literal_r8 = merge(1_c_int, 0_c_int, kind(1.0) == c_double)
    a%idcor = idcor; a%idcor_con = idcor_con; a%idcor_hogan = idcor_hogan
    a%idcor_oreopoulos = idcor_oreopoulos
    a%iovr = iovr; a%iovr_convcld = iovr_convcld; a%iovr_dcorr = iovr_dcorr
    a%iovr_exp = iovr_exp; a%iovr_exprand = iovr_exprand
    a%imfdeepcnv = imfdeepcnv; a%imfdeepcnv_gf = imfdeepcnv_gf; a%imfdeepcnv_samf = imfdeepcnv_samf
    a%julian = julian; a%con_pi = con_pi; a%dcorr_con = dcorr_con
    a%lat = c_loc(lat); a%deltaZc = c_loc(deltaZc)
    a%cld_frac = c_loc(cld_frac); a%cld_cnv_frac = c_loc(cld_cnv_frac)
    a%de_lgth = c_loc(de_lgth); a%cloud_overlap_param = c_loc(cloud_overlap_param)
    a%cnv_cloud_overlap_param = c_loc(cnv_cloud_overlap_param)
    a%precip_overlap_param = c_loc(precip_overlap_param)
    buf = c_null_char
    rc = ccpprad_cloud_overlap(a, buf, int(size(buf), c_int))
    if (rc /= 0) then
       call f_msg(buf, errmsg)
       errflg = 1
    endif
  end subroutine GFS_rrtmgp_cloud_overlap_run
end module GFS_rrtmgp_cloud_overlap
