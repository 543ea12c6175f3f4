! CCPP scheme GFS_rrtmgp_post on the MI355X: same module name and
! GFS_rrtmgp_post_run argument list as
! physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_post.F90:35-389.  The
! heating rates (FP64, compute_heating_rate's formula; SW over the daylit
! rows idxday(1:nDay), zero elsewhere) and the flux bookkeeping (TOA and
! surface records, surface components, htrlwu, the time-weighted fluxr
! diagnostics) run on the GPU in one library call (ccpprad_post of
! include/ccpprad.h), which copies every intent(inout) output in and back so
! that what the reference leaves alone keeps its value.  cld_frac is unused,
! as in the reference.  The flux records are the types of
! module_radlw_parameters / module_radsw_parameters.  Errors set errflg = 1
! (the reference prints heating-rate errors and keeps errflg = 0).
module GFS_rrtmgp_post
  use iso_c_binding
  use ccpprad_c, only: ccpprad_post_args_t, ccpprad_post_hr_args_t, ccpprad_post, f_msg, lgc
  use module_radlw_parameters, only: topflw_type, sfcflw_type
  use module_radsw_parameters, only: topfsw_type, sfcfsw_type, cmpfsw_type
  implicit none
  integer, parameter, private :: kind_phys = c_double

  public GFS_rrtmgp_post_run
contains

  subroutine GFS_rrtmgp_post_run (nCol, nLev, nDay, iSFC, iTOA, idxday, doLWrad, doSWrad,  &
       do_lw_clrsky_hr, do_sw_clrsky_hr, save_diag, fhlwr, fhswr, sfc_alb_nir_dir,         &
       sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif, p_lev, tsfa, coszen, coszdg, &
       fluxlwDOWN_clrsky, fluxlwUP_allsky, fluxlwDOWN_allsky, fluxlwUP_clrsky,             &
       fluxswDOWN_clrsky, fluxswUP_allsky, fluxswDOWN_allsky, fluxswUP_clrsky,             &
       raddt, aerodp, cldsa, mtopa, mbota, cld_frac, cldtaulw, cldtausw, scmpsw, fluxr,    &
       sfcdlw, sfculw, sfcflw, tsflw, htrlw, htrlwu, topflw, nirbmdi, nirdfdi, visbmdi,    &
       visdfdi, nirbmui, nirdfui, visbmui, visdfui, sfcnsw, sfcdsw, htrsw, sfcfsw, topfsw, &
       htrswc, htrlwc, errmsg, errflg)
    integer, intent(in) :: nCol, nLev, nDay, iSFC, iTOA
    integer, intent(in), dimension(:) :: idxday
    integer, intent(in), dimension(:,:) :: mbota, mtopa
    logical, intent(in) :: doLWrad, doSWrad, do_lw_clrsky_hr, do_sw_clrsky_hr, save_diag
    real(kind_phys), intent(in) :: fhlwr, fhswr
    real(kind_phys), dimension(:), intent(in), contiguous, target :: tsfa, coszen, coszdg,   &
         sfc_alb_nir_dir, sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif
    real(kind_phys), dimension(:,:), intent(in), contiguous, target :: p_lev,              &
         fluxlwUP_allsky, fluxlwDOWN_allsky, fluxlwUP_clrsky, fluxlwDOWN_clrsky,           &
         fluxswUP_allsky, fluxswDOWN_allsky, fluxswUP_clrsky, fluxswDOWN_clrsky
    real(kind_phys), intent(in) :: raddt
    real(kind_phys), dimension(:,:), intent(in), contiguous, target :: aerodp, cldsa,      &
         cld_frac, cldtaulw, cldtausw
    type(cmpfsw_type), dimension(:), intent(in), contiguous, target :: scmpsw
    real(kind_phys), dimension(:,:), intent(inout), contiguous, target :: fluxr
    real(kind_phys), dimension(:), intent(inout), contiguous, target :: tsflw, sfcdlw,    &
         sfculw, nirbmdi, nirdfdi, visbmdi, visdfdi, nirbmui, nirdfui, visbmui, visdfui,   &
         sfcnsw, sfcdsw
    real(kind_phys), dimension(:,:), intent(inout), contiguous, target :: htrlw, htrsw, htrlwu
    type(sfcflw_type), dimension(:), intent(inout), contiguous, target :: sfcflw
    type(sfcfsw_type), dimension(:), intent(inout), contiguous, target :: sfcfsw
    type(topfsw_type), dimension(:), intent(inout), contiguous, target :: topfsw
    type(topflw_type), dimension(:), intent(inout), contiguous, target :: topflw
    character(len=*), intent(out) :: errmsg
    integer, intent(out) :: errflg
    real(kind_phys), dimension(:,:), intent(inout), optional, contiguous, target :: htrlwc, htrswc

    type(ccpprad_post_args_t) :: a
    type(ccpprad_post_hr_args_t) :: h
    character(kind=c_char), dimension(512) :: buf
    integer(c_int), dimension(:), allocatable, target :: iday
    integer(c_int), dimension(:,:), allocatable, target :: itop, ibot
    integer :: rc

    errmsg = ''
    errflg = 0
    if (.not. (doLWrad .or. doSWrad)) return
    if (size(p_lev, 1) /= nCol .or. size(p_lev, 2) /= nLev + 1) then
       errmsg = 'GFS_rrtmgp_post_run: p_lev is not (nCol, nLev+1)'
       errflg = 1
       return
    endif
    if ((doLWrad .and. do_lw_clrsky_hr .and. .not. present(htrlwc)) .or.                  &
        (doSWrad .and. do_sw_clrsky_hr .and. .not. present(htrswc))) then
       errmsg = 'GFS_rrtmgp_post_run: a clear-sky heating rate is requested but not passed'
       errflg = 1
       return
    endif
    if (doSWrad .and. nDay > size(idxday)) then
       errmsg = 'GFS_rrtmgp_post_run: nDay exceeds size(idxday)'
       errflg = 1
       return
    endif
    allocate(iday(max(1, nDay)), itop(max(1, size(mtopa, 1)), 3), ibot(max(1, size(mbota, 1)), 3))
    iday = 0
    if (nDay > 0) iday(1:nDay) = int(idxday(1:nDay), c_int)
    itop = 0
    ibot = 0
    if (size(mtopa) > 0) itop(1:size(mtopa, 1), 1:min(3, size(mtopa, 2))) = int(mtopa(:, 1:min(3, size(mtopa, 2))), c_int)
    if (size(mbota) > 0) ibot(1:size(mbota, 1), 1:min(3, size(mbota, 2))) = int(mbota(:, 1:min(3, size(mbota, 2))), c_int)

    a%nCol = nCol; a%nLev = nLev; a%nDay = nDay; a%iSFC = iSFC; a%iTOA = iTOA
    a%doLWrad = lgc(doLWrad); a%doSWrad = lgc(doSWrad); a%save_diag = lgc(save_diag)
    a%naerodp = size(aerodp, 2); a%nfxr = size(fluxr, 2)
    a%fhlwr = fhlwr; a%fhswr = fhswr; a%raddt = raddt
    a%fluxlwUP_allsky = c_loc(fluxlwUP_allsky); a%fluxlwDOWN_allsky = c_loc(fluxlwDOWN_allsky)
    a%fluxlwUP_clrsky = c_loc(fluxlwUP_clrsky); a%fluxlwDOWN_clrsky = c_loc(fluxlwDOWN_clrsky)
    a%fluxswUP_allsky = c_loc(fluxswUP_allsky); a%fluxswDOWN_allsky = c_loc(fluxswDOWN_allsky)
    a%fluxswUP_clrsky = c_loc(fluxswUP_clrsky); a%fluxswDOWN_clrsky = c_loc(fluxswDOWN_clrsky)
    a%scmpsw = c_loc(scmpsw)
    a%sfc_alb_nir_dir = c_loc(sfc_alb_nir_dir); a%sfc_alb_nir_dif = c_loc(sfc_alb_nir_dif)
    a%sfc_alb_uvvis_dir = c_loc(sfc_alb_uvvis_dir); a%sfc_alb_uvvis_dif = c_loc(sfc_alb_uvvis_dif)
    a%tsfa = c_loc(tsfa); a%coszen = c_loc(coszen); a%coszdg = c_loc(coszdg)
    a%htrlw = c_null_ptr; a%p_lev = c_loc(p_lev)
    a%cldsa = c_loc(cldsa); a%cldtaulw = c_loc(cldtaulw); a%cldtausw = c_loc(cldtausw)
    a%aerodp = c_loc(aerodp)
    a%mtopa = c_loc(itop); a%mbota = c_loc(ibot)
    a%topflw = c_loc(topflw); a%sfcflw = c_loc(sfcflw); a%sfcdlw = c_loc(sfcdlw)
    a%sfculw = c_loc(sfculw); a%tsflw = c_loc(tsflw); a%htrlwu = c_loc(htrlwu)
    a%topfsw = c_loc(topfsw); a%sfcfsw = c_loc(sfcfsw)
    a%nirbmdi = c_loc(nirbmdi); a%nirdfdi = c_loc(nirdfdi); a%visbmdi = c_loc(visbmdi)
    a%visdfdi = c_loc(visdfdi); a%nirbmui = c_loc(nirbmui); a%nirdfui = c_loc(nirdfui)
    a%visbmui = c_loc(visbmui); a%visdfui = c_loc(visdfui)
    a%sfcnsw = c_loc(sfcnsw); a%sfcdsw = c_loc(sfcdsw); a%fluxr = c_loc(fluxr)

    h%do_lw_clrsky_hr = lgc(do_lw_clrsky_hr); h%do_sw_clrsky_hr = lgc(do_sw_clrsky_hr)
    h%idxday = c_loc(iday)
    h%htrlw = c_loc(htrlw); h%htrsw = c_loc(htrsw)
    h%htrlwc = c_null_ptr; h%htrswc = c_null_ptr
    if (present(htrlwc)) h%htrlwc = c_loc(htrlwc)
    if (present(htrswc)) h%htrswc = c_loc(htrswc)
    buf = c_null_char
    rc = ccpprad_post(a, h, buf, int(size(buf), c_int))
    if (rc /= 0) then
       call f_msg(buf, errmsg)
       errflg = 1
    endif
  end subroutine GFS_rrtmgp_post_run
end module GFS_rrtmgp_post
