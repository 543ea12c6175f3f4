! Drives the GFS_rrtmgp_cloud_overlap and GFS_rrtmgp_post scheme modules
! the way a CCPP cap calls them (every argument of the reference's list,
! Fortran arrays of the host's shapes).
! Usage: caller_driver overlap|post <input.bin> <output.bin>
! overlap input : int32 ncol, nlev, ncase, yearlen, iovr_dcorr, iovr_exp,
!                 iovr_exprand, idcor_con, idcor_hogan, idcor_oreopoulos,
!                 imfdeepcnv_gf, imfdeepcnv_samf; FP64 con_pi, dcorr_con;
!                 FP64 lat(ncol), deltaZc, cld_frac, cld_cnv_frac (ncol,nlev);
!                 per case int32 idcor, iovr, iovr_convcld, imfdeepcnv and
!                 FP64 julian.
! overlap output: per case int32 errflg; FP64 de_lgth(ncol), then
!                 cloud_overlap_param, cnv_cloud_overlap_param,
!                 precip_overlap_param (ncol,nlev).  The outputs are set to
!                 -7 before each call.
! post input    : int32 ncol, nlev, nday, iSFC, iTOA, doLWrad, doSWrad,
!                 do_lw_clrsky_hr, do_sw_clrsky_hr, save_diag, naerodp, nfxr;
!                 FP64 fhlwr, fhswr, raddt; FP64 fluxlwUP_allsky,
!                 fluxlwDOWN_allsky, fluxlwUP_clrsky, fluxlwDOWN_clrsky, the
!                 four SW ones, p_lev (ncol,nlev+1); sfc_alb_nir_dir,
!                 sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif, tsfa,
!                 coszen, coszdg (ncol); aerodp (ncol,naerodp); cldsa (ncol,5);
!                 cld_frac, cldtaulw, cldtausw (ncol,nlev); scmpsw (6,ncol);
!                 int32 mtopa, mbota (ncol,3), idxday(nday); then the inout
!                 fields as they stand before the call: fluxr (ncol,nfxr);
!                 htrlw, htrsw, htrlwu, htrlwc, htrswc (ncol,nlev); topflw
!                 (2,ncol), sfcflw (4,ncol), topfsw (3,ncol), sfcfsw (4,ncol);
!                 sfcdlw, sfculw, tsflw, nirbmdi, nirdfdi, visbmdi, visdfdi,
!                 nirbmui, nirdfui, visbmui, visdfui, sfcnsw, sfcdsw (ncol).
! post output   : int32 errflg; the inout fields in the same order.
program caller_driver
  use iso_c_binding, only: c_double
  use GFS_rrtmgp_cloud_overlap, only: GFS_rrtmgp_cloud_overlap_run
  use GFS_rrtmgp_post, only: GFS_rrtmgp_post_run
  use module_radlw_parameters, only: topflw_type, sfcflw_type
  use module_radsw_parameters, only: topfsw_type, sfcfsw_type, cmpfsw_type
  implicit none
  integer, parameter :: kp = c_double
  character(len=512) :: mode, fin, fout

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  if (trim(mode) == 'overlap') then
     call run_overlap()
  else
     call run_post()
  endif

contains

  subroutine run_overlap()
    integer :: hdr(12), flg(4), ncol, nlev, ncase, k, u, v, errflg
    real(kp) :: cst(2), julian
    real(kp), allocatable :: lat(:), dzc(:,:), cf(:,:), ccf(:,:), dl(:), cop(:,:), ccop(:,:), pop(:,:)
    real(kp), allocatable :: plev(:,:), play(:,:), tv(:,:)
    character(len=256) :: errmsg

    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) hdr
    read(u) cst
    ncol = hdr(1); nlev = hdr(2); ncase = hdr(3)
    allocate(lat(ncol), dzc(ncol,nlev), cf(ncol,nlev), ccf(ncol,nlev), dl(ncol), cop(ncol,nlev), &
         ccop(ncol,nlev), pop(ncol,nlev), plev(ncol,nlev+1), play(ncol,nlev), tv(ncol,nlev))
    read(u) lat, dzc, cf, ccf
    plev = 0._kp; play = 0._kp; tv = 0._kp
    open(newunit=v, file=trim(fout), access='stream', form='unformatted', status='replace')
    do k = 1, ncase
       read(u) flg
       read(u) julian
       dl = -7._kp; cop = -7._kp; ccop = -7._kp; pop = -7._kp
       call GFS_rrtmgp_cloud_overlap_run(ncol, nlev, hdr(4), .true., .true., julian, lat, plev, &
            play, tv, dzc, cst(1), 9.80665_kp, 287.05_kp, 1.0e-12_kp, cst(2), flg(1), flg(2),   &
            hdr(5), hdr(6), hdr(7), hdr(8), hdr(9), hdr(10), cf, ccf, flg(3), .false., flg(4),   &
            hdr(11), hdr(12), dl, cop, ccop, pop, errmsg, errflg)
       if (errflg /= 0) print *, 'overlap case ', k, ': ', trim(errmsg)
       write(v) errflg
       write(v) dl, cop, ccop, pop
    enddo
    close(u)
    close(v)
  end subroutine run_overlap

  subroutine run_post()
    integer :: hdr(12), ncol, nlev, nday, naer, nfxr, u, errflg
    real(kp) :: rl(3)
    real(kp), allocatable, dimension(:,:) :: lup, ldn, lupc, ldnc, sup, sdn, supc, sdnc, plev, aerodp, &
         cldsa, cf, taulw, tausw, fluxr, htrlw, htrsw, htrlwu, htrlwc, htrswc
    real(kp), allocatable, dimension(:) :: a1, a2, a3, a4, tsfa, cz, czd, sfcdlw, sfculw, tsflw,     &
         nirbmdi, nirdfdi, visbmdi, visdfdi, nirbmui, nirdfui, visbmui, visdfui, sfcnsw, sfcdsw
    integer, allocatable :: mtopa(:,:), mbota(:,:), idxday(:)
    type(cmpfsw_type), allocatable :: scmpsw(:)
    type(topflw_type), allocatable :: topflw(:)
    type(sfcflw_type), allocatable :: sfcflw(:)
    type(topfsw_type), allocatable :: topfsw(:)
    type(sfcfsw_type), allocatable :: sfcfsw(:)
    character(len=256) :: errmsg

    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) hdr
    read(u) rl
    ncol = hdr(1); nlev = hdr(2); nday = hdr(3); naer = hdr(11); nfxr = hdr(12)
    allocate(lup(ncol,nlev+1), ldn(ncol,nlev+1), lupc(ncol,nlev+1), ldnc(ncol,nlev+1), sup(ncol,nlev+1), &
         sdn(ncol,nlev+1), supc(ncol,nlev+1), sdnc(ncol,nlev+1), plev(ncol,nlev+1), aerodp(ncol,naer),  &
         cldsa(ncol,5), cf(ncol,nlev), taulw(ncol,nlev), tausw(ncol,nlev), fluxr(ncol,nfxr),           &
         htrlw(ncol,nlev), htrsw(ncol,nlev), htrlwu(ncol,nlev), htrlwc(ncol,nlev), htrswc(ncol,nlev))
    allocate(a1(ncol), a2(ncol), a3(ncol), a4(ncol), tsfa(ncol), cz(ncol), czd(ncol), sfcdlw(ncol),     &
         sfculw(ncol), tsflw(ncol), nirbmdi(ncol), nirdfdi(ncol), visbmdi(ncol), visdfdi(ncol),        &
         nirbmui(ncol), nirdfui(ncol), visbmui(ncol), visdfui(ncol), sfcnsw(ncol), sfcdsw(ncol))
    allocate(mtopa(ncol,3), mbota(ncol,3), idxday(max(nday,1)), scmpsw(ncol), topflw(ncol), sfcflw(ncol), &
         topfsw(ncol), sfcfsw(ncol))
    read(u) lup, ldn, lupc, ldnc, sup, sdn, supc, sdnc, plev
    read(u) a1, a2, a3, a4, tsfa, cz, czd
    read(u) aerodp, cldsa, cf, taulw, tausw, scmpsw
    read(u) mtopa, mbota
    if (nday > 0) read(u) idxday(1:nday)
    read(u) fluxr, htrlw, htrsw, htrlwu, htrlwc, htrswc, topflw, sfcflw, topfsw, sfcfsw
    read(u) sfcdlw, sfculw, tsflw, nirbmdi, nirdfdi, visbmdi, visdfdi, nirbmui, nirdfui, visbmui, &
         visdfui, sfcnsw, sfcdsw
    close(u)
    call GFS_rrtmgp_post_run(ncol, nlev, nday, hdr(4), hdr(5), idxday, hdr(6) /= 0, hdr(7) /= 0,  &
         hdr(8) /= 0, hdr(9) /= 0, hdr(10) /= 0, rl(1), rl(2), a1, a2, a3, a4, plev, tsfa, cz, czd, &
         ldnc, lup, ldn, lupc, sdnc, sup, sdn, supc, rl(3), aerodp, cldsa, mtopa, mbota, cf, taulw,  &
         tausw, scmpsw, fluxr, sfcdlw, sfculw, sfcflw, tsflw, htrlw, htrlwu, topflw, nirbmdi,        &
         nirdfdi, visbmdi, visdfdi, nirbmui, nirdfui, visbmui, visdfui, sfcnsw, sfcdsw, htrsw,      &
         sfcfsw, topfsw, htrswc, htrlwc, errmsg, errflg)
    if (errflg /= 0) print *, 'post: ', trim(errmsg)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) errflg
    write(u) fluxr, htrlw, htrsw, htrlwu, htrlwc, htrswc, topflw, sfcflw, topfsw, sfcfsw
    write(u) sfcdlw, sfculw, tsflw, nirbmdi, nirdfdi, visbmdi, visdfdi, nirbmui, nirdfui, visbmui, &
         visdfui, sfcnsw, sfcdsw
    close(u)
  end subroutine run_post
end program caller_driver
