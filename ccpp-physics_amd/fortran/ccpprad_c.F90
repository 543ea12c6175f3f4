! ISO_C_BINDING interfaces to the C-ABI of include/ccpprad.h.
module ccpprad_c
  use iso_c_binding
  implicit none

  type, bind(C) :: ccpprad_lw_args_t
     integer(c_int) :: doLWrad, doLWclrsky, top_at_1, doGP_lwscat, use_LW_jacobian
     integer(c_int) :: doGP_sgs_cnv, doGP_sgs_pbl
     integer(c_int) :: nCol, nLay, nGases, rrtmgp_phys_blksz, nGauss_angles
     type(c_ptr)    :: icseed_lw
     integer(c_int) :: iovr, iovr_convcld, iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr
     integer(c_int) :: iovr_exp, iovr_exprand, isubc_lw
     type(c_ptr) :: semis, tsfg, p_lay, p_lev, t_lay, t_lev
     type(c_ptr) :: vmr_o2, vmr_h2o, vmr_o3, vmr_ch4, vmr_n2o, vmr_co2
     type(c_ptr) :: cld_frac, cld_lwp, cld_reliq, cld_iwp, cld_reice
     type(c_ptr) :: cld_swp, cld_resnow, cld_rwp, cld_rerain, precip_frac
     type(c_ptr) :: cld_cnv_lwp, cld_cnv_reliq, cld_cnv_iwp, cld_cnv_reice
     type(c_ptr) :: cld_pbl_lwp, cld_pbl_reliq, cld_pbl_iwp, cld_pbl_reice
     type(c_ptr) :: cloud_overlap_param, aerlw_tau, aerlw_ssa, aerlw_g
     type(c_ptr) :: fluxlwUP_allsky, fluxlwDOWN_allsky, fluxlwUP_clrsky, fluxlwDOWN_clrsky
     type(c_ptr) :: fluxlwUP_jac, fluxlwUP_radtime, fluxlwDOWN_radtime
     type(c_ptr) :: htrlw, htrlwc
  end type ccpprad_lw_args_t

  type, bind(C) :: ccpprad_sw_args_t
     integer(c_int) :: doSWrad, doSWclrsky, top_at_1, doGP_sgs_cnv, doGP_sgs_pbl
     integer(c_int) :: nCol, nDay, nLay, nGases, rrtmgp_phys_blksz
     type(c_ptr)    :: idx, icseed_sw
     integer(c_int) :: iovr, iovr_convcld, iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr
     integer(c_int) :: iovr_exp, iovr_exprand, isubc_sw, iSFC
     type(c_ptr) :: sfc_alb_nir_dir, sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif, coszen
     type(c_ptr) :: p_lay, p_lev, t_lay, t_lev
     type(c_ptr) :: vmr_o2, vmr_h2o, vmr_o3, vmr_ch4, vmr_n2o, vmr_co2
     type(c_ptr) :: cld_frac, cld_lwp, cld_reliq, cld_iwp, cld_reice
     type(c_ptr) :: cld_swp, cld_resnow, cld_rwp, cld_rerain, precip_frac
     type(c_ptr) :: cld_cnv_lwp, cld_cnv_reliq, cld_cnv_iwp, cld_cnv_reice
     type(c_ptr) :: cld_pbl_lwp, cld_pbl_reliq, cld_pbl_iwp, cld_pbl_reice
     type(c_ptr) :: cloud_overlap_param, aersw_tau, aersw_ssa, aersw_g
     real(c_double) :: solcon
     type(c_ptr) :: scmpsw, fluxswUP_allsky, fluxswDOWN_allsky, fluxswUP_clrsky, fluxswDOWN_clrsky
     type(c_ptr) :: cldtausw, htrsw, htrswc
  end type ccpprad_sw_args_t

  type, bind(C) :: ccpprad_post_args_t
     integer(c_int) :: nCol, nLev, nDay, iSFC, iTOA
     integer(c_int) :: doLWrad, doSWrad, save_diag, naerodp, nfxr
     real(c_double) :: fhlwr, fhswr, raddt
     type(c_ptr) :: fluxlwUP_allsky, fluxlwDOWN_allsky, fluxlwUP_clrsky, fluxlwDOWN_clrsky
     type(c_ptr) :: fluxswUP_allsky, fluxswDOWN_allsky, fluxswUP_clrsky, fluxswDOWN_clrsky
     type(c_ptr) :: scmpsw
     type(c_ptr) :: sfc_alb_nir_dir, sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif
     type(c_ptr) :: tsfa, coszen, coszdg, htrlw, p_lev
     type(c_ptr) :: cldsa, cldtaulw, cldtausw, aerodp
     type(c_ptr) :: mtopa, mbota
     type(c_ptr) :: topflw, sfcflw, sfcdlw, sfculw, tsflw, htrlwu
     type(c_ptr) :: topfsw, sfcfsw
     type(c_ptr) :: nirbmdi, nirdfdi, visbmdi, visdfdi, nirbmui, nirdfui, visbmui, visdfui
     type(c_ptr) :: sfcnsw, sfcdsw, fluxr
  end type ccpprad_post_args_t

  type, bind(C) :: ccpprad_post_hr_args_t
     integer(c_int) :: do_lw_clrsky_hr, do_sw_clrsky_hr
     type(c_ptr) :: idxday
     type(c_ptr) :: htrlw, htrlwc, htrsw, htrswc
  end type ccpprad_post_hr_args_t

  type, bind(C) :: ccpprad_cloud_overlap_args_t
     integer(c_int) :: nCol, nLev, yearlen, doSWrad, doLWrad, literal_r8
     integer(c_int) :: idcor, idcor_con, idcor_hogan, idcor_oreopoulos
     integer(c_int) :: iovr, iovr_convcld, iovr_dcorr, iovr_exp, iovr_exprand
     integer(c_int) :: imfdeepcnv, imfdeepcnv_gf, imfdeepcnv_samf
     real(c_double) :: julian, con_pi, dcorr_con
     type(c_ptr) :: lat, deltaZc, cld_frac, cld_cnv_frac
     type(c_ptr) :: de_lgth, cloud_overlap_param, cnv_cloud_overlap_param, precip_overlap_param
  end type ccpprad_cloud_overlap_args_t

  interface
     integer(c_int) function ccpprad_post(args, hr, errmsg, errlen) bind(C, name="ccpprad_post")
       import :: c_int, c_char, ccpprad_post_args_t, ccpprad_post_hr_args_t
       type(ccpprad_post_args_t), intent(in) :: args
       type(ccpprad_post_hr_args_t), intent(in) :: hr
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_post
     integer(c_int) function ccpprad_cloud_overlap(args, errmsg, errlen) bind(C, name="ccpprad_cloud_overlap")
       import :: c_int, c_char, ccpprad_cloud_overlap_args_t
       type(ccpprad_cloud_overlap_args_t), intent(in) :: args
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_cloud_overlap
     integer(c_int) function ccpprad_lw_init_files(gas_file, cloud_file, nrghice, gas_names, ngases, nLay, &
          blksz, ctx, errmsg, errlen) bind(C, name="ccpprad_lw_init_files")
       import :: c_int, c_ptr, c_char
       character(kind=c_char), dimension(*) :: gas_file, cloud_file
       integer(c_int), value :: nrghice
       type(c_ptr), value :: gas_names
       integer(c_int), value :: ngases, nLay, blksz
       type(c_ptr) :: ctx
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_lw_init_files
     integer(c_int) function ccpprad_sw_init_files(gas_file, cloud_file, nrghice, gas_names, ngases, nLay, &
          blksz, ctx, errmsg, errlen) bind(C, name="ccpprad_sw_init_files")
       import :: c_int, c_ptr, c_char
       character(kind=c_char), dimension(*) :: gas_file, cloud_file
       integer(c_int), value :: nrghice
       type(c_ptr), value :: gas_names
       integer(c_int), value :: ngases, nLay, blksz
       type(c_ptr) :: ctx
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_sw_init_files
     integer(c_int) function ccpprad_lw_init_files_pade(gas_file, cloud_file, nrghice, gas_names, ngases, nLay, &
          blksz, ctx, errmsg, errlen) bind(C, name="ccpprad_lw_init_files_pade")
       import :: c_int, c_ptr, c_char
       character(kind=c_char), dimension(*) :: gas_file, cloud_file
       integer(c_int), value :: nrghice
       type(c_ptr), value :: gas_names
       integer(c_int), value :: ngases, nLay, blksz
       type(c_ptr) :: ctx
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_lw_init_files_pade
     integer(c_int) function ccpprad_sw_init_files_pade(gas_file, cloud_file, nrghice, gas_names, ngases, nLay, &
          blksz, ctx, errmsg, errlen) bind(C, name="ccpprad_sw_init_files_pade")
       import :: c_int, c_ptr, c_char
       character(kind=c_char), dimension(*) :: gas_file, cloud_file
       integer(c_int), value :: nrghice
       type(c_ptr), value :: gas_names
       integer(c_int), value :: ngases, nLay, blksz
       type(c_ptr) :: ctx
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_sw_init_files_pade
     integer(c_int) function ccpprad_lw_run(ctx, args, errmsg, errlen) bind(C, name="ccpprad_lw_run")
       import :: c_int, c_ptr, c_char, ccpprad_lw_args_t
       type(c_ptr), value :: ctx
       type(ccpprad_lw_args_t), intent(in) :: args
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_lw_run
     integer(c_int) function ccpprad_sw_run(ctx, args, errmsg, errlen) bind(C, name="ccpprad_sw_run")
       import :: c_int, c_ptr, c_char, ccpprad_sw_args_t
       type(c_ptr), value :: ctx
       type(ccpprad_sw_args_t), intent(in) :: args
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_sw_run
     integer(c_int) function ccpprad_heating_rate(nCol, nLay, flux_up, flux_dn, p_lev, htr, errmsg, &
          errlen) bind(C, name="ccpprad_heating_rate")
       import :: c_int, c_ptr, c_char
       integer(c_int), value :: nCol, nLay
       type(c_ptr), value :: flux_up, flux_dn, p_lev, htr
       character(kind=c_char), dimension(*) :: errmsg
       integer(c_int), value :: errlen
     end function ccpprad_heating_rate
     subroutine ccpprad_finalize(ctx) bind(C, name="ccpprad_finalize")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine ccpprad_finalize
  end interface

contains

  ! Fortran string -> NUL-terminated C string
  function c_str(s) result(r)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s)+1) :: r
    r = trim(s)//c_null_char
  end function c_str

  ! C message buffer -> Fortran errmsg
  subroutine f_msg(buf, msg)
    character(kind=c_char), dimension(:), intent(in) :: buf
    character(len=*), intent(out) :: msg
    integer :: i
    msg = ''
    do i = 1, min(size(buf), len(msg))
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    enddo
  end subroutine f_msg

  integer(c_int) function lgc(x)
    logical, intent(in) :: x
    lgc = merge(1_c_int, 0_c_int, x)
  end function lgc
end module ccpprad_c
