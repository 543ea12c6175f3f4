! CCPP scheme rrtmgp_sw_main on the MI355X: same module name and
! rrtmgp_sw_main_init / rrtmgp_sw_main_run argument lists as
! physics/Radiation/RRTMGP/rrtmgp_sw_main.F90:34-65 and :87-189.  scmpsw is
! cmpfsw_type(nCol) of module_radsw_parameters (Radiation/RRTMG/
! radsw_param.f:111-118; radiation_flux_types.F90 for builds without the
! host's).  Errors set errflg = 1.

module rrtmgp_sw_main
  use mpi_f08,   only: MPI_Comm
  use iso_c_binding
  use ccpprad_c
  use module_radsw_parameters, only: cmpfsw_type
  implicit none
  integer, parameter, private :: kind_phys = c_double
  type(c_ptr), save, private :: sw_ctx = c_null_ptr

  public rrtmgp_sw_main_init, rrtmgp_sw_main_run, rrtmgp_sw_main_release
contains

  subroutine rrtmgp_sw_main_init(rrtmgp_root_dir, rrtmgp_sw_file_gas, rrtmgp_sw_file_clouds,&
       active_gases_array, doGP_cldoptics_PADE, doGP_cldoptics_LUT, doGP_sgs_pbl,           &
       doGP_sgs_cnv, nrghice, mpicomm, mpirank, mpiroot, nLay, rrtmgp_phys_blksz,           &
       errmsg, errflg)
    character(len=128), intent(in) :: rrtmgp_root_dir, rrtmgp_sw_file_clouds, rrtmgp_sw_file_gas
    character(len=*), dimension(:), intent(in) :: active_gases_array
    logical, intent(in) :: doGP_cldoptics_PADE, doGP_cldoptics_LUT, doGP_sgs_pbl, doGP_sgs_cnv
    integer, intent(inout) :: nrghice
    type(MPI_Comm), intent(in) :: mpicomm
    integer, intent(in) :: mpirank, mpiroot, rrtmgp_phys_blksz, nLay
    character(len=*), intent(out) :: errmsg
    integer, intent(out) :: errflg

    character(kind=c_char, len=64), dimension(size(active_gases_array)), target :: names
    type(c_ptr), dimension(size(active_gases_array)), target :: pnames
    character(kind=c_char), dimension(512) :: buf
    character(len=:), allocatable :: cloud
    integer :: i, rc

    errmsg = ''
    errflg = 0
    do i = 1, size(active_gases_array)
       names(i) = trim(active_gases_array(i))//c_null_char
       pnames(i) = c_loc(names(i))
    enddo
    cloud = ''
    if (doGP_cldoptics_LUT .or. doGP_cldoptics_PADE) cloud = trim(rrtmgp_root_dir)//trim(rrtmgp_sw_file_clouds)
    if (c_associated(sw_ctx)) call ccpprad_finalize(sw_ctx)
    buf = c_null_char
    if (doGP_cldoptics_PADE .and. .not. doGP_cldoptics_LUT) then
        rc = ccpprad_sw_init_files_pade(c_str(trim(rrtmgp_root_dir)//trim(rrtmgp_sw_file_gas)), c_str(cloud), &
            int(nrghice, c_int), c_loc(pnames), int(size(active_gases_array), c_int), int(nLay, c_int), &
            int(rrtmgp_phys_blksz, c_int), sw_ctx, buf, int(size(buf), c_int))
    else
        rc = ccpprad_sw_init_files(c_str(trim(rrtmgp_root_dir)//trim(rrtmgp_sw_file_gas)), c_str(cloud), &
            int(nrghice, c_int), c_loc(pnames), int(size(active_gases_array), c_int), int(nLay, c_int), &
            int(rrtmgp_phys_blksz, c_int), sw_ctx, buf, int(size(buf), c_int))
    endif
    if (rc /= 0) then
       call f_msg(buf, errmsg)
       errflg = 1
    endif
  end subroutine rrtmgp_sw_main_init

  subroutine rrtmgp_sw_main_run(doSWrad, doSWclrsky, top_at_1, doGP_sgs_cnv, doGP_sgs_pbl,  &
       nCol, nDay, nLay, nGases, rrtmgp_phys_blksz, idx, icseed_sw, iovr, iovr_convcld,     &
       iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr, iovr_exp, iovr_exprand, isubc_sw,     &
       iSFC, sfc_alb_nir_dir, sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif, coszen,&
       p_lay, p_lev, t_lay, t_lev, vmr_o2, vmr_h2o, vmr_o3, vmr_ch4, vmr_n2o, vmr_co2,      &
       cld_frac, cld_lwp, cld_reliq, cld_iwp, cld_reice, cld_swp, cld_resnow, cld_rwp,      &
       cld_rerain, precip_frac, cld_cnv_lwp, cld_cnv_reliq, cld_cnv_iwp, cld_cnv_reice,     &
       cld_pbl_lwp, cld_pbl_reliq, cld_pbl_iwp, cld_pbl_reice, cloud_overlap_param,         &
       active_gases_array, aersw_tau, aersw_ssa, aersw_g, solcon, scmpsw,                   &
       fluxswUP_allsky, fluxswDOWN_allsky, fluxswUP_clrsky, fluxswDOWN_clrsky, cldtausw,    &
       errmsg, errflg)
    logical, intent(in) :: doSWrad, doSWclrsky, top_at_1, doGP_sgs_pbl, doGP_sgs_cnv
    integer, intent(in) :: nCol, nDay, nLay, nGases, rrtmgp_phys_blksz, iovr, iovr_convcld, &
         iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr, iovr_exp, iovr_exprand, isubc_sw, iSFC
    integer, intent(in), dimension(:), contiguous, target :: idx, icseed_sw
    real(kind_phys), dimension(:), intent(in), contiguous, target :: sfc_alb_nir_dir,       &
         sfc_alb_nir_dif, sfc_alb_uvvis_dir, sfc_alb_uvvis_dif, coszen
    real(kind_phys), dimension(:,:), intent(in), contiguous, target :: p_lay, t_lay, p_lev, &
         t_lev, vmr_o2, vmr_h2o, vmr_o3, vmr_ch4, vmr_n2o, vmr_co2, cld_frac, cld_lwp,       &
         cld_reliq, cld_iwp, cld_reice, cld_swp, cld_resnow, cld_rwp, cld_rerain,           &
         precip_frac, cld_cnv_lwp, cld_cnv_reliq, cld_cnv_iwp, cld_cnv_reice, cld_pbl_lwp,  &
         cld_pbl_reliq, cld_pbl_iwp, cld_pbl_reice, cloud_overlap_param
    real(kind_phys), dimension(:,:,:), intent(in), contiguous, target :: aersw_tau,        &
         aersw_ssa, aersw_g
    character(len=*), dimension(:), intent(in) :: active_gases_array
    real(kind_phys), intent(in) :: solcon
    type(cmpfsw_type), dimension(:), intent(inout), contiguous, target :: scmpsw
    real(kind_phys), dimension(:,:), intent(inout), contiguous, target :: fluxswUP_allsky,  &
         fluxswDOWN_allsky, fluxswUP_clrsky, fluxswDOWN_clrsky, cldtausw
    character(len=*), intent(out) :: errmsg
    integer, intent(out) :: errflg

    type(ccpprad_sw_args_t) :: a
    character(kind=c_char), dimension(512) :: buf
    integer :: rc
    integer(c_int), dimension(:), allocatable, target :: cidx, seeds

    errmsg = ''
    errflg = 0
    if (.not. doSWrad) return
    if (.not. c_associated(sw_ctx)) then
       errmsg = 'rrtmgp_sw_main_run: rrtmgp_sw_main_init was not called'
       errflg = 1
       return
    endif
    allocate(cidx(max(1, nDay)), seeds(max(1, size(icseed_sw))))
    cidx = 1
    if (nDay > 0) cidx(1:nDay) = int(idx(1:nDay), c_int)
    seeds = 0
    if (size(icseed_sw) > 0) seeds(1:size(icseed_sw)) = int(icseed_sw, c_int)
    a%doSWrad = lgc(doSWrad); a%doSWclrsky = lgc(doSWclrsky); a%top_at_1 = lgc(top_at_1)
    a%doGP_sgs_cnv = lgc(doGP_sgs_cnv); a%doGP_sgs_pbl = lgc(doGP_sgs_pbl)
    a%nCol = nCol; a%nDay = nDay; a%nLay = nLay; a%nGases = nGases
    a%rrtmgp_phys_blksz = rrtmgp_phys_blksz
    a%idx = c_loc(cidx); a%icseed_sw = c_loc(seeds)
    a%iovr = iovr; a%iovr_convcld = iovr_convcld; a%iovr_max = iovr_max
    a%iovr_maxrand = iovr_maxrand; a%iovr_rand = iovr_rand; a%iovr_dcorr = iovr_dcorr
    a%iovr_exp = iovr_exp; a%iovr_exprand = iovr_exprand; a%isubc_sw = isubc_sw; a%iSFC = iSFC
    a%sfc_alb_nir_dir = c_loc(sfc_alb_nir_dir); a%sfc_alb_nir_dif = c_loc(sfc_alb_nir_dif)
    a%sfc_alb_uvvis_dir = c_loc(sfc_alb_uvvis_dir); a%sfc_alb_uvvis_dif = c_loc(sfc_alb_uvvis_dif)
    a%coszen = c_loc(coszen)
    a%p_lay = c_loc(p_lay); a%p_lev = c_loc(p_lev); a%t_lay = c_loc(t_lay); a%t_lev = c_loc(t_lev)
    a%vmr_o2 = c_loc(vmr_o2); a%vmr_h2o = c_loc(vmr_h2o); a%vmr_o3 = c_loc(vmr_o3)
    a%vmr_ch4 = c_loc(vmr_ch4); a%vmr_n2o = c_loc(vmr_n2o); a%vmr_co2 = c_loc(vmr_co2)
    a%cld_frac = c_loc(cld_frac); a%cld_lwp = c_loc(cld_lwp); a%cld_reliq = c_loc(cld_reliq)
    a%cld_iwp = c_loc(cld_iwp); a%cld_reice = c_loc(cld_reice); a%cld_swp = c_loc(cld_swp)
    a%cld_resnow = c_loc(cld_resnow); a%cld_rwp = c_loc(cld_rwp); a%cld_rerain = c_loc(cld_rerain)
    a%precip_frac = c_loc(precip_frac)
    a%cld_cnv_lwp = c_loc(cld_cnv_lwp); a%cld_cnv_reliq = c_loc(cld_cnv_reliq)
    a%cld_cnv_iwp = c_loc(cld_cnv_iwp); a%cld_cnv_reice = c_loc(cld_cnv_reice)
    a%cld_pbl_lwp = c_loc(cld_pbl_lwp); a%cld_pbl_reliq = c_loc(cld_pbl_reliq)
    a%cld_pbl_iwp = c_loc(cld_pbl_iwp); a%cld_pbl_reice = c_loc(cld_pbl_reice)
    a%cloud_overlap_param = c_loc(cloud_overlap_param)
    a%aersw_tau = c_loc(aersw_tau); a%aersw_ssa = c_loc(aersw_ssa); a%aersw_g = c_loc(aersw_g)
    a%solcon = solcon
    a%scmpsw = c_loc(scmpsw)
    a%fluxswUP_allsky = c_loc(fluxswUP_allsky); a%fluxswDOWN_allsky = c_loc(fluxswDOWN_allsky)
    a%fluxswUP_clrsky = c_loc(fluxswUP_clrsky); a%fluxswDOWN_clrsky = c_loc(fluxswDOWN_clrsky)
    a%cldtausw = c_loc(cldtausw)
    a%htrsw = c_null_ptr; a%htrswc = c_null_ptr
    buf = c_null_char
    rc = ccpprad_sw_run(sw_ctx, a, buf, int(size(buf), c_int))
    if (rc /= 0) then
       call f_msg(buf, errmsg)
       errflg = 1
    endif
  end subroutine rrtmgp_sw_main_run

  ! Releases the device tables and work buffers.  No counterpart in the
  ! reference scheme, and not a CCPP phase (the name is none of _init, _run,
  ! _finalize, _timestep_init, _timestep_finalize, so the unchanged .meta
  ! needs no argument table for it): the host calls it where it tears
  ! radiation down, e.g. beside GFS_rrtmgp_setup_finalize.
  subroutine rrtmgp_sw_main_release()
    if (c_associated(sw_ctx)) call ccpprad_finalize(sw_ctx)
    sw_ctx = c_null_ptr
  end subroutine rrtmgp_sw_main_release
end module rrtmgp_sw_main
