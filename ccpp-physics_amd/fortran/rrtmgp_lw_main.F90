! CCPP scheme rrtmgp_lw_main on the MI355X: same module name, same
! rrtmgp_lw_main_init / rrtmgp_lw_main_run argument lists as
! physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:43-77 and :104-194, so the
! reference's rrtmgp_lw_main.meta binds unchanged.  The arithmetic runs in
! the gfx950 HIP library through ISO_C_BINDING (include/ccpprad.h).
!
! Deviations from the reference, by design:
!  - errors set errflg = 1 with the library's message (the reference prints
!    rte-rrtmgp errors and leaves errflg = 0, radiation_tools.F90:89-98);
!  - the gas / cloud files are read by the library: the reference's netCDF
!    files in the netCDF classic formats (netCDF-4 files are refused with a
!    message; `nccopy -k cdf5` converts them), or the package's binary table
!    files;
!  - doGP_lwscat together with use_LW_jacobian is rejected with an error
!    (rte_lw computes no Jacobian with two-stream optics);
!  - doGP_cldoptics_PADE without doGP_cldoptics_LUT reads the cloud file's
!    Pade coefficients (ccpprad_lw_init_files_pade).
module rrtmgp_lw_main
  use mpi_f08,   only: MPI_Comm
  use iso_c_binding
  use ccpprad_c
  implicit none
  integer, parameter, private :: kind_phys = c_double
  type(c_ptr), save, private :: lw_ctx = c_null_ptr

  public rrtmgp_lw_main_init, rrtmgp_lw_main_run, rrtmgp_lw_main_release
contains

  subroutine rrtmgp_lw_main_init(rrtmgp_root_dir, rrtmgp_lw_file_gas, rrtmgp_lw_file_clouds,&
       active_gases_array, doGP_cldoptics_PADE, doGP_cldoptics_LUT, doGP_sgs_pbl,           &
       doGP_sgs_cnv, nrghice, mpicomm, mpirank, mpiroot, nLay, rrtmgp_phys_blksz,           &
       errmsg, errflg)
    character(len=128), intent(in) :: rrtmgp_root_dir, rrtmgp_lw_file_clouds, rrtmgp_lw_file_gas
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
    if (doGP_cldoptics_LUT .or. doGP_cldoptics_PADE) cloud = trim(rrtmgp_root_dir)//trim(rrtmgp_lw_file_clouds)
    if (c_associated(lw_ctx)) call ccpprad_finalize(lw_ctx)
    buf = c_null_char
    if (doGP_cldoptics_PADE .and. .not. doGP_cldoptics_LUT) then
        rc = ccpprad_lw_init_files_pade(c_str(trim(rrtmgp_root_dir)//trim(rrtmgp_lw_file_gas)), c_str(cloud), &
            int(nrghice, c_int), c_loc(pnames), int(size(active_gases_array), c_int), int(nLay, c_int), &
            int(rrtmgp_phys_blksz, c_int), lw_ctx, buf, int(size(buf), c_int))
    else
        rc = ccpprad_lw_init_files(c_str(trim(rrtmgp_root_dir)//trim(rrtmgp_lw_file_gas)), c_str(cloud), &
            int(nrghice, c_int), c_loc(pnames), int(size(active_gases_array), c_int), int(nLay, c_int), &
            int(rrtmgp_phys_blksz, c_int), lw_ctx, buf, int(size(buf), c_int))
    endif
    if (rc /= 0) then
       call f_msg(buf, errmsg)
       errflg = 1
    endif
  end subroutine rrtmgp_lw_main_init

  subroutine rrtmgp_lw_main_run(doLWrad, doLWclrsky, top_at_1, doGP_lwscat,              &
       use_LW_jacobian, doGP_sgs_cnv, doGP_sgs_pbl, nCol, nLay, nGases,rrtmgp_phys_blksz,&
       nGauss_angles, icseed_lw, iovr, iovr_convcld, iovr_max, iovr_maxrand, iovr_rand,  &
       iovr_dcorr, iovr_exp, iovr_exprand, isubc_lw, semis, tsfg, p_lay, p_lev, t_lay,   &
       t_lev,  vmr_o2, vmr_h2o, vmr_o3, vmr_ch4, vmr_n2o, vmr_co2,                       &
       cld_frac, cld_lwp, cld_reliq, cld_iwp, cld_reice, cld_swp, cld_resnow,            &
       cld_rwp, cld_rerain, precip_frac, cld_cnv_lwp, cld_cnv_reliq, cld_cnv_iwp,        &
       cld_cnv_reice, cld_pbl_lwp, cld_pbl_reliq, cld_pbl_iwp, cld_pbl_reice,            &
       cloud_overlap_param, active_gases_array, aerlw_tau, aerlw_ssa, aerlw_g,           &
       fluxlwUP_allsky, fluxlwDOWN_allsky, fluxlwUP_clrsky, fluxlwDOWN_clrsky,           &
       fluxlwUP_jac, fluxlwUP_radtime, fluxlwDOWN_radtime, errmsg, errflg)
    logical, intent(in) :: doLWrad, doLWclrsky, top_at_1, use_LW_jacobian, doGP_sgs_pbl, &
         doGP_sgs_cnv, doGP_lwscat
    integer, intent(in) :: nCol, nLay, nGases, rrtmgp_phys_blksz, nGauss_angles, iovr,  &
         iovr_convcld, iovr_max, iovr_maxrand, iovr_rand, iovr_dcorr, iovr_exp,          &
         iovr_exprand, isubc_lw
    integer, intent(in), dimension(:), contiguous, target :: icseed_lw
    real(kind_phys), dimension(:), intent(in), contiguous, target :: semis, tsfg
    real(kind_phys), dimension(:,:), intent(in), contiguous, target :: p_lay, t_lay, p_lev, &
         t_lev, vmr_o2, vmr_h2o, vmr_o3, vmr_ch4, vmr_n2o, vmr_co2, cld_frac, cld_lwp,       &
         cld_reliq, cld_iwp, cld_reice, cld_swp, cld_resnow, cld_rwp, cld_rerain,           &
         precip_frac, cld_cnv_lwp, cld_cnv_reliq, cld_cnv_iwp, cld_cnv_reice, cld_pbl_lwp,  &
         cld_pbl_reliq, cld_pbl_iwp, cld_pbl_reice, cloud_overlap_param
    real(kind_phys), dimension(:,:,:), intent(in), contiguous, target :: aerlw_tau,        &
         aerlw_ssa, aerlw_g
    character(len=*), dimension(:), intent(in) :: active_gases_array
    real(kind_phys), dimension(:,:), intent(inout), contiguous, target :: fluxlwUP_jac,    &
         fluxlwUP_allsky, fluxlwDOWN_allsky, fluxlwUP_clrsky, fluxlwDOWN_clrsky,            &
         fluxlwUP_radtime, fluxlwDOWN_radtime
    character(len=*), intent(out) :: errmsg
    integer, intent(out) :: errflg

    type(ccpprad_lw_args_t) :: a
    character(kind=c_char), dimension(512) :: buf
    integer :: rc
    integer(c_int), dimension(:), allocatable, target :: seeds

    errmsg = ''
    errflg = 0
    if (.not. doLWrad) return
    if (.not. c_associated(lw_ctx)) then
       errmsg = 'rrtmgp_lw_main_run: rrtmgp_lw_main_init was not called'
       errflg = 1
       return
    endif
    a%doLWrad = lgc(doLWrad); a%doLWclrsky = lgc(doLWclrsky); a%top_at_1 = lgc(top_at_1)
    a%doGP_lwscat = lgc(doGP_lwscat); a%use_LW_jacobian = lgc(use_LW_jacobian)
    a%doGP_sgs_cnv = lgc(doGP_sgs_cnv); a%doGP_sgs_pbl = lgc(doGP_sgs_pbl)
    a%nCol = nCol; a%nLay = nLay; a%nGases = nGases; a%rrtmgp_phys_blksz = rrtmgp_phys_blksz
    a%nGauss_angles = nGauss_angles
    allocate(seeds(max(1, size(icseed_lw))))
    seeds = 0
    if (size(icseed_lw) > 0) seeds(1:size(icseed_lw)) = int(icseed_lw, c_int)
    a%icseed_lw = c_loc(seeds)
    a%iovr = iovr; a%iovr_convcld = iovr_convcld; a%iovr_max = iovr_max
    a%iovr_maxrand = iovr_maxrand; a%iovr_rand = iovr_rand; a%iovr_dcorr = iovr_dcorr
    a%iovr_exp = iovr_exp; a%iovr_exprand = iovr_exprand; a%isubc_lw = isubc_lw
    a%semis = c_loc(semis); a%tsfg = c_loc(tsfg)
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
    a%aerlw_tau = c_loc(aerlw_tau); a%aerlw_ssa = c_loc(aerlw_ssa); a%aerlw_g = c_loc(aerlw_g)
    a%fluxlwUP_allsky = c_loc(fluxlwUP_allsky); a%fluxlwDOWN_allsky = c_loc(fluxlwDOWN_allsky)
    a%fluxlwUP_clrsky = c_loc(fluxlwUP_clrsky); a%fluxlwDOWN_clrsky = c_loc(fluxlwDOWN_clrsky)
    a%fluxlwUP_jac = c_loc(fluxlwUP_jac); a%fluxlwUP_radtime = c_loc(fluxlwUP_radtime)
    a%fluxlwDOWN_radtime = c_loc(fluxlwDOWN_radtime)
    a%htrlw = c_null_ptr; a%htrlwc = c_null_ptr
    buf = c_null_char
    rc = ccpprad_lw_run(lw_ctx, a, buf, int(size(buf), c_int))
    if (rc /= 0) then
       call f_msg(buf, errmsg)
       errflg = 1
    endif
  end subroutine rrtmgp_lw_main_run

  ! Releases the device tables and work buffers.  No counterpart in the
  ! reference scheme, and not a CCPP phase (the name is none of _init, _run,
  ! _finalize, _timestep_init, _timestep_finalize, so the unchanged .meta
  ! needs no argument table for it): the host calls it where it tears
  ! radiation down, e.g. beside GFS_rrtmgp_setup_finalize.
  subroutine rrtmgp_lw_main_release()
    if (c_associated(lw_ctx)) call ccpprad_finalize(lw_ctx)
    lw_ctx = c_null_ptr
  end subroutine rrtmgp_lw_main_release
end module rrtmgp_lw_main
