! Cap-like driver: calls rrtmgp_lw_main_init/_run and rrtmgp_sw_main_init/_run
! the way a CCPP host-generated cap does (keyword-free, positional, with the
! reference's argument lists), on fields read from a stream file.
! Usage: ccpp_cap_driver <tables_dir/> <input.bin> <output.bin>
! Input : int32 ncol, nlay, nday, blksz; FP64 p_lay, p_lev, t_lay, t_lev,
!         tsfg, semis, vmr o2 h2o o3 ch4 n2o co2, cld_frac, cld_lwp,
!         cld_reliq, cld_iwp, cld_reice, coszen, alb nir_dir nir_dif
!         uvvis_dir uvvis_dif; int32 idx(nday).
! Output: int32 errflg_lw, errflg_sw; FP64 LW up/dn all sky, SW up/dn all
!         sky (ncol, nlay+1), scmpsw (6, ncol), cldtausw (ncol, nlay).
program ccpp_cap_driver
  use mpi_f08, only: MPI_Comm, MPI_COMM_WORLD
  use rrtmgp_lw_main
  use rrtmgp_sw_main
  use module_radsw_parameters, only: cmpfsw_type
  implicit none
  integer, parameter :: kp = 8
  character(len=128) :: root, fgas_lw, fcld_lw, fgas_sw, fcld_sw
  character(len=512) :: fin, fout
  character(len=3), dimension(6) :: gases = (/'h2o', 'co2', 'o3 ', 'n2o', 'ch4', 'o2 '/)
  integer :: ncol, nlay, nday, blksz, u, nrghice, errflg, errlw, errsw
  character(len=256) :: errmsg
  real(kp), allocatable, dimension(:,:) :: p_lay, p_lev, t_lay, t_lev, vo2, vh2o, vo3, vch4, vn2o, vco2, &
       cf, lwp, rel, iwp, rei, zl, swp, resnow, rwp, rerain, reliq0, reice0, ovl, &
       lup, ldn, lupc, ldnc, ljac, lupr, ldnr, sup, sdn, supc, sdnc, cldtau
  real(kp), allocatable, dimension(:) :: tsfg, semis, coszen, a1, a2, a3, a4
  real(kp), allocatable, dimension(:,:,:) :: aer
  integer, allocatable :: idx(:), seeds(:)
  type(cmpfsw_type), allocatable :: scmpsw(:)

  call get_command_argument(1, root)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  fgas_lw = 'lw_gas.tab'; fcld_lw = 'lw_cloud.tab'; fgas_sw = 'sw_gas.tab'; fcld_sw = 'sw_cloud.tab'

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, nlay, nday, blksz
  allocate(p_lay(ncol,nlay), p_lev(ncol,nlay+1), t_lay(ncol,nlay), t_lev(ncol,nlay+1), tsfg(ncol), semis(ncol))
  allocate(vo2(ncol,nlay), vh2o(ncol,nlay), vo3(ncol,nlay), vch4(ncol,nlay), vn2o(ncol,nlay), vco2(ncol,nlay))
  allocate(cf(ncol,nlay), lwp(ncol,nlay), rel(ncol,nlay), iwp(ncol,nlay), rei(ncol,nlay))
  allocate(coszen(ncol), a1(ncol), a2(ncol), a3(ncol), a4(ncol), idx(max(nday,1)), seeds(ncol))
  read(u) p_lay, p_lev, t_lay, t_lev, tsfg, semis, vo2, vh2o, vo3, vch4, vn2o, vco2
  read(u) cf, lwp, rel, iwp, rei, coszen, a1, a2, a3, a4
  if (nday > 0) read(u) idx(1:nday)
  close(u)
  allocate(zl(ncol,nlay), swp(ncol,nlay), resnow(ncol,nlay), rwp(ncol,nlay), rerain(ncol,nlay))
  allocate(reliq0(ncol,nlay), reice0(ncol,nlay), ovl(ncol,nlay))
  zl = 0; swp = 0; rwp = 0; resnow = 20; rerain = 100; reliq0 = 10; reice0 = 50; ovl = 0.5_kp
  seeds = 0
  allocate(lup(ncol,nlay+1), ldn(ncol,nlay+1), lupc(ncol,nlay+1), ldnc(ncol,nlay+1), ljac(ncol,nlay+1))
  allocate(lupr(ncol,nlay+1), ldnr(ncol,nlay+1), sup(ncol,nlay+1), sdn(ncol,nlay+1), supc(ncol,nlay+1))
  allocate(sdnc(ncol,nlay+1), cldtau(ncol,nlay), scmpsw(ncol))
  lup = 0; ldn = 0; lupc = 0; ldnc = 0; ljac = 0; lupr = 0; ldnr = 0; sup = 0; sdn = 0; supc = 0; sdnc = 0
  cldtau = 0
  scmpsw = cmpfsw_type(0, 0, 0, 0, 0, 0)

  root = trim(root)
  nrghice = 3
  call rrtmgp_lw_main_init(root, fgas_lw, fcld_lw, gases, .false., .true., .false., .false., nrghice, &
       MPI_COMM_WORLD, 0, 0, nlay, blksz, errmsg, errflg)
  if (errflg /= 0) print *, 'lw init: ', trim(errmsg)
  allocate(aer(ncol, nlay, 16))
  aer = 0
  call rrtmgp_lw_main_run(.true., .false., .false., .false., .false., .false., .false., ncol, nlay, 6, blksz, &
       3, seeds, 2, 2, 3, 2, 1, 4, 5, 6, 1, semis, tsfg, p_lay, p_lev, t_lay, t_lev, vo2, vh2o, vo3, vch4, vn2o, &
       vco2, cf, lwp, rel, iwp, rei, swp, resnow, rwp, rerain, zl, zl, reliq0, zl, reice0, zl, reliq0, zl,     &
       reice0, ovl, gases, aer, aer, aer, lup, ldn, lupc, ldnc, ljac, lupr, ldnr, errmsg, errlw)
  if (errlw /= 0) print *, 'lw run: ', trim(errmsg)
  deallocate(aer)
  allocate(aer(ncol, nlay, 14))
  aer = 0
  call rrtmgp_sw_main_init(root, fgas_sw, fcld_sw, gases, .false., .true., .false., .false., nrghice, &
       MPI_COMM_WORLD, 0, 0, nlay, blksz, errmsg, errflg)
  if (errflg /= 0) print *, 'sw init: ', trim(errmsg)
  call rrtmgp_sw_main_run(.true., .false., .false., .false., .false., ncol, nday, nlay, 6, blksz, idx, seeds, &
       2, 2, 3, 2, 1, 4, 5, 6, 1, 1, a1, a2, a3, a4, coszen, p_lay, p_lev, t_lay, t_lev, vo2, vh2o, vo3, vch4, &
       vn2o, vco2, cf, lwp, rel, iwp, rei, swp, resnow, rwp, rerain, zl, zl, reliq0, zl, reice0, zl, reliq0,  &
       zl, reice0, ovl, gases, aer, aer, aer, 1361.0_kp, scmpsw, sup, sdn, supc, sdnc, cldtau, errmsg, errsw)
  if (errsw /= 0) print *, 'sw run: ', trim(errmsg)
  call rrtmgp_lw_main_release()
  call rrtmgp_sw_main_release()

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) errlw, errsw
  write(u) lup, ldn, sup, sdn, scmpsw, cldtau
  close(u)
end program ccpp_cap_driver
