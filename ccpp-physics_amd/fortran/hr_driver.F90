! Drives mo_heating_rates the way GFS_rrtmgp_post_run does
! (physics/Interstitials/UFS_SCM_NEPTUNE/GFS_rrtmgp_post.F90:159-171 for
! the LW heating rates over every column, :249-266 for the SW ones over the
! daylit rows idxday(1:nDay), scattered back into zeroed arrays), then makes
! one call with inconsistent shapes, whose error string must be nonblank.
! Usage: hr_driver <input.bin> <output.bin>
! Input : int32 ncol, nlev, nday; FP64 lw_up, lw_dn, sw_up, sw_dn, p_lev
!         (ncol, nlev+1); int32 idxday(nday).
! Output: int32 three flags (1 = nonblank error string) for the LW, SW and
!         mis-shaped calls; FP64 htrlw, htrsw (ncol, nlev) in K/s.
program hr_driver
  use mo_heating_rates, only: compute_heating_rate
  implicit none
  integer, parameter :: kp = 8
  character(len=512) :: fin, fout
  character(len=128) :: e_lw, e_sw, e_bad
  integer :: ncol, nlev, nday, u
  real(kp), allocatable, dimension(:,:) :: lup, ldn, sup, sdn, plev, htrlw, htrsw, tend, bad
  integer, allocatable :: idxday(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, nlev, nday
  allocate(lup(ncol,nlev+1), ldn(ncol,nlev+1), sup(ncol,nlev+1), sdn(ncol,nlev+1), plev(ncol,nlev+1))
  allocate(htrlw(ncol,nlev), htrsw(ncol,nlev), tend(max(nday,1),nlev), idxday(max(nday,1)), bad(ncol+1,nlev))
  read(u) lup, ldn, sup, sdn, plev
  if (nday > 0) read(u) idxday(1:nday)
  close(u)

  e_lw = compute_heating_rate(lup, ldn, plev, htrlw)
  htrsw(:,:) = 0._kp
  e_sw = ''
  if (nday > 0) then
     e_sw = compute_heating_rate(sup(idxday(1:nday),:), sdn(idxday(1:nday),:), plev(idxday(1:nday),:), &
          tend(1:nday,:))
     htrsw(idxday(1:nday),:) = tend(1:nday,:)
  endif
  e_bad = compute_heating_rate(lup, ldn, plev, bad)
  if (len_trim(e_lw) > 0) print *, 'lw: ', trim(e_lw)
  if (len_trim(e_sw) > 0) print *, 'sw: ', trim(e_sw)
  print *, 'mis-shaped: ', trim(e_bad)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) merge(1, 0, len_trim(e_lw) > 0), merge(1, 0, len_trim(e_sw) > 0), merge(1, 0, len_trim(e_bad) > 0)
  write(u) htrlw, htrsw
  close(u)
end program hr_driver
