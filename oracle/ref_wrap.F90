! Test-harness wrapper (this repository's code, not the reference's): exposes
! the reference's own Fortran routines, compiled unmodified from
! physics/Radiation/mersenne_twister.f and physics/Radiation/radiation_tools.F90,
! through bind(C) so the oracle's restatements can be checked against them.
module ccpprad_ref_wrap
  use iso_c_binding
  use machine, only: kind_phys, kind_dbl_prec
  use mersenne_twister, only: random_setseed, random_number, random_stat
  use radiation_tools, only: cmp_tlev
  implicit none
contains
  ! Seed once, then n draws (the pattern of rrtmgp_lw_main.F90:452-466).
  subroutine ref_mt_draws(seed, n, out) bind(C, name="ref_mt_draws")
    integer(c_int), value :: seed, n
    real(c_double), intent(out) :: out(n)
    type(random_stat) :: stat
    real(kind_dbl_prec) :: h(n)
    call random_setseed(int(seed), stat)
    call random_number(h, stat)
    out = h
  end subroutine
  ! Seed once, then nchunk calls of chunk draws each.
  subroutine ref_mt_draws_chunked(seed, chunk, nchunk, out) bind(C, name="ref_mt_draws_chunked")
    integer(c_int), value :: seed, chunk, nchunk
    real(c_double), intent(out) :: out(chunk, nchunk)
    type(random_stat) :: stat
    real(kind_dbl_prec) :: h(chunk)
    integer :: i
    call random_setseed(int(seed), stat)
    do i = 1, nchunk
       call random_number(h, stat)
       out(:, i) = h
    enddo
  end subroutine
  subroutine ref_cmp_tlev(ncol, nlay, minp, p_lay, t_lay, p_lev, tsfc, t_lev) bind(C, name="ref_cmp_tlev")
    integer(c_int), value :: ncol, nlay
    real(c_double), value :: minp
    real(c_double), intent(in) :: p_lay(ncol, nlay), t_lay(ncol, nlay), p_lev(ncol, nlay+1), tsfc(ncol)
    real(c_double), intent(out) :: t_lev(ncol, nlay+1)
    call cmp_tlev(int(ncol), int(nlay), real(minp, kind_phys), p_lay, t_lay, p_lev, tsfc, t_lev)
  end subroutine
end module
