! compute_heating_rate on the MI355X: the module and function that
! GFS_rrtmgp_post uses (physics/Interstitials/UFS_SCM_NEPTUNE/
! GFS_rrtmgp_post.F90:11, calls at :159-171 and :249-266), with the same
! argument list and the same error-string result, so GFS_rrtmgp_post links
! unchanged.  The heating rate is formed in FP64 on the GPU by the library
! (ccpprad_heating_rate of include/ccpprad.h):
!   heating_rate(i,k) = (flux_up(i,k+1) - flux_up(i,k) - flux_dn(i,k+1)
!                        + flux_dn(i,k)) * grav / (cp_dry * (plev(i,k+1) - plev(i,k)))
! in K/s.  Non-contiguous actual arguments (GFS_rrtmgp_post passes the
! daylit rows fluxswUP_allsky(idxday(1:nDay),:)) are copied by the compiler
! into contiguous temporaries, as the dummies are declared contiguous.
module mo_heating_rates
  use iso_c_binding
  use ccpprad_c, only: ccpprad_heating_rate, f_msg
  implicit none
  private
  integer, parameter :: wp = c_double

  public :: compute_heating_rate
contains

  function compute_heating_rate(flux_up, flux_dn, plev, heating_rate) result(error_msg)
    real(wp), dimension(:,:), intent(in), contiguous, target :: flux_up, flux_dn, plev
    real(wp), dimension(:,:), intent(out), contiguous, target :: heating_rate
    character(len=128) :: error_msg

    character(kind=c_char), dimension(512) :: buf
    integer :: ncol, nlay, rc

    error_msg = ''
    ncol = size(flux_up, 1)
    nlay = size(flux_up, 2) - 1
    if (any([size(flux_dn, 1), size(plev, 1), size(heating_rate, 1)] /= ncol)) then
       error_msg = 'heating_rate: flux and pressure arrays inconsistently sized (ncol)'
       return
    endif
    if (any([size(flux_dn, 2), size(plev, 2), size(heating_rate, 2) + 1] /= nlay + 1)) then
       error_msg = 'heating_rate: flux and pressure arrays inconsistently sized (nlay)'
       return
    endif
    if (ncol == 0 .or. nlay < 1) return
    buf = c_null_char
    rc = ccpprad_heating_rate(int(ncol, c_int), int(nlay, c_int), c_loc(flux_up), c_loc(flux_dn), &
         c_loc(plev), c_loc(heating_rate), buf, int(size(buf), c_int))
    if (rc /= 0) call f_msg(buf, error_msg)
  end function compute_heating_rate
end module mo_heating_rates
