! The flux record types that GFS_rrtmgp_post_run and rrtmgp_sw_main_run take
! (module_radlw_parameters, Radiation/RRTMG/radlw_param.f:78-92, and
! module_radsw_parameters, Radiation/RRTMG/radsw_param.f:79-118): the same
! module and type names, components in the same order, all FP64.  A host
! model that builds the RRTMG parameter modules links those instead and
! leaves this file out; the scheme modules here only need the layout (a
! record of consecutive FP64 components, passed to the C library as an
! (ncomponents, nCol) array).
module module_radlw_parameters
  use iso_c_binding, only: c_double
  implicit none
  type, bind(C) :: topflw_type
     real(c_double) :: upfxc, upfx0
  end type topflw_type
  type, bind(C) :: sfcflw_type
     real(c_double) :: upfxc, upfx0, dnfxc, dnfx0
  end type sfcflw_type
end module module_radlw_parameters

module module_radsw_parameters
  use iso_c_binding, only: c_double
  implicit none
  type, bind(C) :: topfsw_type
     real(c_double) :: upfxc, dnfxc, upfx0
  end type topfsw_type
  type, bind(C) :: sfcfsw_type
     real(c_double) :: upfxc, dnfxc, upfx0, dnfx0
  end type sfcfsw_type
  type, bind(C) :: cmpfsw_type
     real(c_double) :: uvbfc, uvbf0, nirbm, nirdf, visbm, visdf
  end type cmpfsw_type
end module module_radsw_parameters
