! Minimal stand-in for the mpi_f08 module, so the scheme modules keep the
! type(MPI_Comm) dummy of rrtmgp_lw_main_init / rrtmgp_sw_main_init
! (rrtmgp_lw_main.F90:64-65) without an MPI library.  The device tables are
! built per rank from the same files, so no broadcast is needed.
module mpi_f08
  implicit none
  type :: MPI_Comm
     integer :: MPI_VAL = 0
  end type MPI_Comm
  type(MPI_Comm), parameter :: MPI_COMM_WORLD = MPI_Comm(0)
end module mpi_f08
