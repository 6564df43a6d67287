module polyhomocont__config
!! Kinds, constants and status codes of the library (same values as the
!! reference, src/config.f90:11-23).

  use, intrinsic :: iso_fortran_env, only : real64

  implicit none

  private

  integer, parameter, public :: wp = real64
  real(wp), parameter, public :: pi = 3.14159265358979323846_wp

  integer, parameter, public :: status_ok = 0
  integer, parameter, public :: err_invalid_input = 1
  integer, parameter, public :: err_singular_matrix = 2
  integer, parameter, public :: err_numerical = 3
  !> The GPU path could not run (no device, or a HIP error). Sparse systems
  !> are never tracked on the CPU instead.
  integer, parameter, public :: err_device = -1
  !> Internal: a valid sparse system the GPU kernels cannot hold (more than
  !> 15 variables or an exponent above 255); solve tracks it on the host.
  !> Never returned by solve.
  integer, parameter, public :: err_capacity = -2

end module polyhomocont__config
