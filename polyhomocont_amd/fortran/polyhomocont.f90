module polyhomocont
!! Public API: the same names as the reference facade
!! (src/polyhomocont.f90:32-53), so `use polyhomocont` programs written for
!! the reference compile against this library unchanged.

  use polyhomocont__config, only : wp, status_ok, err_invalid_input,  &
    err_singular_matrix, err_numerical, err_device
  use polyhomocont__system, only : poly_system
  use polyhomocont__sparse, only : sparse_system
  use polyhomocont__tracker, only : tracker_options, path_success,  &
    path_failed_min_step, path_failed_max_steps
  use polyhomocont__solve, only : solve, solve_options, solve_result,  &
    path_result, path_endgame_failed
  use polyhomocont__endgame, only : endgame_options, endgame_converged,  &
    endgame_not_converged
  use polyhomocont__device, only : set_device

  implicit none

  private

  public :: wp, status_ok, err_invalid_input, err_singular_matrix
  public :: err_numerical, err_device
  public :: poly_system, sparse_system
  public :: solve, solve_options, solve_result, path_result
  public :: tracker_options, path_success, path_failed_min_step
  public :: path_failed_max_steps, path_endgame_failed
  public :: endgame_options, endgame_converged, endgame_not_converged
  public :: set_device

end module polyhomocont
