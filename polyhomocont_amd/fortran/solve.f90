module polyhomocont__solve
!! solve(sys, res, status, options): all isolated solutions of a square
!! polynomial system by the total-degree homotopy -- the reference's entry
!! point with its option and result types unchanged (src/solve.f90:57-215).
!!
!! Per path the work splits at the reference's per-path unit `endpoint`
!! (src/solve.f90:217-252):
!!   * sparse_system: all paths are tracked on the GPU in one call
!!     (polyhomocont__device); start points and random constants are made
!!     here so they are bit-identical to the reference's;
!!   * any other poly_system (user evaluators): the CPU tracker and end game.
!! Classification, refinement and deduplication then run on the host in
!! path order, as in the reference (src/solve.f90:254-426), because `find`
!! keeps the first match.

  use polyhomocont__config, only : wp, status_ok, err_device
  use polyhomocont__random, only : rng
  use polyhomocont__linalg, only : lu_factor, lu_solve, cond1, norm_inf
  use polyhomocont__system, only : poly_system
  use polyhomocont__sparse, only : sparse_system
  use polyhomocont__homotopy, only : total_degree_homotopy
  use polyhomocont__tracker, only : tracker_options, path_info,  &
    track_path, path_success
  use polyhomocont__endgame, only : endgame_options, endgame_result,  &
    cauchy_endgame, endgame_failed
  use polyhomocont__device, only : device_track, endpoint_records

  implicit none

  private

  public :: solve_options, path_result, solve_result, solve
  public :: path_endgame_failed, classify_endpoints

  integer, parameter :: path_endgame_failed = 3

  type :: solve_options
    type(tracker_options) :: tracker
    type(endgame_options) :: endgame
    logical :: use_endgame = .true.
    integer :: seed = 1
    integer :: refine_maxit = 5
    real(wp) :: infinity_tol = 1.0e-8_wp
    real(wp) :: singular_cond = 1.0e10_wp
    real(wp) :: real_tol = 1.0e-8_wp
    real(wp) :: dedup_tol = 1.0e-8_wp
    real(wp) :: singular_dedup_tol = 1.0e-6_wp
  end type solve_options

  type :: path_result
    integer :: status = path_success
    real(wp) :: t = 1.0_wp
    complex(wp), allocatable :: z(:)
    complex(wp), allocatable :: x(:)
    logical :: at_infinity = .false.
    logical :: singular = .false.
    logical :: is_real = .false.
    integer :: winding_number = 0
    integer :: endgame_status = 0
    real(wp) :: endgame_error = 0.0_wp
    real(wp) :: cond = 0.0_wp
    real(wp) :: residual = 0.0_wp
    integer :: solution_index = 0
    integer :: singular_index = 0
    integer :: naccepted = 0
    integer :: nrejected = 0
  end type path_result

  type :: solve_result
    integer :: n = 0
    integer :: npaths = 0
    type(path_result), allocatable :: paths(:)
    integer :: nsolutions = 0
    complex(wp), allocatable :: solutions(:, :)
    integer :: nreal = 0
    real(wp), allocatable :: real_solutions(:, :)
    integer :: nsingular = 0
    complex(wp), allocatable :: singular_solutions(:, :)
    integer, allocatable :: multiplicities(:)
    integer :: ninfinite = 0
    integer :: nfailed = 0
    integer :: nduplicates = 0
  end type solve_result

contains

  subroutine solve(sys, res, status, options)
    !! status: err_invalid_input for a non-square system or a degree < 1
    !! (path failures are reported in res), err_device if the GPU path for
    !! a sparse system cannot run.
    class(poly_system), intent(inout), target :: sys
    type(solve_result), intent(out) :: res
    integer, intent(out) :: status
    type(solve_options), intent(in), optional :: options
    type(solve_options) :: opts
    type(total_degree_homotopy) :: hom
    type(rng) :: gen
    type(endpoint_records) :: rec
    complex(wp), allocatable :: start(:, :)
    integer :: k

    if (present(options)) opts = options
    call gen%seed(opts%seed)
    call hom%init(sys, gen, status)
    if (status /= status_ok) return

    allocate(start(hom%n + 1, hom%nstart))
    do k = 1, hom%nstart
      call hom%start_solution(k, start(:, k))
    end do

    select type (sys)
    type is (sparse_system)
      call device_track(sys, hom, opts%tracker, opts%endgame,  &
        opts%use_endgame, start, rec, status)
      if (status /= status_ok) return
    class default
      call cpu_endpoints(hom, opts, start, rec)
    end select
    call classify_endpoints(hom, opts, rec, res)
  end subroutine solve

  subroutine cpu_endpoints(hom, opts, start, rec)
    !! `endpoint` for every path on the CPU (user-supplied evaluators).
    type(total_degree_homotopy), intent(inout) :: hom
    type(solve_options), intent(in) :: opts
    complex(wp), intent(in) :: start(:, :)
    type(endpoint_records), intent(inout) :: rec
    type(path_info) :: info
    type(endgame_result) :: eg
    complex(wp) :: z(size(start, 1))
    real(wp) :: t_end
    integer :: k, np

    np = size(start, 2)
    allocate(rec%status(np), rec%winding(np), rec%eg_status(np),  &
      rec%naccepted(np), rec%nrejected(np), rec%t(np), rec%eg_error(np),  &
      rec%eg_residual(np), rec%z(size(start, 1), np))
    t_end = 0.0_wp
    if (opts%use_endgame) t_end = opts%endgame%t_start
    do k = 1, np
      z = start(:, k)
      call track_path(hom, z, 1.0_wp, t_end, opts%tracker, info)
      rec%status(k) = info%status
      rec%t(k) = real(info%t, wp)
      rec%naccepted(k) = info%naccepted
      rec%nrejected(k) = info%nrejected
      rec%winding(k) = 0
      rec%eg_status(k) = 0
      rec%eg_error(k) = 0.0_wp
      rec%eg_residual(k) = 0.0_wp
      if (info%status == path_success .and. opts%use_endgame) then
        call cauchy_endgame(hom, z, opts%endgame, opts%tracker, eg)
        rec%naccepted(k) = rec%naccepted(k) + eg%naccepted
        rec%nrejected(k) = rec%nrejected(k) + eg%nrejected
        if (eg%status == endgame_failed) then
          rec%status(k) = path_endgame_failed
        else
          z = eg%z
          rec%t(k) = 0.0_wp
          rec%winding(k) = eg%winding_number
          rec%eg_status(k) = eg%status
          rec%eg_error(k) = eg%error
          rec%eg_residual(k) = eg%residual
        end if
      end if
      rec%z(:, k) = z
    end do
  end subroutine cpu_endpoints

  subroutine classify_endpoints(hom, opts, rec, res)
    !! The host half of solve (src/solve.f90:181-213): path results from
    !! endpoint records, classification, refinement, deduplication in path
    !! order, real solutions.
    type(total_degree_homotopy), intent(inout) :: hom
    type(solve_options), intent(in) :: opts
    type(endpoint_records), intent(in) :: rec
    type(solve_result), intent(inout) :: res
    complex(wp), allocatable :: sols(:, :), ssols(:, :), z(:)
    integer, allocatable :: mult(:)
    integer :: k, n, np

    n = hom%n
    np = size(rec%status)
    res%n = n
    res%npaths = np
    allocate(res%paths(np), sols(n, np), ssols(n, np), mult(np))
    allocate(z(n + 1))
    do k = 1, np
      associate (p => res%paths(k))
        p%status = rec%status(k)
        p%t = rec%t(k)
        p%naccepted = rec%naccepted(k)
        p%nrejected = rec%nrejected(k)
        p%winding_number = rec%winding(k)
        p%endgame_status = rec%eg_status(k)
        p%endgame_error = rec%eg_error(k)
        z = rec%z(:, k)
        if (p%status /= path_success) then
          res%nfailed = res%nfailed + 1
          p%z = z
        else
          call classify(hom, z, opts, p)
          if (p%at_infinity) then
            res%ninfinite = res%ninfinite + 1
          else if (p%singular) then
            call add_singular(p, ssols, mult, res, opts%singular_dedup_tol)
          else
            call add_solution(p, sols, res, opts%dedup_tol)
          end if
        end if
      end associate
    end do
    res%solutions = sols(:, 1:res%nsolutions)
    res%singular_solutions = ssols(:, 1:res%nsingular)
    res%multiplicities = mult(1:res%nsingular)
    call collect_real(res, opts%real_tol)
  end subroutine classify_endpoints

  subroutine refine(hom, z, maxit)
    !! Newton on H(z, 0) = 0 until the update stalls or reaches rounding
    !! level (4 eps ||z||).
    type(total_degree_homotopy), intent(inout) :: hom
    complex(wp), intent(inout) :: z(:)
    integer, intent(in) :: maxit
    complex(wp) :: h(size(z)), hz(size(z), size(z)), ht(size(z))
    integer :: piv(size(z)), status, it
    real(wp) :: step, last

    last = huge(1.0_wp)
    do it = 1, maxit
      call hom%evaluate(z, (0.0_wp, 0.0_wp), h, hz, ht)
      call lu_factor(hz, piv, status)
      if (status /= status_ok) return
      h = -h
      call lu_solve(hz, piv, h)
      step = norm_inf(h)
      if (step > last) return
      z = z + h
      if (step <= 4.0_wp*epsilon(1.0_wp)*norm_inf(z)) return
      last = step
    end do
  end subroutine refine

  subroutine classify(hom, z, opts, p)
    !! Singular if the winding number exceeds 1 or cond > singular_cond;
    !! regular endpoints are refined; at infinity if |w| <= max(infinity_tol,
    !! 10 min(endgame error, 1e-4)) ||z||; real if ||Im x|| <= real_tol
    !! max(1, ||x||).
    type(total_degree_homotopy), intent(inout) :: hom
    complex(wp), intent(inout) :: z(:)
    type(solve_options), intent(in) :: opts
    type(path_result), intent(inout) :: p
    complex(wp) :: h(size(z)), hz(size(z), size(z)), ht(size(z))
    real(wp) :: zmax
    integer :: n

    n = size(z) - 1
    call hom%evaluate(z, (0.0_wp, 0.0_wp), h, hz, ht)
    p%cond = cond1(hz)
    p%singular = p%winding_number > 1 .or. p%cond > opts%singular_cond
    if (.not. p%singular) then
      call refine(hom, z, opts%refine_maxit)
      call hom%evaluate(z, (0.0_wp, 0.0_wp), h, hz, ht)
      p%cond = cond1(hz)
    end if
    p%z = z
    zmax = norm_inf(z)
    p%residual = maxval(abs(h(1:n))/zmax**hom%deg)
    p%at_infinity = abs(z(n + 1)) <= max(opts%infinity_tol,  &
      10.0_wp*min(p%endgame_error, 1.0e-4_wp))*zmax
    if (.not. p%at_infinity) then
      p%x = z(1:n)/z(n + 1)
      p%is_real = maxval(abs(aimag(p%x))) <= opts%real_tol  &
        *max(1.0_wp, norm_inf(p%x))
    end if
  end subroutine classify

  subroutine add_solution(p, sols, res, tol)
    type(path_result), intent(inout) :: p
    complex(wp), intent(inout) :: sols(:, :)
    type(solve_result), intent(inout) :: res
    real(wp), intent(in) :: tol
    integer :: j
    j = first_match(p%x, sols(:, 1:res%nsolutions), tol)
    if (j > 0) then
      res%nduplicates = res%nduplicates + 1
    else
      res%nsolutions = res%nsolutions + 1
      sols(:, res%nsolutions) = p%x
      j = res%nsolutions
    end if
    p%solution_index = j
  end subroutine add_solution

  subroutine add_singular(p, ssols, mult, res, tol)
    type(path_result), intent(inout) :: p
    complex(wp), intent(inout) :: ssols(:, :)
    integer, intent(inout) :: mult(:)
    type(solve_result), intent(inout) :: res
    real(wp), intent(in) :: tol
    integer :: j
    j = first_match(p%x, ssols(:, 1:res%nsingular), tol)
    if (j > 0) then
      mult(j) = mult(j) + 1
    else
      res%nsingular = res%nsingular + 1
      ssols(:, res%nsingular) = p%x
      mult(res%nsingular) = 1
      j = res%nsingular
    end if
    p%singular_index = j
  end subroutine add_singular

  function first_match(x, pts, tol) result(idx)
    !! First column of pts within relative distance tol of x, else 0.
    complex(wp), intent(in) :: x(:), pts(:, :)
    real(wp), intent(in) :: tol
    integer :: idx
    real(wp) :: scale
    scale = max(1.0_wp, norm_inf(x))
    do idx = 1, size(pts, 2)
      if (norm_inf(x - pts(:, idx)) <= tol*scale) return
    end do
    idx = 0
  end function first_match

  subroutine collect_real(res, tol)
    type(solve_result), intent(inout) :: res
    real(wp), intent(in) :: tol
    logical :: flags(res%nsolutions)
    integer :: j, m
    do j = 1, res%nsolutions
      flags(j) = maxval(abs(aimag(res%solutions(:, j)))) <= tol  &
        *max(1.0_wp, norm_inf(res%solutions(:, j)))
    end do
    res%nreal = count(flags)
    allocate(res%real_solutions(res%n, res%nreal))
    m = 0
    do j = 1, res%nsolutions
      if (flags(j)) then
        m = m + 1
        res%real_solutions(:, m) = real(res%solutions(:, j), wp)
      end if
    end do
  end subroutine collect_real

end module polyhomocont__solve
