module polyhomocont__solve
!! solve(sys, res, status, options): all isolated solutions of a square
!! polynomial system by the total-degree homotopy -- the reference's entry
!! point with its option and result types unchanged (src/solve.f90:57-215).
!!
!! Per path the work splits at the reference's per-path unit `endpoint`
!! (src/solve.f90:217-252):
!!   * sparse_system: all paths are tracked on the GPU in one call
!!     (polyhomocont__device); start points and random constants are made
!!     here so they are bit-identical to the reference's. A valid system
!!     the kernels cannot hold (more than 15 variables, an exponent above
!!     255) and any system after set_device(-1) is tracked on the host;
!!   * any other poly_system (user evaluators): the host tracker.
!! The host tracker runs the paths over OpenMP threads, each thread with
!! its own copy of the system (stateful user evaluators need one copy per
!! thread, as the reference notes for a parallel path loop); each path's
!! record is the same as in a serial loop.
!!
!! Classification then runs per path on all host threads (it depends on
!! nothing but the path), and deduplication runs in path order, because
!! `find` keeps the first match (src/solve.f90:379-400). The first match is
!! found among the stored points inside a window of a fixed linear key,
!! which holds every point within the match distance, so the result equals
!! the reference's linear scan; points with non-finite coordinates keep
!! that linear scan.

  use omp_lib, only : omp_get_max_threads, omp_get_thread_num
  use polyhomocont__config, only : wp, status_ok, err_device, err_capacity
  use polyhomocont__random, only : rng
  use polyhomocont__linalg, only : lu_factor, lu_solve, cond1, norm_inf
  use polyhomocont__system, only : poly_system
  use polyhomocont__sparse, only : sparse_system
  use polyhomocont__homotopy, only : total_degree_homotopy
  use polyhomocont__tracker, only : tracker_options, path_info,  &
    track_path, path_success
  use polyhomocont__endgame, only : endgame_options, endgame_result,  &
    cauchy_endgame, endgame_failed
  use polyhomocont__device, only : device_track, endpoint_records,  &
    host_requested

  implicit none

  private

  public :: solve_options, path_result, solve_result, solve
  public :: path_endgame_failed, classify_endpoints, host_endpoints

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

  !> Per-thread copies of a homotopy and its system.
  type :: workers
    class(poly_system), allocatable :: sys(:)
    type(total_degree_homotopy), allocatable :: hom(:)
  end type workers

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
      if (host_requested()) then
        call host_endpoints(hom, opts, start, rec)
      else
        call device_track(sys, hom, opts%tracker, opts%endgame,  &
          opts%use_endgame, start, rec, status)
        if (status == err_capacity) then
          rec = endpoint_records()
          call host_endpoints(hom, opts, start, rec)
          status = status_ok
        end if
        if (status /= status_ok) return
      end if
    class default
      call host_endpoints(hom, opts, start, rec)
    end select
    call classify_endpoints(hom, opts, rec, res)
  end subroutine solve

  subroutine make_workers(hom, count, w)
    !! `count` copies of hom, each bound to its own copy of hom's system.
    type(total_degree_homotopy), intent(in) :: hom
    integer, intent(in) :: count
    type(workers), intent(out), target :: w
    integer :: t
    allocate(w%sys(count), source=hom%sys)
    allocate(w%hom(count))
    do t = 1, count
      w%hom(t) = hom
      w%hom(t)%sys => w%sys(t)
    end do
  end subroutine make_workers

  integer function thread_count(work)
    integer, intent(in) :: work
    thread_count = max(1, min(omp_get_max_threads(), work))
  end function thread_count

  subroutine host_endpoints(hom, opts, start, rec)
    !! `endpoint` (src/solve.f90:217-252) for every path on the host
    !! tracker, paths spread over the OpenMP threads.
    type(total_degree_homotopy), intent(in) :: hom
    type(solve_options), intent(in) :: opts
    complex(wp), intent(in) :: start(:, :)
    type(endpoint_records), intent(inout) :: rec
    type(workers), target :: w
    integer :: k, np, nth, t

    np = size(start, 2)
    allocate(rec%status(np), rec%winding(np), rec%eg_status(np),  &
      rec%naccepted(np), rec%nrejected(np), rec%t(np), rec%eg_error(np),  &
      rec%eg_residual(np), rec%z(size(start, 1), np))
    if (np == 0) return
    nth = thread_count(np)
    call make_workers(hom, nth, w)
    !$omp parallel do num_threads(nth) schedule(dynamic, 1) private(t)
    do k = 1, np
      t = omp_get_thread_num() + 1
      call host_endpoint(w%hom(t), opts, start(:, k), rec, k)
    end do
    !$omp end parallel do
  end subroutine host_endpoints

  subroutine host_endpoint(hom, opts, z0, rec, k)
    !! One path: track_path to the end game's start, then cauchy_endgame;
    !! the record of path k (src/solve.f90:226-251).
    type(total_degree_homotopy), intent(inout) :: hom
    type(solve_options), intent(in) :: opts
    complex(wp), intent(in) :: z0(:)
    type(endpoint_records), intent(inout) :: rec
    integer, intent(in) :: k
    type(path_info) :: info
    type(endgame_result) :: eg
    complex(wp) :: z(size(z0))
    real(wp) :: t_end

    t_end = 0.0_wp
    if (opts%use_endgame) t_end = opts%endgame%t_start
    z = z0
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
  end subroutine host_endpoint

  subroutine classify_endpoints(hom, opts, rec, res)
    !! The host half of solve (src/solve.f90:181-213): path results from
    !! endpoint records, classification and refinement (per path, over the
    !! OpenMP threads), deduplication in path order, real solutions.
    type(total_degree_homotopy), intent(in) :: hom
    type(solve_options), intent(in) :: opts
    type(endpoint_records), intent(in) :: rec
    type(solve_result), intent(inout) :: res
    type(workers), target :: w
    complex(wp), allocatable :: sols(:, :), ssols(:, :)
    integer, allocatable :: hits(:), shits(:)
    logical, allocatable :: reg(:), sng(:)
    integer :: k, n, np, nth, t

    n = hom%n
    np = size(rec%status)
    res%n = n
    res%npaths = np
    allocate(res%paths(np), sols(n, np), ssols(n, np), hits(np), shits(np))
    allocate(reg(np), sng(np))
    if (np > 0) then
      nth = thread_count(np)
      call make_workers(hom, nth, w)
      !$omp parallel do num_threads(nth) schedule(dynamic, 4) private(t)
      do k = 1, np
        t = omp_get_thread_num() + 1
        call classify_path(w%hom(t), opts, rec, k, res%paths(k))
      end do
      !$omp end parallel do
    end if
    do k = 1, np
      associate (p => res%paths(k))
        reg(k) = .false.
        sng(k) = .false.
        if (p%status /= path_success) then
          res%nfailed = res%nfailed + 1
        else if (p%at_infinity) then
          res%ninfinite = res%ninfinite + 1
        else if (p%singular) then
          sng(k) = .true.
        else
          reg(k) = .true.
        end if
      end associate
    end do
    ! add_singular / add_solution (src/solve.f90:330-377) in path order
    call dedup_in_order(res%paths, sng, opts%singular_dedup_tol, .true.,  &
      ssols, res%nsingular, shits)
    call dedup_in_order(res%paths, reg, opts%dedup_tol, .false., sols,  &
      res%nsolutions, hits)
    res%nduplicates = sum(hits(1:res%nsolutions))
    res%solutions = sols(:, 1:res%nsolutions)
    res%singular_solutions = ssols(:, 1:res%nsingular)
    res%multiplicities = 1 + shits(1:res%nsingular)
    call collect_real(res, opts%real_tol)
  end subroutine classify_endpoints

  subroutine classify_path(hom, opts, rec, k, p)
    !! path_result k from its endpoint record (src/solve.f90:193-201).
    type(total_degree_homotopy), intent(inout) :: hom
    type(solve_options), intent(in) :: opts
    type(endpoint_records), intent(in) :: rec
    integer, intent(in) :: k
    type(path_result), intent(inout) :: p
    complex(wp) :: z(size(rec%z, 1))
    p%status = rec%status(k)
    p%t = rec%t(k)
    p%naccepted = rec%naccepted(k)
    p%nrejected = rec%nrejected(k)
    p%winding_number = rec%winding(k)
    p%endgame_status = rec%eg_status(k)
    p%endgame_error = rec%eg_error(k)
    z = rec%z(:, k)
    if (p%status /= path_success) then
      p%z = z
    else
      call classify(hom, z, opts, p)
    end if
  end subroutine classify_path

  subroutine dedup_in_order(paths, sel, tol, singular, pts, nstored, hits)
    !! For the selected paths in path order: the first stored point within
    !! tol max(1, ||x||) of x (first_match over the stored columns), else x
    !! is stored as a new column. Sets the path's solution_index (or
    !! singular_index); hits(j) counts the later paths matched to column j.
    !!
    !! The stored columns are the stored paths in path order, so the first
    !! matching column is the matching stored path of smallest index. Every
    !! point within the distance r of x lies within W r of x in the key
    !! sum_i a_i Re x_i + b_i Im x_i (a_i, b_i in [0.5, 1.5), W their sum),
    !! so only the paths whose keys fall in that window (plus a margin for
    !! the rounding of the keys) are compared, with the reference's test.
    type(path_result), intent(inout) :: paths(:)
    logical, intent(in) :: sel(:)
    real(wp), intent(in) :: tol
    logical, intent(in) :: singular
    complex(wp), intent(inout) :: pts(:, :)
    integer, intent(out) :: nstored
    integer, intent(out) :: hits(:)
    integer, allocatable :: cand(:), ord(:), col(:), nonfin(:)
    real(wp), allocatable :: key(:), bnd(:), skey(:), a(:), b(:)
    logical, allocatable :: fin(:)
    real(wp) :: r, wsum, half
    integer :: m, c, k, j, i, best, nfin, nnf, n

    nstored = 0
    hits = 0
    m = count(sel)
    if (m == 0) return
    n = size(pts, 1)
    allocate(cand(m), key(m), bnd(m), fin(m), col(size(paths)), nonfin(m))
    allocate(a(n), b(n))
    do i = 1, n
      a(i) = 0.5_wp + modulo(0.6180339887498949_wp*i, 1.0_wp)
      b(i) = 0.5_wp + modulo(0.7548776662466927_wp*i, 1.0_wp)
    end do
    wsum = sum(a) + sum(b)
    col = 0
    c = 0
    do k = 1, size(paths)
      if (.not. sel(k)) cycle
      c = c + 1
      cand(c) = k
      associate (x => paths(k)%x)
        key(c) = sum(a*real(x, wp) + b*aimag(x))
        bnd(c) = sum(a*abs(real(x, wp)) + b*abs(aimag(x)))
        fin(c) = bnd(c) <= huge(1.0_wp)  ! false for NaN and Inf
      end associate
    end do
    nfin = count(fin)
    allocate(ord(nfin), skey(nfin))
    ord = pack([(i, i = 1, m)], fin)
    call sort_by_key(key, ord)
    skey = key(ord)
    nnf = 0
    do c = 1, m
      k = cand(c)
      associate (x => paths(k)%x)
        r = tol*max(1.0_wp, norm_inf(x))
        best = 0
        if (fin(c)) then
          half = wsum*r*(1.0_wp + 1.0e-10_wp) + 1.0e-12_wp*(2.0_wp*bnd(c)  &
            + wsum*r)
          do i = lower_bound(skey, key(c) - half), nfin
            if (skey(i) > key(c) + half) exit
            j = cand(ord(i))
            if (col(j) == 0) cycle
            if (best /= 0 .and. j >= best) cycle
            if (norm_inf(x - paths(j)%x) <= r) best = j
          end do
          do i = 1, nnf
            j = nonfin(i)
            if (best /= 0 .and. j >= best) cycle
            if (norm_inf(x - paths(j)%x) <= r) best = j
          end do
          if (best /= 0) best = col(best)
        else
          do j = 1, nstored
            if (norm_inf(x - pts(:, j)) <= r) then
              best = j
              exit
            end if
          end do
        end if
        if (best /= 0) then
          hits(best) = hits(best) + 1
        else
          nstored = nstored + 1
          pts(:, nstored) = x
          best = nstored
          col(k) = nstored
          if (.not. fin(c)) then
            nnf = nnf + 1
            nonfin(nnf) = k
          end if
        end if
      end associate
      if (singular) then
        paths(k)%singular_index = best
      else
        paths(k)%solution_index = best
      end if
    end do
  end subroutine dedup_in_order

  pure integer function lower_bound(v, x)
    !! First index i with v(i) >= x (size(v) + 1 if none); v ascending.
    real(wp), intent(in) :: v(:), x
    integer :: lo, hi, mid
    lo = 1
    hi = size(v) + 1
    do while (lo < hi)
      mid = (lo + hi)/2
      if (v(mid) < x) then
        lo = mid + 1
      else
        hi = mid
      end if
    end do
    lower_bound = lo
  end function lower_bound

  subroutine sort_by_key(key, idx)
    !! Sorts idx so that key(idx) is ascending (bottom-up merge sort).
    real(wp), intent(in) :: key(:)
    integer, intent(inout) :: idx(:)
    integer, allocatable :: tmp(:)
    integer :: n, width, lo, mid, hi, i, j, k
    n = size(idx)
    allocate(tmp(n))
    width = 1
    do while (width < n)
      lo = 1
      do while (lo <= n)
        mid = min(lo + width, n + 1)
        hi = min(lo + 2*width, n + 1)
        i = lo
        j = mid
        k = lo
        do while (k < hi)
          if (i < mid .and. (j >= hi .or. key(idx(min(i, n))) <=  &
              key(idx(min(j, n))))) then
            tmp(k) = idx(i)
            i = i + 1
          else
            tmp(k) = idx(j)
            j = j + 1
          end if
          k = k + 1
        end do
        lo = hi
      end do
      idx = tmp
      width = 2*width
    end do
  end subroutine sort_by_key

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
