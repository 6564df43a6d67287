module polyhomocont__capi
!! C-callable layer of the Fortran library (include/polyhomocont.h): `solve`
!! and its pieces over flat arrays, for C and Python callers (tests,
!! benchmarks, the serving front end). Fortran callers use `polyhomocont`.

  use, intrinsic :: iso_c_binding
  use omp_lib, only : omp_set_num_threads, omp_get_max_threads
  use polyhomocont__config, only : wp, status_ok
  use polyhomocont__random, only : rng
  use polyhomocont__system, only : poly_system
  use polyhomocont__sparse, only : sparse_system
  use polyhomocont__homotopy, only : total_degree_homotopy
  use polyhomocont__solve, only : solve, solve_options, solve_result,  &
    classify_endpoints
  use polyhomocont__device, only : c_solve_options, c_endpoints,  &
    endpoint_records, set_device

  implicit none

  private

  type, bind(C) :: c_user_system
    integer(c_int32_t) :: n, neq
    type(c_ptr) :: nterms, coefs, exps
  end type c_user_system

  type, bind(C) :: c_solve_summary
    integer(c_int32_t) :: n, npaths, nsolutions, nreal, nsingular,  &
      ninfinite, nfailed, nduplicates
    type(c_ptr) :: solutions, real_solutions, singular_solutions,  &
      multiplicities
  end type c_solve_summary

  type, bind(C) :: c_path_results
    type(c_ptr) :: status, t, z, x, at_infinity, singular, is_real,  &
      winding, eg_status, eg_error, cond, residual, solution_index,  &
      singular_index, naccepted, nrejected
  end type c_path_results

  !> A system with a user-supplied evaluator: only the affine F(x) and its
  !> Jacobian (here those of a sparse system), so solve homogenizes through
  !> the default poly_system evaluate_homogeneous and tracks on the host,
  !> as for the user types of the reference's tests (test/solve.f90:16-57).
  !> `calls` makes it stateful, as user evaluators may be.
  type, extends(poly_system) :: affine_user_system
    type(sparse_system) :: inner
    integer :: calls = 0
  contains
    procedure :: nvars => au_nvars
    procedure :: degrees => au_degrees
    procedure :: evaluate => au_evaluate
  end type affine_user_system

contains

  function au_nvars(self) result(n)
    class(affine_user_system), intent(in) :: self
    integer :: n
    n = self%inner%nvars()
  end function au_nvars

  function au_degrees(self) result(d)
    class(affine_user_system), intent(in) :: self
    integer, allocatable :: d(:)
    d = self%inner%degrees()
  end function au_degrees

  subroutine au_evaluate(self, x, f, jac)
    class(affine_user_system), intent(inout) :: self
    complex(wp), intent(in) :: x(:)
    complex(wp), intent(out) :: f(:)
    complex(wp), intent(out) :: jac(:, :)
    self%calls = self%calls + 1
    call self%inner%evaluate(x, f, jac)
  end subroutine au_evaluate

  subroutine build(csys, sys)
    !! add_equation for each equation of the flat description; rejected
    !! equations are left out (solve then reports err_invalid_input).
    type(c_user_system), intent(in) :: csys
    type(sparse_system), intent(out) :: sys
    integer(c_int32_t), pointer :: nt(:), ex(:)
    real(c_double), pointer :: cf(:)
    integer :: i, off, m, total, st, n

    n = csys%n
    call sys%init(n)
    if (csys%neq < 1) return
    call c_f_pointer(csys%nterms, nt, [csys%neq])
    total = sum(nt)
    call c_f_pointer(csys%coefs, cf, [2*max(total, 1)])
    call c_f_pointer(csys%exps, ex, [max(n*total, 1)])
    off = 0
    do i = 1, csys%neq
      m = nt(i)
      call sys%add_equation(cmplx(cf(2*off+1:2*(off+m):2),  &
        cf(2*off+2:2*(off+m):2), kind=wp),  &
        reshape(ex(n*off+1:n*(off+m)), [n, m]), st)
      off = off + m
    end do
  end subroutine build

  function from_c(copts) result(o)
    type(c_solve_options), intent(in) :: copts
    type(solve_options) :: o
    o%tracker%step_init = copts%tracker%step_init
    o%tracker%step_max = copts%tracker%step_max
    o%tracker%step_min = copts%tracker%step_min
    o%tracker%corrector_tol = copts%tracker%corrector_tol
    o%tracker%corrector_maxit = copts%tracker%corrector_maxit
    o%tracker%contraction = copts%tracker%contraction
    o%tracker%grow_after = copts%tracker%grow_after
    o%tracker%max_steps = copts%tracker%max_steps
    o%endgame%t_start = copts%endgame%t_start
    o%endgame%nsamples = copts%endgame%nsamples
    o%endgame%radius_factor = copts%endgame%radius_factor
    o%endgame%radius_min = copts%endgame%radius_min
    o%endgame%max_winding = copts%endgame%max_winding
    o%endgame%loop_radius = copts%endgame%loop_radius
    o%endgame%winding_agreement = copts%endgame%winding_agreement
    o%endgame%tol = copts%endgame%tol
    o%endgame%residual_tol = copts%endgame%residual_tol
    o%endgame%closure_factor = copts%endgame%closure_factor
    o%use_endgame = copts%use_endgame /= 0
    o%seed = copts%seed
    o%refine_maxit = copts%refine_maxit
    o%infinity_tol = copts%infinity_tol
    o%singular_cond = copts%singular_cond
    o%real_tol = copts%real_tol
    o%dedup_tol = copts%dedup_tol
    o%singular_dedup_tol = copts%singular_dedup_tol
  end function from_c

  subroutine put_z(p, k, z)
    real(c_double), intent(inout) :: p(:)
    integer, intent(in) :: k
    complex(wp), intent(in) :: z(:)
    integer :: j, base
    base = 2*size(z)*(k - 1)
    do j = 1, size(z)
      p(base + 2*j - 1) = real(z(j), wp)
      p(base + 2*j) = aimag(z(j))
    end do
  end subroutine put_z

  subroutine export(res, summ, paths)
    type(solve_result), intent(in) :: res
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int32_t), pointer :: ip(:)
    real(c_double), pointer :: rp(:)
    integer :: k, n, np

    n = res%n
    np = res%npaths
    summ%n = n
    summ%npaths = np
    summ%nsolutions = res%nsolutions
    summ%nreal = res%nreal
    summ%nsingular = res%nsingular
    summ%ninfinite = res%ninfinite
    summ%nfailed = res%nfailed
    summ%nduplicates = res%nduplicates
    if (c_associated(summ%solutions)) then
      call c_f_pointer(summ%solutions, rp, [2*n*max(np, 1)])
      do k = 1, res%nsolutions
        call put_z(rp, k, res%solutions(:, k))
      end do
    end if
    if (c_associated(summ%real_solutions)) then
      call c_f_pointer(summ%real_solutions, rp, [n*max(np, 1)])
      do k = 1, res%nreal
        rp(n*(k - 1) + 1:n*k) = res%real_solutions(:, k)
      end do
    end if
    if (c_associated(summ%singular_solutions)) then
      call c_f_pointer(summ%singular_solutions, rp, [2*n*max(np, 1)])
      do k = 1, res%nsingular
        call put_z(rp, k, res%singular_solutions(:, k))
      end do
    end if
    if (c_associated(summ%multiplicities)) then
      call c_f_pointer(summ%multiplicities, ip, [max(np, 1)])
      ip(1:res%nsingular) = res%multiplicities
    end if
    do k = 1, np
      associate (p => res%paths(k))
        call put_i(paths%status, k, p%status)
        call put_r(paths%t, k, p%t)
        if (c_associated(paths%z)) then
          call c_f_pointer(paths%z, rp, [2*(n + 1)*np])
          call put_z(rp, k, p%z)
        end if
        if (c_associated(paths%x)) then
          call c_f_pointer(paths%x, rp, [2*n*np])
          if (allocated(p%x)) then
            call put_z(rp, k, p%x)
          else
            rp(2*n*(k - 1) + 1:2*n*k) = 0.0_wp
          end if
        end if
        call put_i(paths%at_infinity, k, merge(1, 0, p%at_infinity))
        call put_i(paths%singular, k, merge(1, 0, p%singular))
        call put_i(paths%is_real, k, merge(1, 0, p%is_real))
        call put_i(paths%winding, k, p%winding_number)
        call put_i(paths%eg_status, k, p%endgame_status)
        call put_r(paths%eg_error, k, p%endgame_error)
        call put_r(paths%cond, k, p%cond)
        call put_r(paths%residual, k, p%residual)
        call put_i(paths%solution_index, k, p%solution_index)
        call put_i(paths%singular_index, k, p%singular_index)
        call put_i(paths%naccepted, k, p%naccepted)
        call put_i(paths%nrejected, k, p%nrejected)
      end associate
    end do

  contains

    subroutine put_i(cp, k, v)
      type(c_ptr), intent(in) :: cp
      integer, intent(in) :: k, v
      integer(c_int32_t), pointer :: a(:)
      if (.not. c_associated(cp)) return
      call c_f_pointer(cp, a, [np])
      a(k) = v
    end subroutine put_i

    subroutine put_r(cp, k, v)
      type(c_ptr), intent(in) :: cp
      integer, intent(in) :: k
      real(wp), intent(in) :: v
      real(c_double), pointer :: a(:)
      if (.not. c_associated(cp)) return
      call c_f_pointer(cp, a, [np])
      a(k) = v
    end subroutine put_r

  end subroutine export

  function phc_f_solve(csys, copts, summ, paths) result(status)  &
      bind(C, name="phc_f_solve")
    type(c_user_system), intent(in) :: csys
    type(c_solve_options), intent(in) :: copts
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int) :: status
    type(sparse_system), target :: sys
    type(solve_result) :: res
    integer :: st

    call build(csys, sys)
    call solve(sys, res, st, from_c(copts))
    status = st
    if (st /= status_ok) return
    call export(res, summ, paths)
  end function phc_f_solve

  function phc_f_solve_user(csys, copts, summ, paths) result(status)  &
      bind(C, name="phc_f_solve_user")
    !! solve on the system wrapped as a user evaluator (affine_user_system):
    !! the host tracker over the OpenMP threads.
    type(c_user_system), intent(in) :: csys
    type(c_solve_options), intent(in) :: copts
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int) :: status
    type(affine_user_system), target :: usr
    type(solve_result) :: res
    integer :: st

    call build(csys, usr%inner)
    call solve(usr, res, st, from_c(copts))
    status = st
    if (st /= status_ok) return
    call export(res, summ, paths)
  end function phc_f_solve_user

  function phc_f_set_threads(nthreads) result(previous)  &
      bind(C, name="phc_f_set_threads")
    !! Host threads of later solve calls made by the calling thread (the
    !! OpenMP default when nthreads <= 0 is never set); returns the
    !! previous count.
    integer(c_int32_t), value :: nthreads
    integer(c_int32_t) :: previous
    previous = omp_get_max_threads()
    if (nthreads > 0) call omp_set_num_threads(int(nthreads))
  end function phc_f_set_threads

  function phc_f_homotopy(csys, seed, gamma, b, patch, nstart)  &
      result(status) bind(C, name="phc_f_homotopy")
    type(c_user_system), intent(in) :: csys
    integer(c_int32_t), value :: seed
    type(c_ptr), value :: gamma, b, patch, nstart
    integer(c_int) :: status
    type(sparse_system), target :: sys
    type(total_degree_homotopy) :: hom
    type(rng) :: gen
    real(c_double), pointer :: p(:)
    integer(c_int64_t), pointer :: ns
    integer :: st

    call build(csys, sys)
    call gen%seed(seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= status_ok) return
    if (c_associated(gamma)) then
      call c_f_pointer(gamma, p, [2])
      p(1) = real(hom%gamma, wp)
      p(2) = aimag(hom%gamma)
    end if
    if (c_associated(b)) then
      call c_f_pointer(b, p, [2*hom%n])
      call put_z(p, 1, hom%b)
    end if
    if (c_associated(patch)) then
      call c_f_pointer(patch, p, [2*(hom%n + 1)])
      call put_z(p, 1, hom%patch)
    end if
    if (c_associated(nstart)) then
      call c_f_pointer(nstart, ns)
      ns = hom%nstart
    end if
  end function phc_f_homotopy

  function phc_f_start_points(csys, seed, first, count, start)  &
      result(status) bind(C, name="phc_f_start_points")
    type(c_user_system), intent(in) :: csys
    integer(c_int32_t), value :: seed
    integer(c_int64_t), value :: first, count
    type(c_ptr), value :: start
    integer(c_int) :: status
    type(sparse_system), target :: sys
    type(total_degree_homotopy) :: hom
    type(rng) :: gen
    real(c_double), pointer :: p(:)
    complex(wp), allocatable :: z(:)
    integer :: st, k

    call build(csys, sys)
    call gen%seed(seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= status_ok) return
    allocate(z(hom%n + 1))
    call c_f_pointer(start, p, [2*(hom%n + 1)*max(count, 1_c_int64_t)])
    do k = 1, int(count)
      call hom%start_solution(int(first) + k, z)
      call put_z(p, k, z)
    end do
  end function phc_f_start_points

  function phc_f_pack(csys, nterms, maxexp, first, deg, hexps, coefs)  &
      result(status) bind(C, name="phc_f_pack")
    type(c_user_system), intent(in) :: csys
    integer(c_int32_t), intent(out) :: nterms, maxexp
    type(c_ptr), value :: first, deg, hexps, coefs
    integer(c_int) :: status
    type(sparse_system), target :: sys
    integer(c_int32_t), pointer :: pf(:), pd(:), ph(:)
    real(c_double), pointer :: pc(:)
    integer :: n, nt

    call build(csys, sys)
    n = sys%nvars()
    nt = sys%total_terms()
    nterms = nt
    maxexp = sys%max_exponent()
    status = status_ok
    if (sys%nequations() /= n) status = 1
    if (.not. (c_associated(first) .and. c_associated(deg) .and.  &
      c_associated(hexps) .and. c_associated(coefs))) return
    call c_f_pointer(first, pf, [n + 1])
    call c_f_pointer(deg, pd, [max(n, 1)])
    call c_f_pointer(hexps, ph, [max((n + 1)*nt, 1)])
    call c_f_pointer(coefs, pc, [max(2*nt, 1)])
    pf = 0
    pd = 0
    call sys%pack(pf, pd, ph, pc)
  end function phc_f_pack

  function phc_f_classify(csys, copts, npaths, crec, summ, paths)  &
      result(status) bind(C, name="phc_f_classify")
    !! The host half of solve on endpoint records produced elsewhere (the
    !! serving engine, or phc_track_paths called directly).
    type(c_user_system), intent(in) :: csys
    type(c_solve_options), intent(in) :: copts
    integer(c_int64_t), value :: npaths
    type(c_endpoints), intent(in) :: crec
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int) :: status
    type(sparse_system), target :: sys
    type(total_degree_homotopy) :: hom
    type(solve_options) :: opts
    type(solve_result) :: res
    type(endpoint_records) :: rec
    type(rng) :: gen
    integer(c_int32_t), pointer :: ip(:)
    real(c_double), pointer :: rp(:)
    complex(c_double_complex), pointer :: zp(:, :)
    integer :: st, np, nn

    call build(csys, sys)
    opts = from_c(copts)
    call gen%seed(opts%seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= status_ok) return
    np = int(npaths)
    nn = hom%n + 1
    call c_f_pointer(crec%status, ip, [np])
    rec%status = ip
    call c_f_pointer(crec%winding, ip, [np])
    rec%winding = ip
    call c_f_pointer(crec%eg_status, ip, [np])
    rec%eg_status = ip
    call c_f_pointer(crec%naccepted, ip, [np])
    rec%naccepted = ip
    call c_f_pointer(crec%nrejected, ip, [np])
    rec%nrejected = ip
    call c_f_pointer(crec%t, rp, [np])
    rec%t = rp
    call c_f_pointer(crec%eg_error, rp, [np])
    rec%eg_error = rp
    call c_f_pointer(crec%z, zp, [nn, np])
    rec%z = zp
    call classify_endpoints(hom, opts, rec, res)
    call export(res, summ, paths)
  end function phc_f_classify

  subroutine phc_f_set_device(device) bind(C, name="phc_f_set_device")
    integer(c_int32_t), value :: device
    call set_device(int(device))
  end subroutine phc_f_set_device

end module polyhomocont__capi
