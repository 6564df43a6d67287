module phc_ref_driver
!! TEST INFRASTRUCTURE ONLY -- the checker, never the product.
!!
!! C-callable driver around the *unmodified* reference polyhomocont,
!! compiled from its own sources under /root/reference/src by
!! oracle/Makefile into oracle/_ref/libphc_ref.so. Only tests, the smoke
!! check and the CPU-baseline leg of bench.py load it.
!!
!! The entry points take the same flat structs as include/phc.h and
!! include/polyhomocont.h, with the prefix phc_ref_:
!!   phc_ref_solve         reference solve (src/solve.f90:154-215)
!!   phc_ref_endpoints     the per-path unit of solve, `endpoint`
!!                         (src/solve.f90:217-252), restated here from the
!!                         public track_path + cauchy_endgame because
!!                         `endpoint` itself is private
!!   phc_ref_homotopy,     tdh_init / tdh_start_solution
!!   phc_ref_start_points  (src/homotopy.f90:104-177)
!!   phc_ref_eval          tdh_evaluate (src/homotopy.f90:188-231)
!!   phc_ref_eval_sparse   sparse_evaluate_homogeneous (src/sparse.f90:209)
!!   phc_ref_track_segment track_segment (src/tracker.f90:196-264)
!!   phc_ref_rng           rng draws (src/random.f90:43-109)
!!   phc_ref_lu            lu_factor + lu_solve (src/linalg.f90:26-114)
!!   phc_ref_cond1         cond1 (src/linalg.f90:135-174)

  use, intrinsic :: iso_c_binding
  use polyhomocont, only : wp, sparse_system, solve, solve_options,  &
    solve_result, tracker_options, endgame_options, path_success,  &
    path_endgame_failed
  use polyhomocont__random, only : rng
  use polyhomocont__system, only : poly_system
  use polyhomocont__linalg, only : lu_factor, lu_solve, cond1
  use polyhomocont__homotopy, only : total_degree_homotopy
  use polyhomocont__tracker, only : path_info, track_path, track_segment,  &
    line_segment, arc_segment, segment
  use polyhomocont__endgame, only : endgame_result, cauchy_endgame,  &
    endgame_failed

  implicit none

  private

  type, bind(C) :: c_tracker_options
    real(c_double) :: step_init, step_max, step_min, corrector_tol,  &
      contraction
    integer(c_int32_t) :: corrector_maxit, grow_after, max_steps, reserved
  end type c_tracker_options

  type, bind(C) :: c_endgame_options
    real(c_double) :: t_start, radius_factor, radius_min, loop_radius,  &
      winding_agreement, tol, residual_tol, closure_factor
    integer(c_int32_t) :: nsamples, max_winding
  end type c_endgame_options

  type, bind(C) :: c_solve_options
    type(c_tracker_options) :: tracker
    type(c_endgame_options) :: endgame
    real(c_double) :: infinity_tol, singular_cond, real_tol, dedup_tol,  &
      singular_dedup_tol
    integer(c_int32_t) :: use_endgame, seed, refine_maxit, reserved
  end type c_solve_options

  type, bind(C) :: c_user_system
    integer(c_int32_t) :: n, neq
    type(c_ptr) :: nterms, coefs, exps
  end type c_user_system

  type, bind(C) :: c_endpoints
    type(c_ptr) :: status, t, z, winding, eg_status, eg_error,  &
      eg_residual, naccepted, nrejected, nlin, nbe
  end type c_endpoints

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

  !> A user evaluator over a sparse system: nvars, degrees and the affine
  !> evaluate only.
  type, extends(poly_system) :: affine_user_system
    type(sparse_system) :: inner
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
    call self%inner%evaluate(x, f, jac)
  end subroutine au_evaluate

  subroutine build_system(csys, sys)
    !! add_equation for every equation of the flat system; statuses of
    !! rejected equations are not checked here (solve then reports the
    !! non-square system, as for any caller that ignores them).
    type(c_user_system), intent(in) :: csys
    type(sparse_system), intent(out) :: sys

    integer(c_int32_t), pointer :: nt(:), ex(:)
    real(c_double), pointer :: cf(:)
    complex(wp), allocatable :: c(:)
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
      c = cmplx(cf(2*off+1:2*(off+m):2), cf(2*off+2:2*(off+m):2), kind=wp)
      call sys%add_equation(c, reshape(ex(n*off+1:n*(off+m)), [n, m]), st)
      off = off + m
    end do
  end subroutine build_system

  subroutine to_options(copts, opts)
    type(c_solve_options), intent(in) :: copts
    type(solve_options), intent(out) :: opts

    opts%tracker%step_init = copts%tracker%step_init
    opts%tracker%step_max = copts%tracker%step_max
    opts%tracker%step_min = copts%tracker%step_min
    opts%tracker%corrector_tol = copts%tracker%corrector_tol
    opts%tracker%corrector_maxit = copts%tracker%corrector_maxit
    opts%tracker%contraction = copts%tracker%contraction
    opts%tracker%grow_after = copts%tracker%grow_after
    opts%tracker%max_steps = copts%tracker%max_steps
    opts%endgame%t_start = copts%endgame%t_start
    opts%endgame%nsamples = copts%endgame%nsamples
    opts%endgame%radius_factor = copts%endgame%radius_factor
    opts%endgame%radius_min = copts%endgame%radius_min
    opts%endgame%max_winding = copts%endgame%max_winding
    opts%endgame%loop_radius = copts%endgame%loop_radius
    opts%endgame%winding_agreement = copts%endgame%winding_agreement
    opts%endgame%tol = copts%endgame%tol
    opts%endgame%residual_tol = copts%endgame%residual_tol
    opts%endgame%closure_factor = copts%endgame%closure_factor
    opts%use_endgame = copts%use_endgame /= 0
    opts%seed = copts%seed
    opts%refine_maxit = copts%refine_maxit
    opts%infinity_tol = copts%infinity_tol
    opts%singular_cond = copts%singular_cond
    opts%real_tol = copts%real_tol
    opts%dedup_tol = copts%dedup_tol
    opts%singular_dedup_tol = copts%singular_dedup_tol
  end subroutine to_options

  subroutine put_z(p, k, z)
    !! Stores complex vector z as column k (1-based) of a [.][size(z)][2]
    !! double array.
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

  function phc_ref_homotopy(csys, seed, gamma, b, patch, nstart)  &
      result(status) bind(C, name="phc_ref_homotopy")
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

    call build_system(csys, sys)
    call gen%seed(seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= 0) return
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
  end function phc_ref_homotopy

  function phc_ref_start_points(csys, seed, first, count, start)  &
      result(status) bind(C, name="phc_ref_start_points")
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
    integer :: st, kk

    call build_system(csys, sys)
    call gen%seed(seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= 0) return
    allocate(z(hom%n + 1))
    call c_f_pointer(start, p, [2*(hom%n + 1)*count])
    do kk = 1, int(count)
      call hom%start_solution(int(first) + kk, z)
      call put_z(p, kk, z)
    end do
  end function phc_ref_start_points

  function phc_ref_endpoints(csys, copts, first, count, out)  &
      result(status) bind(C, name="phc_ref_endpoints")
    !! endpoint() (src/solve.f90:217-252) for paths first+1 .. first+count.
    type(c_user_system), intent(in) :: csys
    type(c_solve_options), intent(in) :: copts
    integer(c_int64_t), value :: first, count
    type(c_endpoints), intent(in) :: out
    integer(c_int) :: status

    type(sparse_system), target :: sys
    type(total_degree_homotopy) :: hom
    type(solve_options) :: opts
    type(rng) :: gen
    type(path_info) :: info
    type(endgame_result) :: eg
    complex(wp), allocatable :: z(:)
    integer(c_int32_t), pointer :: pst(:), pwind(:), pegs(:), pacc(:),  &
      prej(:)
    real(c_double), pointer :: pt(:), pz(:), perr(:), pres(:)
    real(wp) :: t_end
    integer :: st, kk, np

    call build_system(csys, sys)
    call to_options(copts, opts)
    call gen%seed(opts%seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= 0) return
    np = int(count)
    allocate(z(hom%n + 1))
    call c_f_pointer(out%status, pst, [np])
    call c_f_pointer(out%t, pt, [np])
    call c_f_pointer(out%z, pz, [2*(hom%n + 1)*np])
    call c_f_pointer(out%winding, pwind, [np])
    call c_f_pointer(out%eg_status, pegs, [np])
    call c_f_pointer(out%eg_error, perr, [np])
    call c_f_pointer(out%naccepted, pacc, [np])
    call c_f_pointer(out%nrejected, prej, [np])
    pres => null()
    if (c_associated(out%eg_residual)) call c_f_pointer(out%eg_residual,  &
      pres, [np])

    t_end = 0.0_wp
    if (opts%use_endgame) t_end = opts%endgame%t_start
    do kk = 1, np
      call hom%start_solution(int(first) + kk, z)
      call track_path(hom, z, 1.0_wp, t_end, opts%tracker, info)
      pst(kk) = info%status
      pt(kk) = real(info%t, wp)
      pacc(kk) = info%naccepted
      prej(kk) = info%nrejected
      pwind(kk) = 0
      pegs(kk) = 0
      perr(kk) = 0.0_wp
      if (associated(pres)) pres(kk) = 0.0_wp
      if (info%status == path_success .and. opts%use_endgame) then
        call cauchy_endgame(hom, z, opts%endgame, opts%tracker, eg)
        pacc(kk) = pacc(kk) + eg%naccepted
        prej(kk) = prej(kk) + eg%nrejected
        if (eg%status == endgame_failed) then
          pst(kk) = path_endgame_failed
        else
          z = eg%z
          pt(kk) = 0.0_wp
          pwind(kk) = eg%winding_number
          pegs(kk) = eg%status
          perr(kk) = eg%error
          if (associated(pres)) pres(kk) = eg%residual
        end if
      end if
      call put_z(pz, kk, z)
    end do
  end function phc_ref_endpoints

  function phc_ref_solve(csys, copts, summ, paths) result(status)  &
      bind(C, name="phc_ref_solve")
    type(c_user_system), intent(in) :: csys
    type(c_solve_options), intent(in) :: copts
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int) :: status
    type(sparse_system), target :: sys
    type(solve_options) :: opts

    call build_system(csys, sys)
    call to_options(copts, opts)
    status = solve_export(sys, opts, summ, paths)
  end function phc_ref_solve

  function phc_ref_solve_user(csys, copts, summ, paths) result(status)  &
      bind(C, name="phc_ref_solve_user")
    !! The reference solve on the system wrapped as a user evaluator (only
    !! the affine F and Jacobian; the default homogenization of
    !! src/system.f90:73-117), the same wrapper as the library's
    !! phc_f_solve_user.
    type(c_user_system), intent(in) :: csys
    type(c_solve_options), intent(in) :: copts
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int) :: status
    type(affine_user_system), target :: usr
    type(solve_options) :: opts

    call build_system(csys, usr%inner)
    call to_options(copts, opts)
    status = solve_export(usr, opts, summ, paths)
  end function phc_ref_solve_user

  function solve_export(sys, opts, summ, paths) result(status)
    class(poly_system), intent(inout), target :: sys
    type(solve_options), intent(in) :: opts
    type(c_solve_summary), intent(inout) :: summ
    type(c_path_results), intent(in) :: paths
    integer(c_int) :: status

    type(solve_result) :: res
    integer(c_int32_t), pointer :: ip(:)
    real(c_double), pointer :: rp(:)
    integer :: st, k, np, n, j

    call solve(sys, res, st, opts)
    status = st
    if (st /= 0) return
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
      call c_f_pointer(summ%solutions, rp, [2*n*np])
      do k = 1, res%nsolutions
        call put_z(rp, k, res%solutions(:, k))
      end do
    end if
    if (c_associated(summ%real_solutions)) then
      call c_f_pointer(summ%real_solutions, rp, [n*np])
      do k = 1, res%nreal
        rp(n*(k-1)+1:n*k) = res%real_solutions(:, k)
      end do
    end if
    if (c_associated(summ%singular_solutions)) then
      call c_f_pointer(summ%singular_solutions, rp, [2*n*np])
      do k = 1, res%nsingular
        call put_z(rp, k, res%singular_solutions(:, k))
      end do
    end if
    if (c_associated(summ%multiplicities)) then
      call c_f_pointer(summ%multiplicities, ip, [np])
      ip(1:res%nsingular) = res%multiplicities
    end if

    do k = 1, np
      associate(p => res%paths(k))
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
            do j = 1, 2*n
              rp(2*n*(k-1) + j) = 0.0_wp
            end do
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

  end function solve_export

  function phc_ref_eval(csys, seed, z, t, h, hz, ht) result(status)  &
      bind(C, name="phc_ref_eval")
    !! tdh_evaluate at z (N complex) and complex t; hz is column-major N x N.
    type(c_user_system), intent(in) :: csys
    integer(c_int32_t), value :: seed
    complex(c_double_complex), intent(in) :: z(*)
    complex(c_double_complex), intent(in) :: t
    complex(c_double_complex), intent(out) :: h(*), hz(*), ht(*)
    integer(c_int) :: status

    type(sparse_system), target :: sys
    type(total_degree_homotopy) :: hom
    type(rng) :: gen
    complex(wp), allocatable :: hh(:), hhz(:, :), hht(:)
    integer :: st, nn

    call build_system(csys, sys)
    call gen%seed(seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= 0) return
    nn = hom%n + 1
    allocate(hh(nn), hhz(nn, nn), hht(nn))
    call hom%evaluate(z(1:nn), t, hh, hhz, hht)
    h(1:nn) = hh
    hz(1:nn*nn) = reshape(hhz, [nn*nn])
    ht(1:nn) = hht
  end function phc_ref_eval

  function phc_ref_eval_sparse(csys, z, f, jac) result(status)  &
      bind(C, name="phc_ref_eval_sparse")
    !! evaluate_homogeneous; jac is column-major n x (n+1).
    type(c_user_system), intent(in) :: csys
    complex(c_double_complex), intent(in) :: z(*)
    complex(c_double_complex), intent(out) :: f(*), jac(*)
    integer(c_int) :: status

    type(sparse_system), target :: sys
    complex(wp), allocatable :: ff(:), jj(:, :)
    integer :: n

    call build_system(csys, sys)
    n = csys%n
    status = 0
    allocate(ff(n), jj(n, n + 1))
    call sys%evaluate_homogeneous(z(1:n+1), ff, jj)
    f(1:n) = ff
    jac(1:n*(n+1)) = reshape(jj, [n*(n+1)])
  end function phc_ref_eval_sparse

  function phc_ref_track_segment(csys, seed, ctopts, kind, t0, t1, radius,  &
      theta0, dtheta, z, info_status, info_s, info_t, info_step,  &
      info_counts) result(status) bind(C, name="phc_ref_track_segment")
    !! track_segment on a line (kind 1: t0 -> t1) or an arc (kind 2).
    type(c_user_system), intent(in) :: csys
    integer(c_int32_t), value :: seed
    type(c_tracker_options), intent(in) :: ctopts
    integer(c_int32_t), value :: kind
    complex(c_double_complex), intent(in) :: t0, t1
    real(c_double), value :: radius, theta0, dtheta
    complex(c_double_complex), intent(inout) :: z(*)
    integer(c_int32_t), intent(out) :: info_status
    real(c_double), intent(out) :: info_s, info_step
    complex(c_double_complex), intent(out) :: info_t
    integer(c_int32_t), intent(out) :: info_counts(2)
    integer(c_int) :: status

    type(sparse_system), target :: sys
    type(total_degree_homotopy) :: hom
    type(rng) :: gen
    type(tracker_options) :: topts
    type(path_info) :: info
    type(segment) :: seg
    complex(wp), allocatable :: zz(:)
    integer :: st, nn

    call build_system(csys, sys)
    call gen%seed(seed)
    call hom%init(sys, gen, st)
    status = st
    if (st /= 0) return
    nn = hom%n + 1
    topts%step_init = ctopts%step_init
    topts%step_max = ctopts%step_max
    topts%step_min = ctopts%step_min
    topts%corrector_tol = ctopts%corrector_tol
    topts%corrector_maxit = ctopts%corrector_maxit
    topts%contraction = ctopts%contraction
    topts%grow_after = ctopts%grow_after
    topts%max_steps = ctopts%max_steps
    if (kind == 1) then
      seg = line_segment(t0, t1)
    else
      seg = arc_segment(radius, theta0, dtheta)
    end if
    zz = z(1:nn)
    call track_segment(hom, zz, seg, topts, info)
    z(1:nn) = zz
    info_status = info%status
    info_s = info%s
    info_t = info%t
    info_step = info%step
    info_counts(1) = info%naccepted
    info_counts(2) = info%nrejected
  end function phc_ref_track_segment

  subroutine phc_ref_rng(seed, kind, count, out) bind(C, name="phc_ref_rng")
    !! kind 0: uniform (count doubles); 1: unit_complex; 2: normal_complex
    !! (count complex numbers, 2*count doubles).
    integer(c_int32_t), value :: seed, kind
    integer(c_int64_t), value :: count
    real(c_double), intent(out) :: out(*)
    type(rng) :: gen
    complex(wp) :: z
    integer :: k

    call gen%seed(seed)
    do k = 1, int(count)
      select case (kind)
      case (0)
        out(k) = gen%uniform()
      case (1)
        z = gen%unit_complex()
        out(2*k - 1) = real(z, wp)
        out(2*k) = aimag(z)
      case default
        z = gen%normal_complex()
        out(2*k - 1) = real(z, wp)
        out(2*k) = aimag(z)
      end select
    end do
  end subroutine phc_ref_rng

  function phc_ref_lu(n, a, ipiv, b) result(status) bind(C, name="phc_ref_lu")
    !! lu_factor(a) then lu_solve(a, ipiv, b); a column-major n x n.
    integer(c_int32_t), value :: n
    complex(c_double_complex), intent(inout) :: a(n, n)
    integer(c_int32_t), intent(out) :: ipiv(n)
    complex(c_double_complex), intent(inout) :: b(n)
    integer(c_int) :: status
    integer :: st, ip(n)

    call lu_factor(a, ip, st)
    ipiv = ip
    status = st
    if (st /= 0) return
    call lu_solve(a, ip, b)
  end function phc_ref_lu

  function phc_ref_cond1(n, a) result(c) bind(C, name="phc_ref_cond1")
    integer(c_int32_t), value :: n
    complex(c_double_complex), intent(in) :: a(n, n)
    real(c_double) :: c
    c = cond1(a)
  end function phc_ref_cond1

end module phc_ref_driver
