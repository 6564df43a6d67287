module polyhomocont__tracker
!! Tracker options, path status codes and segments of the t-plane (the
!! reference's public tracker interface, src/tracker.f90:45-110), plus the
!! CPU predictor-corrector used for systems with user-supplied evaluators.
!! Sparse systems are tracked by the GPU kernels, which implement the same
!! method (RK4 predictor on dz/ds = -H_z^-1 H_t dt/ds, Newton corrector
!! with the fast-convergence guard, step halving/doubling;
!! src/tracker.f90:196-360).

  use polyhomocont__config, only : wp, status_ok
  use polyhomocont__linalg, only : lu_factor, lu_solve, norm_inf
  use polyhomocont__homotopy, only : homotopy

  implicit none

  private

  public :: tracker_options, path_info, segment
  public :: line_segment, arc_segment, track_segment, track_path
  public :: path_success, path_failed_min_step, path_failed_max_steps

  integer, parameter :: path_success = 0
  integer, parameter :: path_failed_min_step = 1
  integer, parameter :: path_failed_max_steps = 2

  type :: tracker_options
    real(wp) :: step_init = 0.05_wp
    real(wp) :: step_max = 0.1_wp
    real(wp) :: step_min = 1.0e-14_wp
    real(wp) :: corrector_tol = 1.0e-10_wp
    integer :: corrector_maxit = 3
    real(wp) :: contraction = 0.5_wp
    integer :: grow_after = 5
    integer :: max_steps = 20000
  end type tracker_options

  type :: segment
    !! t(s), s in [0, 1]: a line t0 -> t1 (is_arc false) or the arc
    !! radius * exp(i (theta0 + s dtheta)).
    logical :: is_arc = .false.
    complex(wp) :: t0 = (1.0_wp, 0.0_wp)
    complex(wp) :: t1 = (0.0_wp, 0.0_wp)
    real(wp) :: radius = 0.0_wp
    real(wp) :: theta0 = 0.0_wp
    real(wp) :: dtheta = 0.0_wp
  contains
    procedure :: at => segment_at
    procedure :: slope => segment_slope
  end type segment

  type :: path_info
    integer :: status = path_success
    real(wp) :: s = 0.0_wp
    complex(wp) :: t = (0.0_wp, 0.0_wp)
    integer :: naccepted = 0
    integer :: nrejected = 0
    real(wp) :: step = 0.0_wp
  end type path_info

contains

  pure function line_segment(t0, t1) result(seg)
    complex(wp), intent(in) :: t0, t1
    type(segment) :: seg
    seg%is_arc = .false.
    seg%t0 = t0
    seg%t1 = t1
  end function line_segment

  pure function arc_segment(radius, theta0, dtheta) result(seg)
    real(wp), intent(in) :: radius, theta0, dtheta
    type(segment) :: seg
    seg%is_arc = .true.
    seg%radius = radius
    seg%theta0 = theta0
    seg%dtheta = dtheta
  end function arc_segment

  pure function segment_at(self, s) result(t)
    !! End points exact: a line returns t1 for s >= 1.
    class(segment), intent(in) :: self
    real(wp), intent(in) :: s
    complex(wp) :: t
    real(wp) :: angle
    if (self%is_arc) then
      angle = self%theta0 + s*self%dtheta
      t = self%radius*cmplx(cos(angle), sin(angle), kind=wp)
    else if (s >= 1.0_wp) then
      t = self%t1
    else
      t = self%t0 + s*(self%t1 - self%t0)
    end if
  end function segment_at

  pure function segment_slope(self, s) result(dt)
    class(segment), intent(in) :: self
    real(wp), intent(in) :: s
    complex(wp) :: dt
    if (self%is_arc) then
      dt = cmplx(0.0_wp, self%dtheta, kind=wp)*self%at(s)
    else
      dt = self%t1 - self%t0
    end if
  end function segment_slope

  subroutine track_path(hom, z, t_start, t_end, opts, info)
    class(homotopy), intent(inout) :: hom
    complex(wp), intent(inout) :: z(:)
    real(wp), intent(in) :: t_start, t_end
    type(tracker_options), intent(in) :: opts
    type(path_info), intent(out) :: info
    call track_segment(hom, z, line_segment(cmplx(t_start, 0.0_wp, kind=wp), &
      cmplx(t_end, 0.0_wp, kind=wp)), opts, info)
  end subroutine track_path

  subroutine track_segment(hom, z, seg, opts, info)
    !! Follows z from t(0) to t(1); on failure z is the last accepted point.
    class(homotopy), intent(inout) :: hom
    complex(wp), intent(inout) :: z(:)
    type(segment), intent(in) :: seg
    type(tracker_options), intent(in) :: opts
    type(path_info), intent(out) :: info
    complex(wp) :: znew(size(z))
    real(wp) :: s, s_next, h, step
    integer :: streak
    logical :: ok

    s = 0.0_wp
    step = min(opts%step_init, opts%step_max)
    streak = 0
    info%naccepted = 0
    info%nrejected = 0
    info%status = path_success
    do while (s < 1.0_wp)
      if (info%naccepted + info%nrejected >= opts%max_steps) then
        info%status = path_failed_max_steps
        exit
      end if
      h = min(step, 1.0_wp - s)
      s_next = s + h
      if (1.0_wp - s_next <= 1.0e-3_wp*step) then
        s_next = 1.0_wp
        h = 1.0_wp - s
      end if
      call rk4_predict(hom, seg, z, s, h, znew, ok)
      if (ok) call newton(hom, znew, seg%at(s_next), opts%corrector_tol,  &
        opts%corrector_maxit, opts%contraction, ok)
      if (ok) then
        z = znew
        s = s_next
        info%naccepted = info%naccepted + 1
        streak = streak + 1
        if (streak >= opts%grow_after) then
          step = min(2.0_wp*step, opts%step_max)
          streak = 0
        end if
      else
        info%nrejected = info%nrejected + 1
        streak = 0
        step = 0.5_wp*step
        if (step < opts%step_min) then
          info%status = path_failed_min_step
          exit
        end if
      end if
    end do
    info%s = s
    info%t = seg%at(s)
    info%step = step
  end subroutine track_segment

  subroutine rk4_predict(hom, seg, z, s, h, znew, ok)
    class(homotopy), intent(inout) :: hom
    type(segment), intent(in) :: seg
    complex(wp), intent(in) :: z(:)
    real(wp), intent(in) :: s, h
    complex(wp), intent(out) :: znew(:)
    logical, intent(out) :: ok
    complex(wp), dimension(size(z)) :: k1, k2, k3, k4

    call slope_field(hom, seg, z, s, k1, ok)
    if (.not. ok) return
    call slope_field(hom, seg, z + 0.5_wp*h*k1, s + 0.5_wp*h, k2, ok)
    if (.not. ok) return
    call slope_field(hom, seg, z + 0.5_wp*h*k2, s + 0.5_wp*h, k3, ok)
    if (.not. ok) return
    call slope_field(hom, seg, z + h*k3, s + h, k4, ok)
    if (.not. ok) return
    znew = z + h/6.0_wp*(k1 + 2.0_wp*k2 + 2.0_wp*k3 + k4)
  end subroutine rk4_predict

  subroutine slope_field(hom, seg, z, s, dz, ok)
    !! dz/ds from H_z dz = -H_t dt/ds.
    class(homotopy), intent(inout) :: hom
    type(segment), intent(in) :: seg
    complex(wp), intent(in) :: z(:)
    real(wp), intent(in) :: s
    complex(wp), intent(out) :: dz(:)
    logical, intent(out) :: ok
    complex(wp) :: h(size(z)), hz(size(z), size(z))
    integer :: piv(size(z)), status

    call hom%evaluate(z, seg%at(s), h, hz, dz)
    call lu_factor(hz, piv, status)
    ok = status == status_ok
    if (.not. ok) return
    dz = -dz*seg%slope(s)
    call lu_solve(hz, piv, dz)
  end subroutine slope_field

  subroutine newton(hom, z, t, tol, maxit, contraction, ok)
    !! Accepts when an update satisfies ||dz|| <= tol ||z|| (z updated
    !! first); rejects an update larger than contraction times the last.
    class(homotopy), intent(inout) :: hom
    complex(wp), intent(inout) :: z(:)
    complex(wp), intent(in) :: t
    real(wp), intent(in) :: tol, contraction
    integer, intent(in) :: maxit
    logical, intent(out) :: ok
    complex(wp) :: h(size(z)), hz(size(z), size(z)), ht(size(z))
    integer :: piv(size(z)), status, it
    real(wp) :: size_dz, size_prev

    ok = .false.
    size_prev = huge(1.0_wp)
    do it = 1, maxit
      call hom%evaluate(z, t, h, hz, ht)
      call lu_factor(hz, piv, status)
      if (status /= status_ok) return
      h = -h
      call lu_solve(hz, piv, h)
      z = z + h
      size_dz = norm_inf(h)
      if (size_dz <= tol*norm_inf(z)) then
        ok = .true.
        return
      end if
      if (it > 1 .and. size_dz > contraction*size_prev) return
      size_prev = size_dz
    end do
  end subroutine newton

end module polyhomocont__tracker
