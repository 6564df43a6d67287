module polyhomocont__endgame
!! Cauchy end game options, results and status codes (the reference's
!! public end game interface, src/endgame.f90:64-120), plus its CPU
!! implementation for systems with user-supplied evaluators. Sparse systems
!! run the same end game inside the GPU kernels.
!!
!! Method (Morgan, Sommese, Wampler, Numer. Math. 58 (1991) 669): move
!! inward on the real axis through r_k = t_start radius_factor^k; estimate
!! the winding number c from consecutive radial differences; once two
!! estimates agree (or below loop_radius), track loops |t| = r of nsamples
!! arcs until the path closes and take the mean of the samples as z(0);
!! stop when consecutive estimates agree to tol and pass the backward-error
!! test (src/endgame.f90:124-378).

  use polyhomocont__config, only : wp, pi
  use polyhomocont__linalg, only : norm_inf
  use polyhomocont__homotopy, only : homotopy
  use polyhomocont__tracker, only : tracker_options, path_info,  &
    track_segment, line_segment, arc_segment, path_success

  implicit none

  private

  public :: endgame_options, endgame_result, cauchy_endgame
  public :: endgame_converged, endgame_not_converged, endgame_failed

  integer, parameter :: endgame_converged = 0
  integer, parameter :: endgame_not_converged = 1
  integer, parameter :: endgame_failed = 2

  type :: endgame_options
    real(wp) :: t_start = 0.1_wp
    integer :: nsamples = 8
    real(wp) :: radius_factor = 0.25_wp
    real(wp) :: radius_min = 1.0e-13_wp
    integer :: max_winding = 16
    real(wp) :: loop_radius = 1.0e-5_wp
    real(wp) :: winding_agreement = 0.2_wp
    real(wp) :: tol = 1.0e-11_wp
    real(wp) :: residual_tol = 1.0e-8_wp
    real(wp) :: closure_factor = 0.1_wp
  end type endgame_options

  type :: endgame_result
    integer :: status = endgame_failed
    complex(wp), allocatable :: z(:)
    integer :: winding_number = 0
    real(wp) :: error = huge(1.0_wp)
    real(wp) :: residual = huge(1.0_wp)
    real(wp) :: radius = 0.0_wp
    integer :: naccepted = 0
    integer :: nrejected = 0
  end type endgame_result

contains

  subroutine cauchy_endgame(hom, z, opts, topts, res)
    !! z: the path at t = t_start on entry, the last real-axis point on
    !! exit; res%z: the estimate of z(0).
    class(homotopy), intent(inout) :: hom
    complex(wp), intent(inout) :: z(:)
    type(endgame_options), intent(in) :: opts
    type(tracker_options), intent(in) :: topts
    type(endgame_result), intent(out) :: res
    type(tracker_options) :: sopts
    type(path_info) :: info
    complex(wp), dimension(size(z)) :: est, last_est, zback
    real(wp) :: r, r_in, best_resid, dist, dist_prev, c_est, c_prev
    integer :: wind, nloops
    logical :: have_last, gated, closed

    ! Arcs and radial segments are short: one step each, step carried over
    sopts = topts
    sopts%step_init = 1.0_wp
    sopts%step_max = 1.0_wp
    sopts%grow_after = 2

    res%status = endgame_failed
    res%z = z
    r = opts%t_start
    have_last = .false.
    best_resid = huge(1.0_wp)
    gated = .false.
    dist_prev = -1.0_wp
    c_est = -1.0_wp
    c_prev = -1.0_wp
    nloops = opts%max_winding

    do
      closed = .false.
      if (gated .or. r < opts%loop_radius) then
        if (.not. gated) nloops = opts%max_winding
        call loops(hom, z, r, opts, nloops, sopts, est, wind, closed, res)
      end if
      if (closed) then
        call keep_estimate()
        if (res%status == endgame_converged) return
        last_est = est
        have_last = .true.
      else
        have_last = .false.
        gated = .false.
      end if

      r_in = opts%radius_factor*r
      if (r_in < opts%radius_min) return
      zback = z
      call track_segment(hom, z, line_segment(cmplx(r, 0.0_wp, kind=wp),  &
        cmplx(r_in, 0.0_wp, kind=wp)), sopts, info)
      sopts%step_init = min(max(info%step, sopts%step_min), 1.0_wp)
      res%naccepted = res%naccepted + info%naccepted
      res%nrejected = res%nrejected + info%nrejected
      if (info%status /= path_success) then
        z = zback
        if (.not. (gated .or. r < opts%loop_radius)) then
          call loops(hom, z, r, opts, opts%max_winding, sopts, est, wind,  &
            closed, res)
          have_last = .false.
          if (closed) call keep_estimate()
        end if
        return
      end if
      r = r_in

      dist = norm_inf(z - zback)
      if (.not. gated .and. dist_prev > 0.0_wp .and. dist > 0.0_wp) then
        c_prev = c_est
        c_est = -1.0_wp
        if (dist_prev > dist) c_est = log(1.0_wp/opts%radius_factor)  &
          /log(dist_prev/dist)
        if (c_est >= 0.5_wp .and. c_est <= opts%max_winding + 0.5_wp  &
          .and. c_prev > 0.0_wp) then
          if (abs(c_est - c_prev) <= opts%winding_agreement  &
            *max(1.0_wp, c_est)) then
            gated = .true.
            nloops = min(opts%max_winding, 2*nint(c_est) + 2)
          end if
        end if
      end if
      dist_prev = dist
    end do

  contains

    subroutine keep_estimate()
      real(wp) :: err, resid
      err = huge(1.0_wp)
      if (have_last) err = norm_inf(est - last_est)/max(norm_inf(est),  &
        tiny(1.0_wp))
      resid = backward_error(hom, est)
      if (res%status == endgame_failed .or. preferable(resid, err,  &
        best_resid, res%error, opts%residual_tol)) then
        res%status = endgame_not_converged
        res%z = est
        res%winding_number = wind
        res%error = err
        res%residual = resid
        res%radius = r
        best_resid = resid
      end if
      if (err <= opts%tol .and. resid <= opts%residual_tol)  &
        res%status = endgame_converged
    end subroutine keep_estimate

  end subroutine cauchy_endgame

  function backward_error(hom, z) result(be)
    !! ||H(z,0)|| / (max_i sum_j |H_z(i,j)| ||z||), maximum norms.
    class(homotopy), intent(inout) :: hom
    complex(wp), intent(in) :: z(:)
    real(wp) :: be
    complex(wp) :: h(size(z)), hz(size(z), size(z)), ht(size(z))
    call hom%evaluate(z, (0.0_wp, 0.0_wp), h, hz, ht)
    be = norm_inf(h)/max(maxval(sum(abs(hz), dim=2))*norm_inf(z),  &
      tiny(1.0_wp))
  end function backward_error

  pure function preferable(resid, err, best_resid, best_err, rtol) result(b)
    !! Residual-valid estimates first, then the smaller difference.
    real(wp), intent(in) :: resid, err, best_resid, best_err, rtol
    logical :: b
    if ((resid <= rtol) .neqv. (best_resid <= rtol)) then
      b = resid <= rtol
    else
      b = err < best_err
    end if
  end function preferable

  subroutine loops(hom, z, r, opts, max_loops, sopts, est, wind, closed, res)
    !! Loops |t| = r from t = r with nsamples arcs each until the path
    !! closes (distance to the start below closure_factor times the
    !! smallest sample spacing, or at corrector level); est = sample mean.
    class(homotopy), intent(inout) :: hom
    complex(wp), intent(in) :: z(:)
    real(wp), intent(in) :: r
    type(endgame_options), intent(in) :: opts
    integer, intent(in) :: max_loops
    type(tracker_options), intent(inout) :: sopts
    complex(wp), intent(out) :: est(:)
    integer, intent(out) :: wind
    logical, intent(out) :: closed
    type(endgame_result), intent(inout) :: res
    type(path_info) :: info
    complex(wp), dimension(size(z)) :: zc, zlast, acc
    real(wp) :: arc, gap, floor
    integer :: nsum, j

    closed = .false.
    arc = 2.0_wp*pi/real(opts%nsamples, wp)
    floor = 10.0_wp*sopts%corrector_tol*norm_inf(z)
    zc = z
    acc = (0.0_wp, 0.0_wp)
    nsum = 0
    gap = huge(1.0_wp)
    do wind = 1, max_loops
      do j = 1, opts%nsamples
        acc = acc + zc
        nsum = nsum + 1
        zlast = zc
        call track_segment(hom, zc, arc_segment(r, (j - 1)*arc, arc),  &
          sopts, info)
        sopts%step_init = min(max(info%step, sopts%step_min), 1.0_wp)
        res%naccepted = res%naccepted + info%naccepted
        res%nrejected = res%nrejected + info%nrejected
        if (info%status /= path_success) return
        gap = min(gap, norm_inf(zc - zlast))
      end do
      if (norm_inf(zc - z) <= max(opts%closure_factor*gap, floor)) then
        est = acc/real(nsum, wp)
        closed = .true.
        return
      end if
    end do
  end subroutine loops

end module polyhomocont__endgame
