module polyhomocont__device
!! ISO_C_BINDING layer over libphc_hip.so (include/phc.h): the bind(C)
!! mirrors of the C structs and the calls `solve` makes to track the paths
!! of a sparse_system on the GPU.
!!
!! The device context is the C library's process-wide context of the
!! device (phc_ctx_shared: created once per device under the library's
!! lock, kept until the process ends), so this module holds no mutable
!! state and `solve` stays reentrant. The device index is the library's
!! atomic default (set_device; -1 asks solve to track on the host).

  use, intrinsic :: iso_c_binding
  use polyhomocont__config, only : wp, status_ok, err_invalid_input,  &
    err_device, err_capacity
  use polyhomocont__sparse, only : sparse_system
  use polyhomocont__homotopy, only : total_degree_homotopy
  use polyhomocont__tracker, only : tracker_options
  use polyhomocont__endgame, only : endgame_options

  implicit none

  private

  public :: c_tracker_options, c_endgame_options, c_solve_options
  public :: c_endpoints, c_system, c_homotopy
  public :: device_track, set_device, device_error_message, host_requested
  public :: endpoint_records

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

  type, bind(C) :: c_system
    integer(c_int32_t) :: n, nterms, maxexp, reserved
    type(c_ptr) :: first, deg, hexps, coefs
  end type c_system

  type, bind(C) :: c_homotopy
    type(c_ptr) :: gamma, b, patch
  end type c_homotopy

  type, bind(C) :: c_endpoints
    type(c_ptr) :: status, t, z, winding, eg_status, eg_error,  &
      eg_residual, naccepted, nrejected, nlin, nbe
  end type c_endpoints

  !> Endpoint records of the paths of one solve (Fortran-owned storage).
  type :: endpoint_records
    integer(c_int32_t), allocatable :: status(:), winding(:), eg_status(:),  &
      naccepted(:), nrejected(:)
    real(c_double), allocatable :: t(:), eg_error(:), eg_residual(:)
    complex(c_double_complex), allocatable :: z(:, :)
  end type endpoint_records

  interface
    function phc_ctx_shared(device, ctx) result(rc)  &
        bind(C, name="phc_ctx_shared")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function phc_ctx_shared
    subroutine phc_set_default_device(device)  &
        bind(C, name="phc_set_default_device")
      import :: c_int
      integer(c_int), value :: device
    end subroutine phc_set_default_device
    function phc_get_default_device() result(device)  &
        bind(C, name="phc_get_default_device")
      import :: c_int
      integer(c_int) :: device
    end function phc_get_default_device
    function phc_track_paths(ctx, sys, hom, opts, npaths, start, out)  &
        result(rc) bind(C, name="phc_track_paths")
      import :: c_ptr, c_int, c_int64_t, c_system, c_homotopy,  &
        c_solve_options, c_endpoints
      type(c_ptr), value :: ctx
      type(c_system), intent(in) :: sys
      type(c_homotopy), intent(in) :: hom
      type(c_solve_options), intent(in) :: opts
      integer(c_int64_t), value :: npaths
      type(c_ptr), value :: start
      type(c_endpoints), intent(in) :: out
      integer(c_int) :: rc
    end function phc_track_paths
    function phc_check_system(sys, opts) result(rc)  &
        bind(C, name="phc_check_system")
      import :: c_int, c_system, c_solve_options
      type(c_system), intent(in) :: sys
      type(c_solve_options), intent(in) :: opts
      integer(c_int) :: rc
    end function phc_check_system
    function phc_thread_last_error() result(msg)  &
        bind(C, name="phc_thread_last_error")
      import :: c_ptr
      type(c_ptr) :: msg
    end function phc_thread_last_error
  end interface

contains

  subroutine set_device(device)
    !! HIP device of later solve calls (default 0). -1 tracks sparse
    !! systems on the host CPU instead, by explicit request only: solve
    !! never falls back to the host when the device it was asked for fails.
    integer, intent(in) :: device
    call phc_set_default_device(int(device, c_int))
  end subroutine set_device

  function host_requested() result(yes)
    logical :: yes
    yes = phc_get_default_device() < 0
  end function host_requested

  function device_error_message() result(msg)
    !! Message of the calling thread's last failed library call.
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: k
    msg = ""
    cp = phc_thread_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do k = 1, 512
      if (p(k) == c_null_char) exit
      msg = msg//p(k)
    end do
  end function device_error_message

  function get_context(ctx) result(rc)
    type(c_ptr), intent(out) :: ctx
    integer :: rc
    rc = phc_ctx_shared(phc_get_default_device(), ctx)
    if (rc /= 0) then
      ctx = c_null_ptr
      rc = err_device
    end if
  end function get_context

  function to_c_options(topts, eopts, use_endgame) result(o)
    type(tracker_options), intent(in) :: topts
    type(endgame_options), intent(in) :: eopts
    logical, intent(in) :: use_endgame
    type(c_solve_options) :: o
    o%tracker = c_tracker_options(topts%step_init, topts%step_max,  &
      topts%step_min, topts%corrector_tol, topts%contraction,  &
      topts%corrector_maxit, topts%grow_after, topts%max_steps, 0)
    o%endgame = c_endgame_options(eopts%t_start, eopts%radius_factor,  &
      eopts%radius_min, eopts%loop_radius, eopts%winding_agreement,  &
      eopts%tol, eopts%residual_tol, eopts%closure_factor,  &
      eopts%nsamples, eopts%max_winding)
    o%infinity_tol = 0.0_wp
    o%singular_cond = 0.0_wp
    o%real_tol = 0.0_wp
    o%dedup_tol = 0.0_wp
    o%singular_dedup_tol = 0.0_wp
    o%use_endgame = merge(1, 0, use_endgame)
    o%seed = 0
    o%refine_maxit = 0
    o%reserved = 0
  end function to_c_options

  subroutine device_track(sys, hom, topts, eopts, use_endgame, start, rec,  &
      status)
    !! Tracks every path whose start point is a column of `start` (N x
    !! npaths) on the GPU; rec receives one endpoint record per path.
    !! status: status_ok, err_invalid_input (rejected by the library),
    !! err_capacity (a valid system the kernels cannot hold: more than 15
    !! variables or an exponent above 255) or err_device.
    type(sparse_system), intent(in) :: sys
    type(total_degree_homotopy), intent(in) :: hom
    type(tracker_options), intent(in) :: topts
    type(endgame_options), intent(in) :: eopts
    logical, intent(in) :: use_endgame
    complex(c_double_complex), intent(in), target, contiguous :: start(:, :)
    type(endpoint_records), intent(inout), target :: rec
    integer, intent(out) :: status
    integer(c_int32_t), allocatable, target :: first(:), deg(:), hexps(:)
    real(c_double), allocatable, target :: coefs(:)
    complex(c_double_complex), allocatable, target :: gam(:), b(:), patch(:)
    type(c_system) :: cs
    type(c_homotopy) :: ch
    type(c_solve_options) :: co
    type(c_endpoints) :: ce
    integer :: n, nt, np, rc
    type(c_ptr) :: ctx

    n = sys%nvars()
    nt = sys%total_terms()
    np = size(start, 2)
    allocate(first(n + 1), deg(n), hexps((n + 1)*nt), coefs(2*nt))
    call sys%pack(first, deg, hexps, coefs)
    cs = c_system(n, nt, sys%max_exponent(), 0, c_loc(first), c_loc(deg),  &
      c_loc(hexps), c_loc(coefs))
    gam = [hom%gamma]
    b = hom%b
    patch = hom%patch
    ch = c_homotopy(c_loc(gam), c_loc(b), c_loc(patch))
    co = to_c_options(topts, eopts, use_endgame)
    ! systems the kernels cannot hold never need the device
    rc = phc_check_system(cs, co)
    if (rc == -2) then  ! PHC_ERR_CAPACITY
      status = err_capacity
      return
    end if
    status = get_context(ctx)
    if (status /= status_ok) return

    allocate(rec%status(np), rec%winding(np), rec%eg_status(np),  &
      rec%naccepted(np), rec%nrejected(np), rec%t(np), rec%eg_error(np),  &
      rec%eg_residual(np), rec%z(n + 1, np))
    ce%status = c_loc(rec%status)
    ce%t = c_loc(rec%t)
    ce%z = c_loc(rec%z)
    ce%winding = c_loc(rec%winding)
    ce%eg_status = c_loc(rec%eg_status)
    ce%eg_error = c_loc(rec%eg_error)
    ce%eg_residual = c_loc(rec%eg_residual)
    ce%naccepted = c_loc(rec%naccepted)
    ce%nrejected = c_loc(rec%nrejected)
    ce%nlin = c_null_ptr
    ce%nbe = c_null_ptr
    if (np == 0) then
      status = status_ok
      return
    end if
    rc = phc_track_paths(ctx, cs, ch, co, int(np, c_int64_t),  &
      c_loc(start), ce)
    if (rc == 0) then
      status = status_ok
    else if (rc > 0) then
      status = err_invalid_input
    else if (rc == -2) then  ! PHC_ERR_CAPACITY
      status = err_capacity
    else
      status = err_device
    end if
  end subroutine device_track

end module polyhomocont__device
