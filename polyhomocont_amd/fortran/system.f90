module polyhomocont__system
!! Abstract square polynomial system F: C^n -> C^n (the reference
!! interface, src/system.f90:30-69). Extensions provide nvars, degrees and
!! the affine evaluator; evaluate_homogeneous defaults to homogenizing the
!! affine evaluator,
!!   F^h_i = w^d_i F_i(x),  dF^h_i/dz_j = w^(d_i-1) dF_i/dx_j,
!!   dF^h_i/dw = w^(d_i-1) (d_i F_i(x) - sum_j x_j dF_i/dx_j),  x = z/w,
!! with w = 0 replaced by epsilon * max(||z||, tiny) (src/system.f90:73-117).
!! Systems with user evaluators are tracked by the CPU tracker; the built-in
!! sparse_system overrides evaluate_homogeneous and is tracked on the GPU.

  use polyhomocont__config, only : wp
  use polyhomocont__linalg, only : norm_inf

  implicit none

  private

  public :: poly_system

  type, abstract :: poly_system
  contains
    procedure(nvars_if), deferred :: nvars
    procedure(degrees_if), deferred :: degrees
    procedure(evaluate_if), deferred :: evaluate
    procedure :: evaluate_homogeneous => homogenized_evaluate
  end type poly_system

  abstract interface
    function nvars_if(self) result(n)
      import :: poly_system
      class(poly_system), intent(in) :: self
      integer :: n
    end function nvars_if
    function degrees_if(self) result(d)
      import :: poly_system
      class(poly_system), intent(in) :: self
      integer, allocatable :: d(:)
    end function degrees_if
    subroutine evaluate_if(self, x, f, jac)
      import :: poly_system, wp
      class(poly_system), intent(inout) :: self
      complex(wp), intent(in) :: x(:)
      complex(wp), intent(out) :: f(:)
      complex(wp), intent(out) :: jac(:, :)
    end subroutine evaluate_if
  end interface

contains

  subroutine homogenized_evaluate(self, z, f, jac)
    class(poly_system), intent(inout) :: self
    complex(wp), intent(in) :: z(:)
    complex(wp), intent(out) :: f(:)
    complex(wp), intent(out) :: jac(:, :)
    integer, allocatable :: deg(:)
    complex(wp), allocatable :: x(:), fx(:), jx(:, :)
    complex(wp) :: w, wpow
    integer :: n, i

    n = size(z) - 1
    deg = self%degrees()
    allocate(fx(n), jx(n, n))
    w = z(n + 1)
    if (w == (0.0_wp, 0.0_wp)) w = cmplx(epsilon(1.0_wp)*  &
      max(norm_inf(z), tiny(1.0_wp)), 0.0_wp, kind=wp)
    x = z(1:n)/w
    call self%evaluate(x, fx, jx)
    do i = 1, n
      wpow = w**(deg(i) - 1)
      f(i) = wpow*w*fx(i)
      jac(i, 1:n) = wpow*jx(i, :)
      jac(i, n + 1) = wpow*(deg(i)*fx(i) - sum(x*jx(i, :)))
    end do
  end subroutine homogenized_evaluate

end module polyhomocont__system
