module polyhomocont__homotopy
!! Total-degree homotopy with the gamma trick on a random affine patch
!! (the reference's total_degree_homotopy, src/homotopy.f90:78-231):
!!   H_i(z, t)   = gamma t G_i(z) + (1 - t) F^h_i(z),  i = 1..n,
!!   H_{n+1}(z)  = a . z - 1,
!!   G_i         = z_i^d_i - b_i w^d_i,
!! in projective coordinates z = (z_1..z_n, w). init draws gamma, b and the
!! patch from the generator in the reference's order; start_solution gives
!! the k-th root combination (mixed radix, first variable fastest) scaled
!! onto the patch. evaluate is the host evaluator (classification,
!! refinement, CPU tracker); the GPU assembles the same rows per lane.

  use, intrinsic :: iso_fortran_env, only : int64
  use polyhomocont__config, only : wp, pi, status_ok, err_invalid_input
  use polyhomocont__random, only : rng
  use polyhomocont__system, only : poly_system

  implicit none

  private

  public :: homotopy, total_degree_homotopy

  type, abstract :: homotopy
  contains
    procedure(nunknowns_if), deferred :: nunknowns
    procedure(hevaluate_if), deferred :: evaluate
  end type homotopy

  abstract interface
    function nunknowns_if(self) result(nn)
      import :: homotopy
      class(homotopy), intent(in) :: self
      integer :: nn
    end function nunknowns_if
    subroutine hevaluate_if(self, z, t, h, hz, ht)
      import :: homotopy, wp
      class(homotopy), intent(inout) :: self
      complex(wp), intent(in) :: z(:)
      complex(wp), intent(in) :: t
      complex(wp), intent(out) :: h(:), hz(:, :), ht(:)
    end subroutine hevaluate_if
  end interface

  type, extends(homotopy) :: total_degree_homotopy
    class(poly_system), pointer :: sys => null()
    integer :: n = 0
    integer, allocatable :: deg(:)
    complex(wp) :: gamma = (1.0_wp, 0.0_wp)
    complex(wp), allocatable :: b(:)
    complex(wp), allocatable :: patch(:)
    integer :: nstart = 0
  contains
    procedure :: init => td_init
    procedure :: start_solution => td_start
    procedure :: nunknowns => td_nunknowns
    procedure :: evaluate => td_evaluate
  end type total_degree_homotopy

contains

  subroutine td_init(self, sys, gen, status)
    !! err_invalid_input for a non-square system, a degree < 1 or a Bezout
    !! number beyond the default integer range.
    class(total_degree_homotopy), intent(inout) :: self
    class(poly_system), intent(inout), target :: sys
    type(rng), intent(inout) :: gen
    integer, intent(out) :: status
    integer(int64) :: bezout
    integer :: i

    status = err_invalid_input
    self%sys => sys
    self%n = sys%nvars()
    self%deg = sys%degrees()
    if (self%n < 1 .or. size(self%deg) /= self%n) return
    if (any(self%deg < 1)) return
    bezout = 1_int64
    do i = 1, self%n
      bezout = bezout*int(self%deg(i), int64)
      if (bezout > int(huge(1), int64)) return
    end do
    self%nstart = int(bezout)

    self%gamma = gen%unit_complex()
    if (allocated(self%b)) deallocate(self%b)
    allocate(self%b(self%n))
    do i = 1, self%n
      self%b(i) = gen%unit_complex()
    end do
    if (allocated(self%patch)) deallocate(self%patch)
    allocate(self%patch(self%n + 1))
    do i = 1, self%n + 1
      self%patch(i) = gen%normal_complex()
    end do
    self%patch = self%patch/sqrt(sum(abs(self%patch)**2))
    status = status_ok
  end subroutine td_init

  subroutine td_start(self, k, z)
    class(total_degree_homotopy), intent(in) :: self
    integer, intent(in) :: k
    complex(wp), intent(out) :: z(:)
    integer :: left, j, i
    real(wp) :: angle

    left = k - 1
    do i = 1, self%n
      j = mod(left, self%deg(i))
      left = left/self%deg(i)
      angle = (atan2(aimag(self%b(i)), real(self%b(i), wp))  &
        + 2.0_wp*pi*real(j, wp))/real(self%deg(i), wp)
      z(i) = cmplx(cos(angle), sin(angle), kind=wp)
    end do
    z(self%n + 1) = (1.0_wp, 0.0_wp)
    z = z/sum(self%patch*z)
  end subroutine td_start

  function td_nunknowns(self) result(nn)
    class(total_degree_homotopy), intent(in) :: self
    integer :: nn
    nn = self%n + 1
  end function td_nunknowns

  subroutine td_evaluate(self, z, t, h, hz, ht)
    class(total_degree_homotopy), intent(inout) :: self
    complex(wp), intent(in) :: z(:)
    complex(wp), intent(in) :: t
    complex(wp), intent(out) :: h(:), hz(:, :), ht(:)
    complex(wp) :: fh(self%n), jf(self%n, self%n + 1)
    complex(wp) :: gt, onemt, zpow, wpow, g, w
    integer :: i, d, n

    n = self%n
    w = z(n + 1)
    call self%sys%evaluate_homogeneous(z, fh, jf)
    gt = self%gamma*t
    onemt = 1.0_wp - t
    hz(1:n, :) = onemt*jf
    do i = 1, n
      d = self%deg(i)
      zpow = z(i)**(d - 1)
      wpow = w**(d - 1)
      g = zpow*z(i) - self%b(i)*wpow*w
      h(i) = gt*g + onemt*fh(i)
      ht(i) = self%gamma*g - fh(i)
      hz(i, i) = hz(i, i) + gt*d*zpow
      hz(i, n + 1) = hz(i, n + 1) - gt*self%b(i)*d*wpow
    end do
    h(n + 1) = sum(self%patch*z) - 1.0_wp
    hz(n + 1, :) = self%patch
    ht(n + 1) = (0.0_wp, 0.0_wp)
  end subroutine td_evaluate

end module polyhomocont__homotopy
