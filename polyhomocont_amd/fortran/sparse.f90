module polyhomocont__sparse
!! sparse_system: the built-in sparse polynomial system, with the
!! reference's builder (src/sparse.f90:65-195) -- equations are appended
!! one at a time as lists of complex coefficients and exponent vectors;
!! zero terms are dropped, shapes/negative exponents/constant equations
!! rejected, and the homogenized exponents (w exponent d_i - |a| last) kept
!! next to the affine ones.
!!
!! Its homogenized form is what the GPU evaluates: `pack` exports exactly
!! the arrays of a phc_system (include/phc.h). The host evaluators below
!! (prefix/suffix products over a power table, src/sparse.f90:223-280) are
!! used by the classification of endpoints and by tests.

  use, intrinsic :: iso_c_binding, only : c_int32_t, c_double
  use polyhomocont__config, only : wp, status_ok, err_invalid_input
  use polyhomocont__system, only : poly_system

  implicit none

  private

  public :: sparse_system

  type, extends(poly_system) :: sparse_system
    private
    integer :: n = 0
    integer :: neq = 0
    integer :: maxexp = 0
    integer, allocatable :: first(:)
    integer, allocatable :: deg(:)
    integer, allocatable :: exps(:, :)
    integer, allocatable :: hexps(:, :)
    complex(wp), allocatable :: coefs(:)
  contains
    procedure :: init => sp_init
    procedure, private :: sp_add_complex
    procedure, private :: sp_add_real
    generic :: add_equation => sp_add_complex, sp_add_real
    procedure :: nequations => sp_nequations
    procedure :: nterms => sp_nterms
    procedure :: nvars => sp_nvars
    procedure :: degrees => sp_degrees
    procedure :: evaluate => sp_evaluate
    procedure :: evaluate_homogeneous => sp_evaluate_homogeneous
    procedure :: total_terms => sp_total_terms
    procedure :: max_exponent => sp_max_exponent
    procedure :: pack => sp_pack
  end type sparse_system

contains

  subroutine sp_init(self, n)
    class(sparse_system), intent(inout) :: self
    integer, intent(in) :: n

    self%n = n
    self%neq = 0
    self%maxexp = 0
    self%first = [1]
    self%deg = [integer ::]
    if (allocated(self%exps)) deallocate(self%exps)
    allocate(self%exps(n, 0))
    if (allocated(self%hexps)) deallocate(self%hexps)
    allocate(self%hexps(n + 1, 0))
    self%coefs = [complex(wp) ::]
  end subroutine sp_init

  subroutine sp_add_complex(self, coefs, exps, status)
    class(sparse_system), intent(inout) :: self
    complex(wp), intent(in) :: coefs(:)
    integer, intent(in) :: exps(:, :)
    integer, intent(out) :: status
    logical :: nonzero(size(coefs))
    integer, allocatable :: e(:, :), h(:, :)
    integer :: kept, d, k, j, tot

    status = err_invalid_input
    if (.not. allocated(self%first)) return
    if (self%neq >= self%n) return
    if (size(exps, 1) /= self%n .or. size(exps, 2) /= size(coefs)) return
    if (any(exps < 0)) return
    nonzero = coefs /= (0.0_wp, 0.0_wp)
    d = 0
    do k = 1, size(coefs)
      if (nonzero(k)) d = max(d, sum(exps(:, k)))
    end do
    if (d < 1) return

    kept = count(nonzero)
    tot = size(self%coefs)
    allocate(e(self%n, tot + kept), h(self%n + 1, tot + kept))
    e(:, 1:tot) = self%exps
    h(:, 1:tot) = self%hexps
    j = tot
    do k = 1, size(coefs)
      if (.not. nonzero(k)) cycle
      j = j + 1
      e(:, j) = exps(:, k)
      h(1:self%n, j) = exps(:, k)
      h(self%n + 1, j) = d - sum(exps(:, k))
      self%maxexp = max(self%maxexp, maxval(h(:, j)))
    end do
    call move_alloc(e, self%exps)
    call move_alloc(h, self%hexps)
    self%coefs = [self%coefs, pack(coefs, nonzero)]
    self%neq = self%neq + 1
    self%first = [self%first, self%first(self%neq) + kept]
    self%deg = [self%deg, d]
    status = status_ok
  end subroutine sp_add_complex

  subroutine sp_add_real(self, coefs, exps, status)
    class(sparse_system), intent(inout) :: self
    real(wp), intent(in) :: coefs(:)
    integer, intent(in) :: exps(:, :)
    integer, intent(out) :: status

    call self%sp_add_complex(cmplx(coefs, 0.0_wp, kind=wp), exps, status)
  end subroutine sp_add_real

  function sp_nequations(self) result(m)
    class(sparse_system), intent(in) :: self
    integer :: m
    m = self%neq
  end function sp_nequations

  function sp_nterms(self, i) result(m)
    class(sparse_system), intent(in) :: self
    integer, intent(in) :: i
    integer :: m
    m = self%first(i + 1) - self%first(i)
  end function sp_nterms

  function sp_nvars(self) result(n)
    class(sparse_system), intent(in) :: self
    integer :: n
    n = self%n
  end function sp_nvars

  function sp_degrees(self) result(d)
    class(sparse_system), intent(in) :: self
    integer, allocatable :: d(:)
    d = self%deg
  end function sp_degrees

  function sp_total_terms(self) result(m)
    class(sparse_system), intent(in) :: self
    integer :: m
    m = 0
    if (allocated(self%coefs)) m = size(self%coefs)
  end function sp_total_terms

  function sp_max_exponent(self) result(m)
    class(sparse_system), intent(in) :: self
    integer :: m
    m = self%maxexp
  end function sp_max_exponent

  subroutine sp_pack(self, first, deg, hexps, coefs)
    !! The homogenized system as a phc_system: 0-based term offsets,
    !! degrees, hexps column-major (N x nterms) and interleaved coefficients.
    class(sparse_system), intent(in) :: self
    integer(c_int32_t), intent(out) :: first(:), deg(:), hexps(:)
    real(c_double), intent(out) :: coefs(:)
    integer :: k, m

    first(1:self%neq + 1) = self%first(1:self%neq + 1) - 1
    deg(1:self%neq) = self%deg(1:self%neq)
    m = size(self%coefs)
    hexps(1:(self%n + 1)*m) = reshape(self%hexps, [(self%n + 1)*m])
    do k = 1, m
      coefs(2*k - 1) = real(self%coefs(k), wp)
      coefs(2*k) = aimag(self%coefs(k))
    end do
  end subroutine sp_pack

  subroutine sp_evaluate(self, x, f, jac)
    class(sparse_system), intent(inout) :: self
    complex(wp), intent(in) :: x(:)
    complex(wp), intent(out) :: f(:)
    complex(wp), intent(out) :: jac(:, :)
    call term_sums(self%neq, self%first, self%exps, self%coefs, self%maxexp,  &
      x, f, jac)
  end subroutine sp_evaluate

  subroutine sp_evaluate_homogeneous(self, z, f, jac)
    !! Exact homogenized values, also at w = 0.
    class(sparse_system), intent(inout) :: self
    complex(wp), intent(in) :: z(:)
    complex(wp), intent(out) :: f(:)
    complex(wp), intent(out) :: jac(:, :)
    call term_sums(self%neq, self%first, self%hexps, self%coefs,  &
      self%maxexp, z, f, jac)
  end subroutine sp_evaluate_homogeneous

  subroutine term_sums(neq, first, exps, coefs, maxexp, v, f, jac)
    !! f_i = sum_k c_k v^e_k and jac(i, j) = df_i/dv_j. The derivative of a
    !! term with respect to v_j is c e_j v_j^(e_j-1) times the product of
    !! the other factors, formed as (prefix up to j-1) * (suffix from j+1)
    !! so no division by v_j occurs.
    integer, intent(in) :: neq, maxexp
    integer, intent(in) :: first(:), exps(:, :)
    complex(wp), intent(in) :: coefs(:), v(:)
    complex(wp), intent(out) :: f(:), jac(:, :)
    complex(wp) :: pw(0:maxexp, size(v)), lead(0:size(v)),  &
      trail(1:size(v)+1), c
    integer :: m, i, j, k, e, p

    m = size(v)
    pw(0, :) = (1.0_wp, 0.0_wp)
    do p = 1, maxexp
      pw(p, :) = pw(p - 1, :)*v
    end do
    f = (0.0_wp, 0.0_wp)
    jac = (0.0_wp, 0.0_wp)
    do i = 1, neq
      do k = first(i), first(i + 1) - 1
        c = coefs(k)
        lead(0) = (1.0_wp, 0.0_wp)
        do j = 1, m
          lead(j) = lead(j - 1)*pw(exps(j, k), j)
        end do
        trail(m + 1) = (1.0_wp, 0.0_wp)
        do j = m, 1, -1
          trail(j) = pw(exps(j, k), j)*trail(j + 1)
        end do
        f(i) = f(i) + c*lead(m)
        do j = 1, m
          e = exps(j, k)
          if (e > 0) jac(i, j) = jac(i, j) + c*e*pw(e - 1, j)*lead(j - 1)  &
            *trail(j + 1)
        end do
      end do
    end do
  end subroutine term_sums

end module polyhomocont__sparse
