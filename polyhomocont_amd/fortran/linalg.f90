module polyhomocont__linalg
!! Small dense complex linear algebra used on the host: the classification
!! and refinement of endpoints (condition numbers, Newton steps) and the
!! CPU tracker for user-supplied systems. Same pivot rule and operation
!! order as the reference (src/linalg.f90:26-189), so host results agree
!! with it bit for bit; the GPU kernels implement the same factorization
!! row-per-lane.

  use polyhomocont__config, only : wp, status_ok, err_singular_matrix

  implicit none

  private

  public :: lu_factor, lu_solve, solve_linear, cond1, norm_inf

contains

  subroutine lu_factor(a, ipiv, status)
    !! P*A = L*U in place, pivot = first entry of largest modulus in the
    !! column; row i was exchanged with row ipiv(i) at step i.
    complex(wp), intent(inout) :: a(:, :)
    integer, intent(out) :: ipiv(:)
    integer, intent(out) :: status
    complex(wp) :: row(size(a, 2)), pivot
    real(wp) :: best, mag
    integer :: m, col, r, prow

    status = status_ok
    m = size(a, 1)
    do col = 1, m
      prow = col
      best = abs(a(col, col))
      do r = col + 1, m
        mag = abs(a(r, col))
        if (mag > best) then
          best = mag
          prow = r
        end if
      end do
      ipiv(col) = prow
      if (best == 0.0_wp) then
        status = err_singular_matrix
        return
      end if
      if (prow /= col) then
        row = a(col, :)
        a(col, :) = a(prow, :)
        a(prow, :) = row
      end if
      pivot = a(col, col)
      do r = col + 1, m
        a(r, col) = a(r, col)/pivot
        a(r, col+1:m) = a(r, col+1:m) - a(r, col)*a(col, col+1:m)
      end do
    end do
  end subroutine lu_factor

  subroutine lu_solve(a, ipiv, b)
    !! Solves A x = b with the factors of lu_factor; x overwrites b.
    complex(wp), intent(in) :: a(:, :)
    integer, intent(in) :: ipiv(:)
    complex(wp), intent(inout) :: b(:)
    complex(wp) :: swap
    integer :: m, r

    m = size(a, 1)
    do r = 1, m
      if (ipiv(r) /= r) then
        swap = b(r)
        b(r) = b(ipiv(r))
        b(ipiv(r)) = swap
      end if
    end do
    do r = 2, m
      b(r) = b(r) - sum(a(r, 1:r-1)*b(1:r-1))
    end do
    do r = m, 1, -1
      b(r) = (b(r) - sum(a(r, r+1:m)*b(r+1:m)))/a(r, r)
    end do
  end subroutine lu_solve

  subroutine solve_linear(a, b, x, status)
    complex(wp), intent(in) :: a(:, :)
    complex(wp), intent(in) :: b(:)
    complex(wp), intent(out) :: x(:)
    integer, intent(out) :: status
    complex(wp) :: f(size(a, 1), size(a, 2))
    integer :: p(size(a, 1))

    f = a
    call lu_factor(f, p, status)
    if (status /= status_ok) return
    x = b
    call lu_solve(f, p, x)
  end subroutine solve_linear

  function cond1(a) result(c)
    !! ||A||_1 ||A^-1||_1 with the inverse formed column by column;
    !! huge(1.0_wp) if A is exactly singular.
    complex(wp), intent(in) :: a(:, :)
    real(wp) :: c
    complex(wp) :: f(size(a, 1), size(a, 2)), e(size(a, 1))
    integer :: p(size(a, 1))
    real(wp) :: na, ninv
    integer :: m, col, status

    m = size(a, 1)
    na = 0.0_wp
    do col = 1, m
      na = max(na, sum(abs(a(:, col))))
    end do
    f = a
    call lu_factor(f, p, status)
    if (status /= status_ok) then
      c = huge(1.0_wp)
      return
    end if
    ninv = 0.0_wp
    do col = 1, m
      e = (0.0_wp, 0.0_wp)
      e(col) = (1.0_wp, 0.0_wp)
      call lu_solve(f, p, e)
      ninv = max(ninv, sum(abs(e)))
    end do
    c = na*ninv
  end function cond1

  pure function norm_inf(v) result(r)
    complex(wp), intent(in) :: v(:)
    real(wp) :: r
    integer :: k

    r = 0.0_wp
    do k = 1, size(v)
      r = max(r, abs(v(k)))
    end do
  end function norm_inf

end module polyhomocont__linalg
