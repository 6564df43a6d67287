module polyhomocont__random
!! Seeded generator for the random constants of the homotopy.
!!
!! L'Ecuyer's combined multiplicative congruential generator (Commun. ACM
!! 31 (1988) 742) with the seeding, warm-up and derived draws of the
!! reference (src/random.f90:43-109), so that gamma, b_i and the patch --
!! and hence every path -- are the reference's for the same seed. The
!! state is an object; nothing global is touched.

  use, intrinsic :: iso_fortran_env, only : int64
  use polyhomocont__config, only : wp, pi

  implicit none

  private

  public :: rng

  integer(int64), parameter :: mod1 = 2147483563_int64
  integer(int64), parameter :: mod2 = 2147483399_int64
  integer(int64), parameter :: mul1 = 40014_int64
  integer(int64), parameter :: mul2 = 40692_int64

  type :: rng
    integer(int64) :: s1 = 1234567_int64
    integer(int64) :: s2 = 7654321_int64
  contains
    procedure :: seed => gen_seed
    procedure :: uniform => gen_uniform
    procedure :: unit_complex => gen_unit_complex
    procedure :: normal_complex => gen_normal_complex
  end type rng

contains

  subroutine gen_seed(self, seed)
    class(rng), intent(inout) :: self
    integer, intent(in) :: seed
    integer(int64) :: v
    real(wp) :: discard
    integer :: k

    v = abs(int(seed, int64))
    self%s1 = 1_int64 + mod(v, mod1 - 1_int64)
    self%s2 = 1_int64 + mod(7919_int64*v + 104729_int64, mod2 - 1_int64)
    do k = 1, 16
      discard = self%uniform()
    end do
  end subroutine gen_seed

  function gen_uniform(self) result(u)
    !! Deviate in (0, 1).
    class(rng), intent(inout) :: self
    real(wp) :: u
    integer(int64) :: d

    self%s1 = mod(mul1*self%s1, mod1)
    self%s2 = mod(mul2*self%s2, mod2)
    d = self%s1 - self%s2
    if (d < 1_int64) d = d + (mod1 - 1_int64)
    u = real(d, wp)/real(mod1, wp)
  end function gen_uniform

  function gen_unit_complex(self) result(z)
    class(rng), intent(inout) :: self
    complex(wp) :: z
    real(wp) :: angle

    angle = 2.0_wp*pi*self%uniform()
    z = cmplx(cos(angle), sin(angle), kind=wp)
  end function gen_unit_complex

  function gen_normal_complex(self) result(z)
    !! Box-Muller: both parts normal with variance 1/2.
    class(rng), intent(inout) :: self
    complex(wp) :: z
    real(wp) :: radius, angle

    radius = sqrt(-log(self%uniform()))
    angle = 2.0_wp*pi*self%uniform()
    z = cmplx(radius*cos(angle), radius*sin(angle), kind=wp)
  end function gen_normal_complex

end module polyhomocont__random
