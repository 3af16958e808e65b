! spc_random.f90 -- xoroshiro128+ generator with the interface the driver needs
! (set_seed, next, normal): the stream of the reference generator
! (m_random.f90:90-198, 321-340), used to draw the per-sample seeds and, for
! trace_sampling_serial, the trace vectors themselves.
module spc_random
    use, intrinsic :: iso_fortran_env, only: dp => real64, i8 => int64
    implicit none
    private

    type, public :: rng_t
        integer(i8) :: s(2) = [123456789_i8, 987654321_i8]
        real(dp)    :: stored = 0.0_dp
        logical     :: have = .false.
    contains
        procedure :: set_seed
        procedure :: next
        procedure :: unif_01
        procedure :: normal
    end type

contains

    ! x + y modulo 2^64 on two's-complement integers, without signed overflow
    pure function wrap_add(x, y) result(z)
        integer(i8), intent(in) :: x, y
        integer(i8) :: z
        integer(i8), parameter :: mask = 4294967295_i8   ! 2^32 - 1
        integer(i8) :: lo, hi
        lo = iand(x, mask) + iand(y, mask)
        hi = iand(ishft(x, -32), mask) + iand(ishft(y, -32), mask) + ishft(lo, -32)
        z = ior(ishft(hi, 32), iand(lo, mask))
    end function

    function next(self) result(res)
        class(rng_t), intent(inout) :: self
        integer(i8) :: res, a, b
        a = self%s(1)
        b = self%s(2)
        res = wrap_add(a, b)
        b = ieor(b, a)
        self%s(1) = ieor(ieor(ishftc(a, 55), b), ishft(b, 14))
        self%s(2) = ishftc(b, 36)
    end function

    subroutine set_seed(self, seed)
        class(rng_t), intent(inout) :: self
        integer(i8), intent(in) :: seed(2)
        ! jump polynomial of xoroshiro128+ (2^64 steps)
        integer(i8), parameter :: jmp(2) = [-4707382666127344949_i8, -2852180941702784734_i8]
        integer(i8) :: t(2), dummy
        integer :: i, b
        self%s = seed
        t = 0
        do i = 1, 2
            do b = 0, 63
                if (btest(jmp(i), b)) t = ieor(t, self%s)
                dummy = self%next()
            end do
        end do
        self%s = t
    end subroutine

    function unif_01(self) result(u)
        class(rng_t), intent(inout) :: self
        real(dp) :: u
        integer(i8) :: x
        x = ior(ishft(1023_i8, 52), ishft(self%next(), -12))
        u = transfer(x, u) - 1.0_dp
    end function

    ! Marsaglia polar method; the second value is kept for the next call
    function normal(self) result(z)
        class(rng_t), intent(inout) :: self
        real(dp) :: z, a, b, s
        if (self%have) then
            self%have = .false.
            z = self%stored
            return
        end if
        do
            a = 2 * self%unif_01() - 1
            b = 2 * self%unif_01() - 1
            s = a * a + b * b
            if (s < 1.0_dp .and. s > 0.0_dp) exit
        end do
        s = sqrt(-2 * log(s) / s)
        z = a * s
        self%stored = b * s
        self%have = .true.
    end function

end module spc_random
