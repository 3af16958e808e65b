! ref_shim.f90 -- C entry points over the reference modules that compile on
! their own (m_random.f90, observables.f90), used only to pin the oracle.
! Test infrastructure: built into oracle/_ref/, never part of the product.
module ref_shim
    use, intrinsic :: iso_c_binding
    use m_random
    use class_observables
    implicit none
contains

    ! m_random.f90:90-98 set_seed, then n calls of next()
    subroutine ref_rng_next(seed, n, out) bind(C, name='ref_rng_next')
        integer(c_int64_t), intent(in)        :: seed(2)
        integer(c_int), value                 :: n
        integer(c_int64_t), intent(out)       :: out(n)
        type(rng_t) :: rng
        integer :: i
        call rng%set_seed(seed)
        do i = 1, n
            out(i) = rng%int_8()
        end do
    end subroutine

    ! set_seed, then n calls of unif_01 / normal / circle / sphere (kind 0..3)
    subroutine ref_rng_draw(seed, kind, n, out) bind(C, name='ref_rng_draw')
        integer(c_int64_t), intent(in) :: seed(2)
        integer(c_int), value          :: kind, n
        real(c_double), intent(out)    :: out(*)
        type(rng_t) :: rng
        integer :: i
        call rng%set_seed(seed)
        do i = 1, n
            select case (kind)
            case (0)
                out(i) = rng%unif_01()
            case (1)
                out(i) = rng%normal()
            case (2)
                out(2*i-1:2*i) = rng%circle(1.0_c_double)
            case (3)
                out(3*i-2:3*i) = rng%sphere(1.0_c_double)
            end select
        end do
    end subroutine

    ! observables.f90:74-101 get_kinetics
    subroutine ref_get_kinetics(n, ps, pt0, ptp, ptm, dt, ks, kt, iden, phis, phit) &
            bind(C, name='ref_get_kinetics')
        integer(c_int), value       :: n
        real(c_double), intent(in)  :: ps(n), pt0(n), ptp(n), ptm(n)
        real(c_double), value       :: dt, ks, kt
        real(c_double), intent(out) :: iden(n), phis(n/2), phit(n/2)
        type(observables) :: res
        call res%malloc(n)
        res%P_S = ps
        res%P_T0 = pt0
        res%P_Tp = ptp
        res%P_Tm = ptm
        call res%get_kinetics(dt, ks, kt)
        iden = res%iden
        phis = res%Phi_S
        phit = res%Phi_T
    end subroutine

end module ref_shim
