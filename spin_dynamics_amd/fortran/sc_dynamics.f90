! sc_dynamics.f90 -- module sc_dynamics of src/sc (sc_dynamics.f90:17-100):
! the semiclassical singlet and triplet correlation functions, computed on the
! device by spc_semiclassical_run (classical electron and nuclear spin
! vectors, isotropic couplings a_iso as src/sc uses).
module sc_dynamics
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use, intrinsic :: iso_c_binding
    use spinchem_capi
    use variables
    use class_observables, only: observables, write_data
    implicit none
    private
    public :: semi_classical

    integer(c_int32_t), allocatable, target, save :: mult_(:)
    real(c_double), allocatable, target, save :: aiso_(:)

contains

    ! N_samples samples seeded from sim%seed (two next() per sample), N_steps =
    ! ceiling((t_end+dt)/dt) steps of dt (sc_dynamics.f90:47, 78): P_S(t) =
    ! 4<P_S(0) P_S(t)>, P_T(t) = 4<P_S(0) P_T(t)> (N_steps + 1 values), the
    ! Simpson yields of get_kinetics, and the files S_prob.data, T_prob.data,
    ! total_prob.data, S_yield.data, T_yield.data.
    subroutine semi_classical(sim, sys, res, out)
        type(sim_param), intent(inout)     :: sim
        type(sys_param), intent(in)        :: sys
        type(observables), intent(out)     :: res
        character(*), optional, intent(in) :: out
        type(spc_sc_system) :: s
        integer :: n1, n2, k, h
        integer(c_int64_t) :: n_steps, seed(2)
        real(dp), allocatable :: pt(:), phis(:), phit(:)
        n1 = size(sys%e1%g_I)
        n2 = size(sys%e2%g_I)
        if (allocated(mult_)) deallocate(mult_, aiso_)
        allocate(mult_(max(n1 + n2, 1)), aiso_(max(n1 + n2, 1)))
        do k = 1, n1
            mult_(k) = sys%e1%g_I(k)
            aiso_(k) = sys%e1%a_iso(k)
        end do
        do k = 1, n2
            mult_(n1 + k) = sys%e2%g_I(k)
            aiso_(n1 + k) = sys%e2%a_iso(k)
        end do
        s%n_nuc1 = n1
        s%n_nuc2 = n2
        s%mult = c_loc(mult_)
        s%a_iso = c_loc(aiso_)
        s%B = -sys%e1%w
        s%field_dir = sys%field_dir
        s%J = sys%J
        s%kS = sys%kS
        s%kT = sys%kT
        n_steps = spc_n_steps(sim%t_end, sim%dt)
        call res%malloc(int(n_steps + 1))
        allocate(pt(n_steps + 1), phis(n_steps + 1), phit(n_steps + 1))
        seed = sim%seed
        call check(spc_semiclassical_run(device_context(), s, int(sim%N_samples, c_int64_t), seed, sim%dt, n_steps, &
                                         res%P_S, pt, phis, phit))
        res%P_T0 = pt
        res%P_Tp = 0.0_dp
        res%P_Tm = 0.0_dp
        h = int(n_steps + 1) / 2
        res%iden = res%P_S + pt
        res%Phi_S = phis(1:h)
        res%Phi_T = phit(1:h)
        if (present(out)) then
            call write_data(res%P_S, out, 'S_prob.data')
            call write_data(pt, out, 'T_prob.data')
            call write_data(res%iden, out, 'total_prob.data')
            call write_data(res%Phi_S, out, 'S_yield.data')
            call write_data(res%Phi_T, out, 'T_yield.data')
        end if
    end subroutine

end module sc_dynamics
