! dynamics.f90 -- the engines of module dynamics (qm_dynamics.f90:15-19) with
! their names and argument lists: trace_sampling_para, trace_sampling and
! exact_dynamics.  The propagation runs on the device through the C-ABI; the
! outputs are the sample-averaged traces plus, as get_kinetics computes them,
! the total population and the Simpson yields.
module dynamics
    use, intrinsic :: iso_fortran_env, only: dp => real64, i8 => int64
    use, intrinsic :: iso_c_binding
    use spinchem_capi
    use spc_random
    use variables
    use class_observables
    implicit none
    private
    public :: trace_sampling_para, trace_sampling, exact_dynamics, sample_SUZ

    ! flattened tensors kept alive while the library reads them
    integer(c_int32_t), allocatable, target, save :: mult_(:)
    real(c_double), allocatable, target, save :: hfc_(:)

contains

    integer function state_index(s)
        character(*), intent(in) :: s
        select case (trim(s))
        case ('singlet'); state_index = 0
        case ('T0'); state_index = 1
        case ('Tp'); state_index = 2
        case ('Tm'); state_index = 3
        case default
            print *, 'Error: Invalid initial radical pair state entered.'
            stop
        end select
    end function

    ! sys_param (field B = -w, main.f90:55-56) -> spc_system_set, or
    ! spc_system_set_block for a symmetry block (multiplets of equivalent
    ! spin-1/2 nuclei)
    subroutine set_system(sys, block)
        type(sys_param), intent(in) :: sys
        logical, intent(in), optional :: block
        type(spc_spin_system) :: s
        integer :: n1, n2, k, a, b
        n1 = size(sys%e1%g_I)
        n2 = size(sys%e2%g_I)
        if (allocated(mult_)) deallocate(mult_, hfc_)
        allocate(mult_(max(n1 + n2, 1)), hfc_(max(9 * (n1 + n2), 1)))
        do k = 1, n1 + n2
            if (k <= n1) then
                mult_(k) = sys%e1%g_I(k)
            else
                mult_(k) = sys%e2%g_I(k - n1)
            end if
            do b = 1, 3
                do a = 1, 3
                    if (k <= n1) then
                        hfc_(9*(k-1) + 3*(b-1) + a) = sys%e1%A(k, a, b)
                    else
                        hfc_(9*(k-1) + 3*(b-1) + a) = sys%e2%A(k - n1, a, b)
                    end if
                end do
            end do
        end do
        s%n_nuc1 = n1
        s%n_nuc2 = n2
        s%mult = c_loc(mult_)
        s%hfc = c_loc(hfc_)
        s%J = sys%J
        s%D = reshape(sys%D, [9])
        s%kS = sys%kS
        s%kT = sys%kT
        s%B = -sys%e1%w
        s%field_dir = sys%field_dir
        s%iso_rule = 1
        if (present(block)) then
            if (block) then
                call check(spc_system_set_block(device_context(), s))
                return
            end if
        end if
        call check(spc_system_set(device_context(), s))
    end subroutine

    function sim_c(sim, n, kind) result(c)
        type(sim_param), intent(in) :: sim
        integer, intent(in) :: n, kind
        type(spc_sim_params) :: c
        c%n_samples = n
        c%n_krylov = sim%N_krylov
        c%t_end = sim%t_end
        c%dt = sim%dt
        c%tol = sim%tol
        c%init_state = state_index(sim%init_state)
        c%vec_kind = kind
    end function

    subroutine finish(sys, res, phis, phit, out)
        type(sys_param), intent(in) :: sys
        type(observables), intent(inout) :: res
        real(dp), intent(in) :: phis(:), phit(:)
        character(*), intent(in), optional :: out
        integer :: h
        h = size(res%P_S) / 2
        res%iden = res%P_S + res%P_T0 + res%P_Tp + res%P_Tm
        res%Phi_S = phis(1:h)
        res%Phi_T = phit(1:h)
        if (present(out)) call res%output(out)
    end subroutine

    ! qm_dynamics.f90:257-384: per-sample seeds drawn from rng, then the
    ! device propagation of all samples.  block: sys is a symmetry block.
    subroutine trace_sampling_para(sys, sim, rng, res, out, block)
        type(sys_param), intent(in)            :: sys
        type(sim_param), intent(inout)         :: sim
        type(rng_t), intent(inout)             :: rng
        type(observables), intent(out)         :: res
        character(*), optional, intent(in)     :: out
        logical, optional, intent(in)          :: block
        integer(c_int64_t), allocatable :: seeds(:,:)
        real(dp), allocatable :: phis(:), phit(:)
        integer(c_int64_t) :: n_out
        integer :: i, kind
        call set_system(sys, block)
        n_out = spc_n_steps(sim%t_end, sim%dt) + 1
        call res%malloc(int(n_out))
        allocate(seeds(2, max(sim%N_samples, 1)), phis(n_out), phit(n_out))
        do i = 1, sim%N_samples
            seeds(1, i) = rng%next()
            seeds(2, i) = rng%next()
        end do
        kind = SPC_VEC_SUZ
        if (sim%vec_kind == 'phase') kind = SPC_VEC_PHASE
        call check(spc_trace_sample_seeds(device_context(), sim_c(sim, sim%N_samples, kind), seeds, &
                                          res%P_S, res%P_T0, res%P_Tp, res%P_Tm, n_out, phis, phit))
        call finish(sys, res, phis, phit, out)
    end subroutine

    ! qm_dynamics.f90:416-430
    subroutine sample_SUZ(rng, Z_ket)
        type(rng_t), intent(inout) :: rng
        complex(dp), intent(inout) :: Z_ket(:)
        integer :: i
        real(dp) :: re, im
        do i = 1, size(Z_ket)
            re = rng%normal()
            im = rng%normal()
            Z_ket(i) = cmplx(re, im, kind=dp)
        end do
        Z_ket = Z_ket / sqrt(dot_product(Z_ket, Z_ket))
    end subroutine

    ! qm_dynamics.f90:139-255: every sample drawn in turn from the master rng
    subroutine trace_sampling(sys, sim, rng, res, out)
        type(sys_param), intent(in)            :: sys
        type(sim_param), intent(inout)         :: sim
        type(rng_t), intent(inout)             :: rng
        type(observables), intent(out)         :: res
        character(*), optional, intent(in)     :: out
        complex(c_double_complex), allocatable, target :: tv(:,:)
        real(dp), allocatable :: phis(:), phit(:)
        integer(c_int64_t) :: n_out, seed(2)
        integer :: i, Z
        call set_system(sys)
        Z = sys%Z1 * sys%Z2
        n_out = spc_n_steps(sim%t_end, sim%dt) + 1
        call res%malloc(int(n_out))
        allocate(tv(Z, max(sim%N_samples, 1)), phis(n_out), phit(n_out))
        do i = 1, sim%N_samples
            call sample_SUZ(rng, tv(:, i))
        end do
        seed = 0
        call check(spc_trace_sample(device_context(), sim_c(sim, sim%N_samples, SPC_VEC_INJECTED), seed, c_loc(tv), &
                                    res%P_S, res%P_T0, res%P_Tp, res%P_Tm, n_out, phis, phit))
        call finish(sys, res, phis, phit, out)
    end subroutine

    ! qm_dynamics.f90:24-137: the Z nuclear basis states
    subroutine exact_dynamics(sys, sim, res, out)
        type(sys_param), intent(in)            :: sys
        type(sim_param), intent(inout)         :: sim
        type(observables), intent(out)         :: res
        character(*), optional, intent(in)     :: out
        real(dp), allocatable :: phis(:), phit(:)
        integer(c_int64_t) :: n_out
        call set_system(sys)
        n_out = spc_n_steps(sim%t_end, sim%dt) + 1
        call res%malloc(int(n_out))
        allocate(phis(n_out), phit(n_out))
        call check(spc_exact_trace(device_context(), sim_c(sim, 0, SPC_VEC_SUZ), &
                                   res%P_S, res%P_T0, res%P_Tp, res%P_Tm, n_out, phis, phit))
        call finish(sys, res, phis, phit, out)
    end subroutine

end module dynamics
