! symmetry.f90 -- symmetrised_dynamics with the interface of module symmetry
! (symmetry.f90:16-139 of the reference): the spin-1/2 nuclei of each
! electron approximated by M1 (M2) sets of equivalent nuclei (spc_shrink,
! moments.f90:15-93); every combination of one total-spin multiplet per set
! (spc_symmetry_blocks) trace-sampled on the device with its own master seed
! and weighted by its share of the nuclear space; the weighted sum and its
! Simpson yields written to out.  Electron 2's set counts are copied into
! the block list and multiplets of any multiplicity are propagated, as the
! method intends (the reference assigns an empty section at symmetry.f90:59
! and getHiso_sp stops above multiplicity 3); multiplets carry 1H's nuclear
! Zeeman term.
module symmetry
    use, intrinsic :: iso_fortran_env, only: dp => real64, i8 => int64
    use, intrinsic :: iso_c_binding
    use spinchem_capi
    use spc_random
    use variables
    use class_observables
    use dynamics
    implicit none
    private
    public :: symmetrised_dynamics

contains

    ! moments.f90:15-93 through spc_shrink
    subroutine shrink_sets(a_iso, M, a_bar, n_bar)
        real(dp), intent(in)               :: a_iso(:)
        integer, intent(inout)             :: M
        real(dp), allocatable, intent(out) :: a_bar(:)
        integer, allocatable, intent(out)  :: n_bar(:)
        integer(c_int32_t) :: Mc
        integer(c_int32_t), allocatable :: nb(:)
        real(c_double), allocatable :: ab(:), a(:)
        if (size(a_iso) > 0 .and. M >= size(a_iso)) stop 'Error: Symetry reduction requires M<N'
        Mc = int(M, c_int32_t)
        allocate(ab(max(M, 1)), nb(max(M, 1)), a(max(size(a_iso), 1)))
        a = 0.0_dp
        a(1:size(a_iso)) = a_iso
        if (spc_shrink(int(size(a_iso), c_int32_t), a, Mc, ab, nb) /= SPC_OK) &
            stop 'Error: equivalent-nuclei reduction found no integer set sizes'
        M = int(Mc)
        a_bar = ab(1:M)
        n_bar = int(nb(1:M))
    end subroutine

    ! symmetry.f90:146-178: the block's system, tensors a_bar 1
    subroutine reduce_system(sys, g_I1, a1, g_I2, a2, red)
        type(sys_param), intent(in)  :: sys
        integer, intent(in)          :: g_I1(:), g_I2(:)
        real(dp), intent(in)         :: a1(:), a2(:)
        type(sys_param), intent(out) :: red
        integer :: k
        red%J = sys%J
        red%D = sys%D
        red%kS = sys%kS
        red%kT = sys%kT
        red%field_dir = sys%field_dir
        red%e1%w = sys%e1%w
        red%e2%w = sys%e2%w
        red%e1%g_I = g_I1
        red%e2%g_I = g_I2
        red%e1%a_iso = a1
        red%e2%a_iso = a2
        allocate(red%e1%A(size(g_I1), 3, 3), red%e2%A(size(g_I2), 3, 3))
        red%e1%A = 0.0_dp
        red%e2%A = 0.0_dp
        do k = 1, size(g_I1)
            red%e1%A(k, 1, 1) = a1(k)
            red%e1%A(k, 2, 2) = a1(k)
            red%e1%A(k, 3, 3) = a1(k)
        end do
        do k = 1, size(g_I2)
            red%e2%A(k, 1, 1) = a2(k)
            red%e2%A(k, 2, 2) = a2(k)
            red%e2%A(k, 3, 3) = a2(k)
        end do
        red%e1%isotropic = .true.
        red%e2%isotropic = .true.
        red%Z1 = product(g_I1)
        red%Z2 = product(g_I2)
    end subroutine

    subroutine symmetrised_dynamics(sys, sim, rng, out)
        type(sys_param), intent(in)    :: sys
        type(sim_param), intent(inout) :: sim
        type(rng_t), intent(inout)     :: rng
        character(*), intent(in)       :: out
        real(dp), allocatable :: a1_bar(:), a2_bar(:)
        integer, allocatable :: n1_bar(:), n2_bar(:)
        integer(c_int32_t), allocatable :: n_bar(:)
        integer(c_int32_t), allocatable, target :: K(:,:)
        integer(c_int64_t), allocatable, target :: Zc(:), wc(:)
        integer(c_int64_t) :: nblk
        integer(i8), allocatable :: seeds(:,:)
        integer(i8) :: Z, maxi
        type(observables) :: res, cur
        type(sys_param) :: red
        real(dp) :: w_k
        integer :: i, n_out, start, nsets
        character(len=16) :: tmp
        character(:), allocatable :: folder

        if (any(sys%e1%g_I /= 2) .or. any(sys%e2%g_I /= 2)) &
            stop 'Method has not been implemented for nuclei with I>1/2.'
        Z = product(int(sys%e1%g_I, i8)) * product(int(sys%e2%g_I, i8))
        n_out = int(spc_n_steps(sim%t_end, sim%dt)) + 1
        call res%malloc(n_out)
        call res%set(0.0_dp)

        call shrink_sets(sys%e1%a_iso, sim%M1, a1_bar, n1_bar)
        call shrink_sets(sys%e2%a_iso, sim%M2, a2_bar, n2_bar)
        nsets = sim%M1 + sim%M2
        allocate(n_bar(max(nsets, 1)))
        n_bar(1:sim%M1) = int(n1_bar, c_int32_t)
        n_bar(sim%M1 + 1:nsets) = int(n2_bar, c_int32_t)

        ! cartesian_product + sort_col_prod (symmetry.f90:64-65)
        call check(spc_symmetry_blocks(int(nsets, c_int32_t), n_bar, 0_c_int64_t, nblk, c_null_ptr, c_null_ptr, &
                                       c_null_ptr))
        allocate(K(max(nsets, 1), nblk), Zc(nblk), wc(nblk))
        call check(spc_symmetry_blocks(int(nsets, c_int32_t), n_bar, nblk, nblk, c_loc(K), c_loc(Zc), c_loc(wc)))

        allocate(seeds(2, nblk))
        do i = 1, int(nblk)
            seeds(1, i) = rng%next()
            seeds(2, i) = rng%next()
        end do
        start = 1
        if (sim%isRestart) start = sim%restart
        maxi = 0
        do i = start, int(nblk)
            w_k = real(wc(i) * Zc(i), kind=dp) / real(Z, kind=dp)
            if (w_k >= sim%block_tol) then
                write(tmp, '(I0)') i
                folder = out // '/hamiltonian_' // trim(tmp)
                call execute_command_line('mkdir -p ' // folder)
                call reduce_system(sys, int(K(1:sim%M1, i)), a1_bar, int(K(sim%M1 + 1:nsets, i)), a2_bar, red)
                maxi = max(maxi, int(Zc(i), i8))
                print '(A, I0, A, I0)', 'starting trace sampling of hamiltonian ', i, '/', nblk
                print '(A, I0)', 'Z = ', Zc(i)
                print '(A, I0, A, F0.16)', 'weight_', i, ' = ', w_k
                call rng%set_seed(seeds(:, i))
                call trace_sampling_para(red, sim, rng, cur, folder, block=.true.)
                call cur%scale(w_k)
                call res%update(cur)
            else
                print '(A, I0, A, I0)', 'skipped hamiltonian ', i
                print '(A, I0)', 'Z = ', Zc(i)
                print '(A, I0, A, F0.16)', 'weight_', i, ' = ', w_k
            end if
        end do
        print '(A, I0)', 'Nuclear subspace of largest Hamiltonian: ', maxi

        call res%get_kinetics(sim%dt, sys%kS, sys%kT)
        call res%output(out)
    end subroutine

end module symmetry
