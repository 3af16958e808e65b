! variables.f90 -- spin-system and simulation parameters and the .ini reader,
! with the names, units and messages of module variables (spinsys.f90:9-254).
! The .ini file is parsed by the library's reader (spc_ini_*).
module variables
    use, intrinsic :: iso_fortran_env, only: dp => real64, i8 => int64
    use, intrinsic :: iso_c_binding
    use spinchem_capi
    use class_observables, only: output_style
    implicit none
    private
    public :: g_e, gamma_e, gamma_1H, gamma_14N, electron, sys_param, sim_param, read_inp

    real(dp), parameter :: g_e = 2.002319_dp
    real(dp), parameter :: gamma_e = 176.08596_dp
    real(dp), parameter :: gamma_1H = 0.26752218708_dp
    real(dp), parameter :: gamma_14N = 0.0193297792_dp

    type electron
        real(dp)              :: g = g_e
        real(dp)              :: w = 0.0_dp          ! Larmor frequency (-B)
        integer, allocatable  :: g_I(:)              ! nuclear multiplicities
        real(dp), allocatable :: A(:,:,:)            ! A(k,a,b), mT
        real(dp), allocatable :: a_iso(:)
        logical               :: isotropic = .true.
    end type

    type sys_param
        type(electron) :: e1, e2
        real(dp)       :: J = 0.0_dp
        real(dp)       :: D(3,3) = 0.0_dp
        real(dp)       :: kS = 0.0_dp, kT = 0.0_dp   ! rates / gamma_e
        integer        :: Z1 = 1, Z2 = 1
        real(dp)       :: field_dir(3) = [0.0_dp, 0.0_dp, 1.0_dp]
    end type

    type sim_param
        character(:), allocatable :: type
        real(dp), allocatable     :: B(:)
        character(:), allocatable :: init_state
        integer(i8)               :: seed(2) = [42_i8, 99_i8]
        integer                   :: N_samples = 0
        real(dp)                  :: t_end = 0.0_dp, dt = 0.0_dp   ! rad/mT
        integer                   :: N_krylov = 0
        real(dp)                  :: tol = 0.0_dp
        integer                   :: M1 = 0, M2 = 0
        real(dp)                  :: block_tol = 0.0_dp
        character(:), allocatable :: output_folder
        integer                   :: restart = 0
        logical                   :: isRestart = .false.
        character(:), allocatable :: vec_kind              ! 'SUZ' (reference) or 'phase'
    end type

contains

    logical function has(ini, sec, opt)
        type(c_ptr), intent(in) :: ini
        character(*), intent(in) :: sec, opt
        has = spc_ini_count_values(ini, c_str(sec), c_str(opt)) >= 0
    end function

    integer function nvals(ini, sec, opt)
        type(c_ptr), intent(in) :: ini
        character(*), intent(in) :: sec, opt
        nvals = spc_ini_count_values(ini, c_str(sec), c_str(opt))
    end function

    subroutine get_real(ini, sec, opt, val, msg)
        type(c_ptr), intent(in) :: ini
        character(*), intent(in) :: sec, opt
        real(dp), intent(out) :: val(:)
        character(*), intent(in), optional :: msg
        if (spc_ini_get_doubles(ini, c_str(sec), c_str(opt), val, int(size(val), c_int32_t)) /= SPC_OK) then
            if (present(msg)) stop msg
            val = 0.0_dp
        end if
    end subroutine

    function get_string(ini, sec, opt, msg) result(s)
        type(c_ptr), intent(in) :: ini
        character(*), intent(in) :: sec, opt
        character(*), intent(in), optional :: msg
        character(:), allocatable :: s
        character(kind=c_char) :: buf(1000)
        integer :: i, n
        if (spc_ini_get_string(ini, c_str(sec), c_str(opt), buf, 1000_c_int32_t) /= SPC_OK) then
            if (present(msg)) stop msg
            s = ''
            return
        end if
        n = 0
        do i = 1, 1000
            if (buf(i) == c_null_char) exit
            n = i
        end do
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = buf(i)
        end do
    end function

    ! spinsys.f90:60-135
    subroutine read_inp(filename, sys, sim)
        character(*), intent(in)       :: filename
        type(sys_param), intent(out)   :: sys
        type(sim_param), intent(out)   :: sim
        type(c_ptr) :: ini
        real(dp)    :: tmp(9), t1(1)
        integer(c_int64_t) :: seed(2)
        logical     :: dirExists

        if (spc_ini_open(c_str(filename), ini) /= SPC_OK) stop 'Error: Input file does not exist.'
        call check_inp(ini)

        call get_real(ini, 'system variables', 'J', t1, 'Error: Exchange coupling is missing')
        sys%J = t1(1)
        call get_real(ini, 'system variables', 'D', tmp, 'Error: Dipolar coupling is missing')
        sys%D = reshape(tmp, [3,3])
        call get_real(ini, 'system variables', 'kS', t1, 'Error: Singlet recombination rate is missing')
        sys%kS = t1(1) / gamma_e
        call get_real(ini, 'system variables', 'kT', t1, 'Error: Triplet recombination rate is missing')
        sys%kT = t1(1) / gamma_e

        sim%type = get_string(ini, 'simulation parameters', 'simulation_type', 'Error: Simulation type is missing')
        if (nvals(ini, 'simulation parameters', 'B') < 0) stop 'Error: List of experiment magnetic fields is missing'
        allocate(sim%B(nvals(ini, 'simulation parameters', 'B')))
        call get_real(ini, 'simulation parameters', 'B', sim%B, 'Error: List of experiment magnetic fields is missing')
        sim%init_state = get_string(ini, 'simulation parameters', 'initial_state', 'Error: Initial state is missing')
        call get_real(ini, 'simulation parameters', 'dt', t1, 'Error: Propagation timestep is missing')
        sim%dt = t1(1)
        if (has(ini, 'simulation parameters', 'seed')) then
            if (spc_ini_get_int64s(ini, c_str('simulation parameters'), c_str('seed'), seed, 2_c_int32_t) == SPC_OK) &
                sim%seed = seed
        end if
        if (has(ini, 'simulation parameters', 'N_samples')) then
            call get_real(ini, 'simulation parameters', 'N_samples', t1)
            sim%N_samples = nint(t1(1))
        end if
        call get_real(ini, 'simulation parameters', 'simulation_time', t1)
        sim%t_end = (t1(1) * gamma_e) / 1000.0_dp
        sim%dt = (sim%dt * gamma_e) / 1000.0_dp
        if (has(ini, 'simulation parameters', 'block_tolerance')) then
            call get_real(ini, 'simulation parameters', 'block_tolerance', t1)
            sim%block_tol = t1(1)
        end if
        if (has(ini, 'simulation parameters', 'N_krylov')) then
            call get_real(ini, 'simulation parameters', 'N_krylov', t1)
            sim%N_krylov = nint(t1(1))
        end if
        if (has(ini, 'simulation parameters', 'integrator_tolerance')) then
            call get_real(ini, 'simulation parameters', 'integrator_tolerance', t1)
            sim%tol = t1(1)
        end if
        if (has(ini, 'simulation parameters', 'M1')) then
            call get_real(ini, 'simulation parameters', 'M1', t1)
            sim%M1 = nint(t1(1))
        end if
        if (has(ini, 'simulation parameters', 'M2')) then
            call get_real(ini, 'simulation parameters', 'M2', t1)
            sim%M2 = nint(t1(1))
        end if
        sim%isRestart = has(ini, 'simulation parameters', 'restart')
        if (sim%isRestart) then
            call get_real(ini, 'simulation parameters', 'restart', t1)
            sim%restart = nint(t1(1))
        end if
        sim%vec_kind = 'SUZ'
        if (has(ini, 'simulation parameters', 'trace_vectors')) &
            sim%vec_kind = get_string(ini, 'simulation parameters', 'trace_vectors')
        if (has(ini, 'simulation parameters', 'field_direction')) &
            call get_real(ini, 'simulation parameters', 'field_direction', sys%field_dir)

        if (has(ini, 'simulation parameters', 'output_format')) then
            ! extension key: 'gfortran' writes the .data files as the reference's gfortran builds do
            if (get_string(ini, 'simulation parameters', 'output_format') == 'gfortran') output_style = 1
        end if
        sim%output_folder = get_string(ini, 'simulation parameters', 'output_folder')
        inquire(file=sim%output_folder // '/.', exist=dirExists)
        if (.not. dirExists) stop 'Error: Output folder does not exist.'

        call electron_init(ini, 1, sys%e1)
        call electron_init(ini, 2, sys%e2)
        sys%Z1 = product(sys%e1%g_I)
        sys%Z2 = product(sys%e2%g_I)
        call spc_ini_close(ini)
    end subroutine

    ! spinsys.f90:137-163
    subroutine check_inp(ini)
        type(c_ptr), intent(in) :: ini
        character(len=24) :: sections(4)
        integer :: i
        sections = [character(len=24) :: 'system variables', 'simulation parameters', 'electron 1', 'electron 2']
        if (spc_ini_num_sections(ini) < size(sections)) then
            print *, 'Error: Number of sections in input file is less than expected.'
            stop
        else if (spc_ini_num_sections(ini) > size(sections)) then
            print *, 'Error: Number of sections in input file is more than expected.'
            stop
        end if
        do i = 1, size(sections)
            if (spc_ini_has_section(ini, c_str(trim(sections(i)))) == 0) then
                print *, 'Error: Section "', trim(sections(i)), '" is missing from input file.'
                stop
            end if
        end do
    end subroutine

    ! spinsys.f90:165-235 (N_I copies of each multiplicity in I; A1..An)
    subroutine electron_init(ini, n, e)
        type(c_ptr), intent(in)     :: ini
        integer, intent(in)         :: n
        type(electron), intent(out) :: e
        character(len=20) :: sec
        character(len=10) :: opt
        real(dp), allocatable :: mult(:), cnt(:)
        real(dp) :: A_tmp(9), t1(1)
        integer :: i, k, pos, nI
        write(sec, '(A,I0)') 'electron ', n
        if (has(ini, trim(sec), 'g')) then
            call get_real(ini, trim(sec), 'g', t1)
            e%g = t1(1)
        end if
        nI = max(nvals(ini, trim(sec), 'I'), 0)
        if (nvals(ini, trim(sec), 'N_I') /= nvals(ini, trim(sec), 'I')) then
            if (nI > 0 .or. nvals(ini, trim(sec), 'N_I') > 0) stop 'Sections N_I and I must have the sam length.'
        end if
        allocate(mult(nI), cnt(nI))
        if (nI > 0) then
            call get_real(ini, trim(sec), 'I', mult)
            call get_real(ini, trim(sec), 'N_I', cnt)
        end if
        allocate(e%g_I(nint(sum(cnt))))
        pos = 0
        do i = 1, nI
            do k = 1, nint(cnt(i))
                pos = pos + 1
                e%g_I(pos) = nint(mult(i))
            end do
        end do
        allocate(e%A(size(e%g_I), 3, 3), e%a_iso(size(e%g_I)))
        do i = 1, size(e%g_I)
            write(opt, '(A,I0)') 'A', i
            if (spc_ini_get_doubles(ini, c_str(sec), c_str(opt), A_tmp, 9_c_int32_t) /= SPC_OK) then
                print '(A,I0,A,I0,A)', 'Error: Hyperfine copuling A', i, ' for electron ', n, ' is missing.'
                stop
            end if
            e%A(i,:,:) = reshape(A_tmp, [3,3])
            e%a_iso(i) = (A_tmp(1) + A_tmp(5) + A_tmp(9)) / 3.0_dp
        end do
        e%isotropic = .true.
        do i = 1, size(e%g_I)
            if (e%A(i,1,1) /= e%A(i,2,2)) e%isotropic = .false.
        end do
    end subroutine

end module variables
