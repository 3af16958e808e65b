! main.f90 -- `spinchem input.ini`: the driver of app/main.f90 (lines 1-88):
! read the input, loop over the field list with the generator re-seeded per
! field, make <output_folder>/<F0.3(B)>, dispatch on simulation_type.
program main
    use, intrinsic :: iso_fortran_env, only: dp => real64, i8 => int64
    use variables
    use spc_random
    use class_observables
    use dynamics
    use sc_dynamics
    use symmetry
    implicit none

    character(len=600)        :: arg
    character(:), allocatable :: input_file, folder
    character(len=16)         :: guz
    type(sys_param)           :: sys
    type(sim_param)           :: sim
    type(rng_t)               :: rng
    type(observables)         :: res
    logical                   :: inpExists
    integer                   :: i
    integer(i8)               :: c0, c1, rate
    integer                   :: secs

    call system_clock(c0, rate)
    print *, 'Program initialised'
    call get_command_argument(1, arg)
    input_file = trim(arg)
    inquire(file=input_file, exist=inpExists)
    if (.not. inpExists) stop 'Error: Input file does not exist.'

    call read_inp(input_file, sys, sim)

    do i = 1, size(sim%B)
        call rng%set_seed(sim%seed)
        sys%e1%w = -sim%B(i)
        sys%e2%w = -sim%B(i)
        write(guz, '(F0.3)') sim%B(i)
        folder = sim%output_folder // '/' // trim(guz)
        call execute_command_line('mkdir -p ' // folder)

        select case (sim%type)
        case ('trace_sampling')
            call trace_sampling_para(sys, sim, rng, res, folder)
        case ('trace_sampling_serial')
            call trace_sampling(sys, sim, rng, res, folder)
        case ('exact_dynamics')
            call exact_dynamics(sys, sim, res, folder)
        case ('semi_classical')
            call semi_classical(sim, sys, res, folder)
        case ('symmetry_dynamics')
            print *, sim%B(i)
            call symmetrised_dynamics(sys, sim, rng, folder)
        case default
            stop 'Error: Invalid simulation type entered.'
        end select
    end do

    call system_clock(c1)
    secs = int((c1 - c0) / rate)
    if (sim%type /= 'symmetry_dynamics') then
        print '(A15,I2,A1,I2,A1,I2)', 'Time elapsed:  ', secs / 3600, ':', mod(secs / 60, 60), ':', mod(secs, 60)
    end if
end program main
