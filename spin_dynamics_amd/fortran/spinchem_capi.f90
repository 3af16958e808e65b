! spinchem_capi.f90 -- iso_c_binding interface to libspinchem_hip.so
! (include/spinchem.h): the C structs, the entry points, and helpers for
! C strings and the process-wide device context.
module spinchem_capi
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: SPC_OK = 0, SPC_VEC_SUZ = 0, SPC_VEC_PHASE = 1, SPC_VEC_INJECTED = 2

    type, bind(C), public :: spc_spin_system
        integer(c_int32_t) :: n_nuc1 = 0, n_nuc2 = 0
        type(c_ptr)        :: mult = c_null_ptr
        type(c_ptr)        :: hfc = c_null_ptr
        real(c_double)     :: J = 0, D(9) = 0, kS = 0, kT = 0, B = 0, field_dir(3) = [0d0, 0d0, 1d0]
        integer(c_int32_t) :: iso_rule = 1
    end type

    type, bind(C), public :: spc_sc_system
        integer(c_int32_t) :: n_nuc1 = 0, n_nuc2 = 0
        type(c_ptr)        :: mult = c_null_ptr
        type(c_ptr)        :: a_iso = c_null_ptr
        real(c_double)     :: B = 0, field_dir(3) = [0d0, 0d0, 1d0], J = 0, kS = 0, kT = 0
    end type

    type, bind(C), public :: spc_sim_params
        integer(c_int32_t) :: n_samples = 0, n_krylov = 0
        real(c_double)     :: t_end = 0, dt = 0, tol = 0
        integer(c_int32_t) :: init_state = 0, vec_kind = 0
    end type

    public :: spc_ctx_create, spc_ctx_destroy, spc_last_error, spc_system_set, spc_system_set_block, spc_shrink, &
              spc_symmetry_blocks, spc_n_steps, &
              spc_trace_sample, spc_trace_sample_seeds, spc_exact_trace, spc_expm_small, spc_semiclassical_run, &
              spc_ini_open, spc_ini_close, spc_ini_num_sections, spc_ini_has_section, &
              spc_ini_count_values, spc_ini_get_string, spc_ini_get_doubles, spc_ini_get_int64s, &
              c_str, f_str, device_context, check

    interface
        integer(c_int) function spc_ctx_create(device, ctx) bind(C, name='spc_ctx_create')
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine spc_ctx_destroy(ctx) bind(C, name='spc_ctx_destroy')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        type(c_ptr) function spc_last_error(ctx) bind(C, name='spc_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function spc_system_set(ctx, sys) bind(C, name='spc_system_set')
            import :: c_int, c_ptr, spc_spin_system
            type(c_ptr), value :: ctx
            type(spc_spin_system), intent(in) :: sys
        end function
        integer(c_int) function spc_system_set_block(ctx, sys) bind(C, name='spc_system_set_block')
            import :: c_int, c_ptr, spc_spin_system
            type(c_ptr), value :: ctx
            type(spc_spin_system), intent(in) :: sys
        end function
        integer(c_int) function spc_shrink(n, a_iso, M, a_bar, n_bar) bind(C, name='spc_shrink')
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: a_iso(*)
            integer(c_int32_t), intent(inout) :: M
            real(c_double), intent(out) :: a_bar(*)
            integer(c_int32_t), intent(out) :: n_bar(*)
        end function
        integer(c_int) function spc_symmetry_blocks(n_sets, n_bar, capacity, n_blocks, mult, Z_block, w_block) &
                bind(C, name='spc_symmetry_blocks')
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            integer(c_int32_t), value :: n_sets
            integer(c_int32_t), intent(in) :: n_bar(*)
            integer(c_int64_t), value :: capacity
            integer(c_int64_t), intent(out) :: n_blocks
            type(c_ptr), value :: mult, Z_block, w_block
        end function
        integer(c_int64_t) function spc_n_steps(t_end, dt) bind(C, name='spc_n_steps')
            import :: c_int64_t, c_double
            real(c_double), value :: t_end, dt
        end function
        integer(c_int) function spc_trace_sample(ctx, sim, seed, tv, ps, pt0, ptp, ptm, n_out, phis, phit) &
                bind(C, name='spc_trace_sample')
            import :: c_int, c_ptr, c_int64_t, c_double, spc_sim_params
            type(c_ptr), value :: ctx
            type(spc_sim_params), intent(in) :: sim
            integer(c_int64_t), intent(in) :: seed(2)
            type(c_ptr), value :: tv
            real(c_double), intent(out) :: ps(*), pt0(*), ptp(*), ptm(*)
            integer(c_int64_t), value :: n_out
            real(c_double), intent(out) :: phis(*), phit(*)
        end function
        integer(c_int) function spc_trace_sample_seeds(ctx, sim, seeds, ps, pt0, ptp, ptm, n_out, phis, phit) &
                bind(C, name='spc_trace_sample_seeds')
            import :: c_int, c_ptr, c_int64_t, c_double, spc_sim_params
            type(c_ptr), value :: ctx
            type(spc_sim_params), intent(in) :: sim
            integer(c_int64_t), intent(in) :: seeds(*)
            real(c_double), intent(out) :: ps(*), pt0(*), ptp(*), ptm(*)
            integer(c_int64_t), value :: n_out
            real(c_double), intent(out) :: phis(*), phit(*)
        end function
        integer(c_int) function spc_exact_trace(ctx, sim, ps, pt0, ptp, ptm, n_out, phis, phit) &
                bind(C, name='spc_exact_trace')
            import :: c_int, c_ptr, c_int64_t, c_double, spc_sim_params
            type(c_ptr), value :: ctx
            type(spc_sim_params), intent(in) :: sim
            real(c_double), intent(out) :: ps(*), pt0(*), ptp(*), ptm(*)
            integer(c_int64_t), value :: n_out
            real(c_double), intent(out) :: phis(*), phit(*)
        end function
        integer(c_int) function spc_expm_small(ctx, m, A, E) bind(C, name='spc_expm_small')
            import :: c_int, c_ptr, c_int32_t, c_double_complex
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: m
            complex(c_double_complex), intent(in) :: A(*)
            complex(c_double_complex), intent(out) :: E(*)
        end function
        integer(c_int) function spc_ini_open(path, ini) bind(C, name='spc_ini_open')
            import :: c_int, c_ptr, c_char
            character(kind=c_char), intent(in) :: path(*)
            type(c_ptr), intent(out) :: ini
        end function
        subroutine spc_ini_close(ini) bind(C, name='spc_ini_close')
            import :: c_ptr
            type(c_ptr), value :: ini
        end subroutine
        integer(c_int) function spc_ini_num_sections(ini) bind(C, name='spc_ini_num_sections')
            import :: c_int, c_ptr
            type(c_ptr), value :: ini
        end function
        integer(c_int) function spc_ini_has_section(ini, sec) bind(C, name='spc_ini_has_section')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ini
            character(kind=c_char), intent(in) :: sec(*)
        end function
        integer(c_int) function spc_ini_count_values(ini, sec, opt) bind(C, name='spc_ini_count_values')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ini
            character(kind=c_char), intent(in) :: sec(*), opt(*)
        end function
        integer(c_int) function spc_ini_get_string(ini, sec, opt, buf, buflen) bind(C, name='spc_ini_get_string')
            import :: c_int, c_ptr, c_char, c_int32_t
            type(c_ptr), value :: ini
            character(kind=c_char), intent(in) :: sec(*), opt(*)
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int32_t), value :: buflen
        end function
        integer(c_int) function spc_ini_get_doubles(ini, sec, opt, out, n) bind(C, name='spc_ini_get_doubles')
            import :: c_int, c_ptr, c_char, c_double, c_int32_t
            type(c_ptr), value :: ini
            character(kind=c_char), intent(in) :: sec(*), opt(*)
            real(c_double), intent(out) :: out(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function spc_ini_get_int64s(ini, sec, opt, out, n) bind(C, name='spc_ini_get_int64s')
            import :: c_int, c_ptr, c_char, c_int64_t, c_int32_t
            type(c_ptr), value :: ini
            character(kind=c_char), intent(in) :: sec(*), opt(*)
            integer(c_int64_t), intent(out) :: out(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
        integer(c_int) function spc_semiclassical_run(ctx, sys, n_samples, seed, dt, n_steps, ps, pt, phis, phit) &
                bind(C, name='spc_semiclassical_run')
            import :: c_int, c_ptr, c_int64_t, c_double, spc_sc_system
            type(c_ptr), value :: ctx
            type(spc_sc_system), intent(in) :: sys
            integer(c_int64_t), value :: n_samples
            integer(c_int64_t), intent(in) :: seed(2)
            real(c_double), value :: dt
            integer(c_int64_t), value :: n_steps
            real(c_double), intent(out) :: ps(*), pt(*), phis(*), phit(*)
        end function
    end interface

    type(c_ptr), save :: ctx_ = c_null_ptr

contains

    ! Fortran string -> NUL-terminated C string
    function c_str(s) result(cs)
        character(len=*), intent(in) :: s
        character(kind=c_char, len=len_trim(s) + 1) :: cs
        cs = trim(s) // c_null_char
    end function

    ! C string pointer -> Fortran string
    function f_str(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        if (.not. c_associated(p)) then
            s = ''
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = chars(i)
        end do
    end function

    ! The process-wide context on device 0 (created on first use).
    function device_context() result(ctx)
        type(c_ptr) :: ctx
        integer(c_int) :: rc
        if (.not. c_associated(ctx_)) then
            rc = spc_ctx_create(0_c_int, ctx_)
            if (rc /= SPC_OK) then
                print *, 'Error: no HIP device available for the spinchem library (code ', rc, ')'
                stop 1
            end if
        end if
        ctx = ctx_
    end function

    ! Stops with the library's message on a nonzero status.
    subroutine check(rc)
        integer(c_int), intent(in) :: rc
        if (rc /= SPC_OK) then
            print '(A)', f_str(spc_last_error(ctx_))
            stop 1
        end if
    end subroutine

end module spinchem_capi
