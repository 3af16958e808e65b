! matrix_exponential.f90 -- expm(A, E) with the interface of module
! matrix_exponential (matrix_exponential.f90:14-17), evaluated on the device
! (Pade(6,6), scaling and squaring) through spc_expm_small.
module matrix_exponential
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use, intrinsic :: iso_c_binding
    use spinchem_capi
    implicit none
    private
    public :: expm

contains

    subroutine expm(A, E)
        complex(dp), intent(in)               :: A(:,:)
        complex(dp), allocatable, intent(out) :: E(:,:)
        complex(c_double_complex), allocatable :: Ac(:,:)
        integer :: m
        m = size(A, 1)
        allocate(E(m, m))
        Ac = A
        call check(spc_expm_small(device_context(), int(m, c_int32_t), Ac, E))
    end subroutine

end module matrix_exponential
