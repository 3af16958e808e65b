! class_observables.f90 -- population traces, Simpson yields and the .data
! files, with the interface of class_observables (observables.f90:9-148).
module class_observables
    use, intrinsic :: iso_fortran_env, only: dp => real64
    implicit none
    private
    public :: observables, write_data, output_style

    ! how write_data renders a value: 0, this compiler's list-directed output
    ! (write(u,*), as observables.f90:143); 1, gfortran's list-directed output
    ! of a real(8), the form the reference's own builds write: a blank, then
    ! G25.17E3 where F editing applies (0.1 <= |x| < 1e17, and 0) and 1PE25.16E3
    ! elsewhere (17 significant digits either way)
    integer :: output_style = 0

    type observables
        real(dp), allocatable :: P_S(:), P_T0(:), P_Tp(:), P_Tm(:)
        real(dp), allocatable :: iden(:), Phi_S(:), Phi_T(:)
    contains
        procedure :: malloc
        procedure :: set
        procedure :: scale
        procedure :: update
        procedure :: get_kinetics
        procedure :: output
    end type

contains

    subroutine malloc(self, n)
        class(observables), intent(inout) :: self
        integer, intent(in) :: n
        if (allocated(self%P_S)) deallocate(self%P_S, self%P_T0, self%P_Tp, self%P_Tm)
        allocate(self%P_S(n), self%P_T0(n), self%P_Tp(n), self%P_Tm(n))
    end subroutine

    subroutine set(self, x)
        class(observables), intent(inout) :: self
        real(dp), intent(in) :: x
        self%P_S = x
        self%P_T0 = x
        self%P_Tp = x
        self%P_Tm = x
    end subroutine

    subroutine scale(self, x)
        class(observables), intent(inout) :: self
        real(dp), intent(in) :: x
        self%P_S = self%P_S * x
        self%P_T0 = self%P_T0 * x
        self%P_Tp = self%P_Tp * x
        self%P_Tm = self%P_Tm * x
    end subroutine

    subroutine update(self, obs)
        class(observables), intent(inout) :: self
        class(observables), intent(in) :: obs
        self%P_S = self%P_S + obs%P_S
        self%P_T0 = self%P_T0 + obs%P_T0
        self%P_Tp = self%P_Tp + obs%P_Tp
        self%P_Tm = self%P_Tm + obs%P_Tm
    end subroutine

    ! cumulative Simpson yields over size/2 pairs (observables.f90:74-101)
    subroutine get_kinetics(self, dt, kS, kT)
        class(observables), intent(inout) :: self
        real(dp), intent(in) :: dt, kS, kT
        real(dp), allocatable :: P_T(:)
        integer :: i, h
        h = size(self%P_S) / 2
        if (allocated(self%iden)) deallocate(self%iden)
        if (allocated(self%Phi_S)) deallocate(self%Phi_S, self%Phi_T)
        allocate(self%iden(size(self%P_S)), self%Phi_S(h), self%Phi_T(h))
        P_T = self%P_Tp + self%P_T0 + self%P_Tm
        self%iden = self%P_S + P_T
        if (h < 1) return
        self%Phi_S(1) = 0.0_dp
        self%Phi_T(1) = 0.0_dp
        do i = 1, h - 1
            self%Phi_S(i+1) = self%Phi_S(i) + (1.0_dp/3.0_dp)*dt*kS*(self%P_S(2*i-1) + 4*self%P_S(2*i) + self%P_S(2*i+1))
            self%Phi_T(i+1) = self%Phi_T(i) + (1.0_dp/3.0_dp)*dt*kT*(P_T(2*i-1) + 4*P_T(2*i) + P_T(2*i+1))
        end do
    end subroutine

    subroutine output(self, path)
        class(observables), intent(inout) :: self
        character(*), intent(in) :: path
        call write_data(self%P_S, path, 'S_prob.data')
        call write_data(self%P_Tp, path, 'Tp_prob.data')
        call write_data(self%P_T0, path, 'T0_prob.data')
        call write_data(self%P_Tm, path, 'Tm_prob.data')
        if (allocated(self%iden)) call write_data(self%iden, path, 'total_prob.data')
        if (allocated(self%Phi_S)) call write_data(self%Phi_S, path, 'S_yield.data')
        if (allocated(self%Phi_T)) call write_data(self%Phi_T, path, 'T_yield.data')
    end subroutine

    ! one list-directed value per line (observables.f90:133-148)
    subroutine write_data(x, path, name)
        real(dp), intent(in) :: x(:)
        character(*), intent(in) :: path, name
        integer :: i, u
        open(newunit=u, file=trim(path) // '/' // trim(name), status='replace')
        do i = 1, size(x)
            if (output_style == 1) then
                if (x(i) == 0.0_dp .or. (abs(x(i)) >= 0.1_dp .and. abs(x(i)) < 1.0e17_dp)) then
                    write(u, '(1X,G25.17E3)') x(i)
                else
                    write(u, '(1X,1PE25.16E3)') x(i)
                end if
            else
                write(u, *) x(i)
            end if
        end do
        close(u)
    end subroutine

end module class_observables
