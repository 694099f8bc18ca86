! pressure.f90 -- pressure_workspace and pressure_solver with the reference interface
! (pressure_workspace.f90:5-24, pressure_solver.f90:31-74).  The solver is chosen at run time
! (input.ini [pressure] solver) instead of by the USE_REDBLACK / USE_CUFFT build options.
module pressure_workspace
    use, intrinsic :: iso_c_binding
    use :: mdx_c, only: mdx_solve_stats
    implicit none

    type :: pressure_solver_type
        integer(C_INT) :: nIter = 3              ! red-black iterations (fixed-count SOLA mode)
        real(C_DOUBLE) :: sor = 1.5              ! relaxation factor
        ! extensions
        integer(C_INT) :: solver = 1             ! 0 fft, 1 sola (fixed nIter), 2 sor (to tolerance)
        real(C_DOUBLE) :: tol = 1.0d-8           ! ||div||_inf tolerance of the sor solver
        integer(C_INT) :: max_iter = 200000
        integer(C_INT) :: check_every = 20
        integer(C_INT) :: precision = 0          ! 0 fp64, 1 fp32 sweeps in fp64 defect correction
        integer(C_INT) :: inner_iters = 20
        type(mdx_solve_stats) :: last            ! statistics of the last solve
    end type pressure_solver_type
end module pressure_workspace

module pressure_solver
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type
    use :: pressure_workspace, only: pressure_solver_type
    use :: ibmm, only: ibm_type
    use :: boundary, only: boundary_type
    use :: mdx_c
    implicit none

    private
    public :: pressure_solver_type, init_pressure_solver, pressure_projection, destroy_pressure_solver

contains

    subroutine init_pressure_solver(ps, g)
        type(pressure_solver_type), intent(inout) :: ps
        type(grid_type), intent(in) :: g
        ! workspaces and FFT tables are created with the device state (enter_field_data)
        if (ps%solver /= 0 .and. (mod(g%nx,2) /= 0 .or. mod(g%nz,2) /= 0)) then
            error stop "red-black pressure solver requires even nx and nz for periodic coloring"
        end if
    end subroutine init_pressure_solver

    subroutine pressure_projection(ps, f, g, dt_gamma, ibm, bc)
        type(pressure_solver_type), intent(inout) :: ps
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        real(C_DOUBLE), intent(in) :: dt_gamma
        type(ibm_type), intent(in) :: ibm
        type(boundary_type), intent(in) :: bc
        call mdx_check(mdx_pressure_projection(mdx_h, dt_gamma, ps%last), "pressure_projection")
        if (ps%solver == 2 .and. ps%last%converged == 0) then
            print *, "pressure_projection: SOR stopped at max_iter with ||div||_inf =", ps%last%div_inf
            error stop "pressure solve did not reach its tolerance"
        end if
    end subroutine pressure_projection

    subroutine destroy_pressure_solver(ps)
        type(pressure_solver_type), intent(inout) :: ps
    end subroutine destroy_pressure_solver

end module pressure_solver
