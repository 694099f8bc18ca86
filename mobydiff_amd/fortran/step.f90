! step.f90 -- RK3 coefficients, momentum predictor and CFL with the reference interface
! (step.f90:21-32,210); the stencil runs on the GPU through the C ABI.
module step
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type
    use :: ibmm, only: ibm_type
    use :: boundary, only: boundary_type
    use :: mdx_c
    implicit none

    ! low-storage RK3: alpha + beta = gamma per stage, sum of gamma = 1
    real(C_DOUBLE), parameter :: rk_alpha(3) = [64.0d0/120.0d0, 50.0d0/120.0d0, 90.0d0/120.0d0]
    real(C_DOUBLE), parameter :: rk_beta(3) = [0.0d0, -34.0d0/120.0d0, -50.0d0/120.0d0]
    real(C_DOUBLE), parameter :: rk_gamma(3) = [64.0d0/120.0d0, 16.0d0/120.0d0, 40.0d0/120.0d0]

contains

    subroutine momentum(f, g, dt_alpha, dt_beta, dt_gamma, ibm, bc)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        real(C_DOUBLE), intent(in) :: dt_alpha, dt_beta, dt_gamma
        type(ibm_type), intent(in) :: ibm
        type(boundary_type), intent(in) :: bc
        call mdx_check(mdx_momentum(mdx_h, dt_alpha, dt_beta, dt_gamma), "momentum")
        if (g%flow_rate) call mdx_check(mdx_flowrate_shift(mdx_h, dt_gamma, c_null_ptr), "momentum (flow rate)")
    end subroutine momentum

    real(C_DOUBLE) function get_cfl(f, g)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        real(C_DOUBLE) :: cfl
        call mdx_check(mdx_get_cfl(mdx_h, cfl), "get_cfl")
        get_cfl = cfl
    end function get_cfl

end module step
