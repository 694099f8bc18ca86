! ibm.f90 -- immersed-boundary state with the reference interface (module ibmm, ibm.f90:27-81,133):
! the coefficient arrays stay on the host as in the reference, are computed by the library's host
! code (same algorithm: body test, bisection to the interface, second-order distance weights) and
! uploaded by enter_ibm_data.
module ibmm
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type
    use :: mdx_c
    implicit none

    real(C_DOUBLE), parameter :: SOLID = 1.0d30

    type :: ibm_type
        integer :: n_wave_x, n_wave_z
        real(C_DOUBLE) :: amp_x, phase_x
        real(C_DOUBLE) :: amp_z, phase_z
        real(C_DOUBLE), allocatable :: coef_u(:,:,:), coef_v(:,:,:), coef_w(:,:,:)
    end type ibm_type

contains

    subroutine init_ibm(ibm, g)
        type(ibm_type), intent(inout) :: ibm
        type(grid_type), intent(in) :: g
        ibm%n_wave_x = 1
        ibm%n_wave_z = 1
        ibm%amp_x = 5*g%dy
        ibm%amp_z = 5*g%dy
        ibm%phase_x = 0.0d0
        ibm%phase_z = 0.0d0
        if (mdx_geo%kind == 0) then
            if (mdx_geo%amp_x > 0.0d0) ibm%amp_x = mdx_geo%amp_x
            if (mdx_geo%n_wave_x > 0) ibm%n_wave_x = mdx_geo%n_wave_x
            ibm%phase_x = mdx_geo%phase_x
        end if
        allocate(ibm%coef_u(0:g%nx+1,0:g%ny+1,0:g%nz+1))
        allocate(ibm%coef_v(0:g%nx+1,0:g%ny+1,0:g%nz+1))
        allocate(ibm%coef_w(0:g%nx+1,0:g%ny+1,0:g%nz+1))
        ibm%coef_u = 0.0d0; ibm%coef_v = 0.0d0; ibm%coef_w = 0.0d0
    end subroutine init_ibm

    subroutine set_ibm_coeff(g, ibm, coeff, dix, diy, diz)
        type(grid_type), intent(in) :: g
        type(ibm_type), intent(in) :: ibm
        real(C_DOUBLE), intent(inout), contiguous :: coeff(:,:,:)
        integer, intent(in) :: dix, diy, diz
        type(mdx_geometry) :: geo
        if (mdx_geo_enabled == 0) then
            coeff = 0.0d0
            return
        end if
        geo = mdx_geo
        if (geo%kind == 0) then
            geo%amp_x = ibm%amp_x
            geo%phase_x = ibm%phase_x
            geo%n_wave_x = ibm%n_wave_x
        end if
        call mdx_check(mdx_ibm_coefficient(mdx_h, geo, int(dix, c_int32_t), int(diy, c_int32_t), &
                                           int(diz, c_int32_t), coeff), "set_ibm_coeff")
    end subroutine set_ibm_coeff

    subroutine enter_ibm_data(ibm, g)
        type(ibm_type), intent(inout), target :: ibm
        type(grid_type), intent(in) :: g
        if (mdx_geo_enabled == 0) then
            call mdx_check(mdx_set_ibm(mdx_h, c_null_ptr, c_null_ptr, c_null_ptr), "enter_ibm_data")
        else
            call mdx_check(mdx_set_ibm(mdx_h, c_loc(ibm%coef_u), c_loc(ibm%coef_v), c_loc(ibm%coef_w)), &
                           "enter_ibm_data")
        end if
    end subroutine enter_ibm_data

    subroutine exit_ibm_data(ibm, g)
        type(ibm_type), intent(inout) :: ibm
        type(grid_type), intent(in) :: g
    end subroutine exit_ibm_data

end module ibmm
