! boundary.f90 -- BC table and ghost-cell updates with the reference interface (boundary.f90:6-25,
! 173-176).  The ghost fills run on the GPU through the C ABI; the table is passed to the library at
! enter_field_data.
module boundary
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type
    use :: mdx_c
    implicit none

    type :: boundary_type
        logical(C_BOOL) :: isPeriodic(1:3)
        integer(C_INT) :: bcType(1:3,0:1,0:3)     ! (direction, side, variable p/u/v/w); 0 Dirichlet, 1 Neumann
        real(C_DOUBLE) :: bcValue(1:3,0:1,0:3)
    end type boundary_type

contains

    subroutine init_bc(bc)
        type(boundary_type), intent(inout) :: bc
        bc%isPeriodic(1:3) = .true.
        bc%bcType(:,:,:) = 0
        bc%bcValue(:,:,:) = 0.0d0
    end subroutine init_bc

    ! FFT-solver ghost fill: y no-slip walls, x/z periodic (boundary.f90:23-170)
    subroutine apply_bc(f, g)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        call mdx_check(mdx_apply_bc(mdx_h), "apply_bc")
    end subroutine apply_bc

    ! table-driven ghost fill of the red-black solver (boundary.f90:173-368)
    subroutine apply_bc_redblack(f, g, bc)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        type(boundary_type), intent(in) :: bc
        call mdx_check(mdx_apply_bc_redblack(mdx_h), "apply_bc_redblack")
    end subroutine apply_bc_redblack

end module boundary
