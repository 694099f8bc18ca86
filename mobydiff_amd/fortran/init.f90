! init.f90 -- grid and field types (the interface of init.f90:5-71 in the reference), host side.
! The components keep the reference's names and meaning so driver code using g%nx, f%un, ...
! compiles unchanged; the device copies live in libmobydiff_hip (see gpu_runtime).
module init
    use, intrinsic :: iso_c_binding
    implicit none

    type :: grid_type
        integer :: nx, ny, nz
        integer :: nsteps
        real(C_DOUBLE) :: lx, ly, lz
        real(C_DOUBLE) :: dx, dy, dz
        real(C_DOUBLE) :: re, dt, t_final, t_current, cfl, cflmax, dtmax
        real(C_DOUBLE) :: forcing_x, forcing_y, forcing_z
        ! extensions: tanh wall-normal stretching (0 = uniform) and constant flow rate
        real(C_DOUBLE) :: stretch_beta = 0.0d0
        logical :: flow_rate = .false.
        real(C_DOUBLE) :: ub_target = 1.0d0, flow_height = 0.0d0
    end type grid_type

    type :: field_type
        real(C_DOUBLE), allocatable :: un(:,:,:), us(:,:,:), oldrhsu(:,:,:)
        real(C_DOUBLE), allocatable :: vn(:,:,:), vs(:,:,:), oldrhsv(:,:,:)
        real(C_DOUBLE), allocatable :: wn(:,:,:), ws(:,:,:), oldrhsw(:,:,:)
        real(C_DOUBLE), allocatable :: pn(:,:,:)
    end type field_type

contains

    ! defaults of the reference driver (init.f90:23-46)
    subroutine init_grid(g)
        type(grid_type), intent(inout) :: g
        g%nx = 200; g%ny = 200; g%nz = 200
        g%nsteps = 10000
        g%lx = 1.0d0; g%ly = 1.0d0; g%lz = 1.0d0
        g%re = 100.0d0
        g%forcing_x = 1.0d0; g%forcing_y = 0.0d0; g%forcing_z = 0.0d0
        g%dt = 1.0d-4
        g%cflmax = 0.0d0
        g%dtmax = 1.0d-3
        g%t_current = 0.0d0
        g%cfl = 0.0d0
        call finalize_grid(g)
    end subroutine init_grid

    subroutine finalize_grid(g)
        type(grid_type), intent(inout) :: g
        g%dx = g%lx / real(g%nx, C_DOUBLE)
        g%dy = g%ly / real(g%ny, C_DOUBLE)
        g%dz = g%lz / real(g%nz, C_DOUBLE)
        g%t_final = g%dt * real(g%nsteps, C_DOUBLE)
    end subroutine finalize_grid

    subroutine init_field(f, g)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        allocate(f%un(0:g%nx+1,0:g%ny+1,0:g%nz+1), f%us(0:g%nx+1,0:g%ny+1,0:g%nz+1), f%oldrhsu(1:g%nx,1:g%ny,1:g%nz))
        allocate(f%vn(0:g%nx+1,0:g%ny+1,0:g%nz+1), f%vs(0:g%nx+1,0:g%ny+1,0:g%nz+1), f%oldrhsv(1:g%nx,1:g%ny,1:g%nz))
        allocate(f%wn(0:g%nx+1,0:g%ny+1,0:g%nz+1), f%ws(0:g%nx+1,0:g%ny+1,0:g%nz+1), f%oldrhsw(1:g%nx,1:g%ny,1:g%nz))
        allocate(f%pn(0:g%nx+1,0:g%ny+1,0:g%nz+1))
        f%un = 0.0d0; f%us = 0.0d0; f%oldrhsu = 0.0d0
        f%vn = 0.0d0; f%vs = 0.0d0; f%oldrhsv = 0.0d0
        f%wn = 0.0d0; f%ws = 0.0d0; f%oldrhsw = 0.0d0
        f%pn = 0.0d0
    end subroutine init_field

end module init
