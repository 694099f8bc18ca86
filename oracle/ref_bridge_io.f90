! ref_bridge_io.f90 -- C-callable entry points into the UNMODIFIED reference output module
! (io.f90) and its HDF5 field writer (field_hdf5.c), compiled from their sources in place by
! oracle/Makefile.  Test infrastructure only: it runs the reference's write_vtk (io.f90:168-182:
! center_vel, then data_output), write_field (io.f90:105-137) and read_field (io.f90:139-166) on
! arbitrary fields, so the test suite can compare this build's VTK and HDF5 files with the
! reference's own.
module ref_bridge_io
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type, init_grid, finalize_grid, init_field
    use :: io, only: output_workspace_type, init_output_workspace, destroy_output_workspace, &
                     write_vtk, write_field, read_field
    implicit none
contains

    subroutine make_grid(g, nx, ny, nz, lx, ly, lz, re, dt, t_current, cfl)
        type(grid_type), intent(inout) :: g
        integer(c_int), intent(in) :: nx, ny, nz
        real(c_double), intent(in) :: lx, ly, lz, re, dt, t_current, cfl
        call init_grid(g)
        g%nx = nx; g%ny = ny; g%nz = nz
        g%lx = lx; g%ly = ly; g%lz = lz
        g%re = re
        g%dt = dt
        call finalize_grid(g)
        g%t_current = t_current
        g%cfl = cfl
    end subroutine make_grid

    function from_c(s) result(t)
        character(kind=c_char), intent(in) :: s(*)
        character(len=:), allocatable :: t
        integer :: n, i
        n = 0
        do while (s(n + 1) /= c_null_char)
            n = n + 1
        end do
        allocate(character(len=n) :: t)
        do i = 1, n
            t(i:i) = s(i)
        end do
    end function from_c

    ! write_vtk(out, f, g, step, prefix): the file <prefix>_<step>.vtk
    subroutine ref_write_vtk(nx, ny, nz, lx, ly, lz, re, dt, t_current, cfl, un, vn, wn, step, prefix) &
            bind(C, name="ref_write_vtk")
        integer(c_int), value :: nx, ny, nz, step
        real(c_double), value :: lx, ly, lz, re, dt, t_current, cfl
        real(c_double), intent(in) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1), wn(0:nx+1,0:ny+1,0:nz+1)
        character(kind=c_char), intent(in) :: prefix(*)
        type(grid_type) :: g
        type(field_type) :: f
        type(output_workspace_type) :: out
        call make_grid(g, nx, ny, nz, lx, ly, lz, re, dt, t_current, cfl)
        call init_field(f, g)
        f%un = un; f%vn = vn; f%wn = wn
        call init_output_workspace(out, g, 1)
        call write_vtk(out, f, g, int(step), from_c(prefix))
        call destroy_output_workspace(out, g)
    end subroutine ref_write_vtk

    ! write_field(f, g, step, prefix): the file <prefix>_<step>.h5
    subroutine ref_write_field(nx, ny, nz, lx, ly, lz, re, dt, t_current, un, vn, wn, pn, step, prefix) &
            bind(C, name="ref_write_field")
        integer(c_int), value :: nx, ny, nz, step
        real(c_double), value :: lx, ly, lz, re, dt, t_current
        real(c_double), intent(in) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1), &
                                      wn(0:nx+1,0:ny+1,0:nz+1), pn(0:nx+1,0:ny+1,0:nz+1)
        character(kind=c_char), intent(in) :: prefix(*)
        type(grid_type) :: g
        type(field_type) :: f
        call make_grid(g, nx, ny, nz, lx, ly, lz, re, dt, t_current, 0.0d0)
        call init_field(f, g)
        f%un = un; f%vn = vn; f%wn = wn; f%pn = pn
        call write_field(f, g, int(step), from_c(prefix))
    end subroutine ref_write_field

    ! read_field(f, g, file_name)
    subroutine ref_read_field(nx, ny, nz, un, vn, wn, pn, path) bind(C, name="ref_read_field")
        integer(c_int), value :: nx, ny, nz
        real(c_double), intent(out) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1), &
                                       wn(0:nx+1,0:ny+1,0:nz+1), pn(0:nx+1,0:ny+1,0:nz+1)
        character(kind=c_char), intent(in) :: path(*)
        type(grid_type) :: g
        type(field_type) :: f
        call make_grid(g, nx, ny, nz, 1.0d0, 1.0d0, 1.0d0, 1.0d0, 1.0d0, 0.0d0, 0.0d0)
        call init_field(f, g)
        call read_field(f, g, from_c(path))
        un = f%un; vn = f%vn; wn = f%wn; pn = f%pn
    end subroutine ref_read_field

end module ref_bridge_io
