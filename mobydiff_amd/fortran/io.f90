! io.f90 -- output and checkpoint files with the reference's interface (io.f90:1-275): VTK files of
! the cell-centred velocity every `interval` steps and HDF5 field files every `field_interval`
! steps, plus read_field for a restart.
!   * center_vel runs on the device (mdx_center_velocity) and brings only the three cell-centred
!     arrays to the host; data_output writes the legacy ASCII VTK layout of io.f90:224-259 (the same
!     header lines and number formats, so files compare byte for byte), emitting the cell records
!     from one implied-do output list.  A tanh-stretched grid (which the reference does not have)
!     is written as a RECTILINEAR_GRID with the face coordinates instead of STRUCTURED_POINTS.
!   * write_field downloads un, vn, wn, pn (the `target update from` of io.f90:124-127) and writes
!     them with mdx_h5_write_field (libmobydiff_io, the layout of field_hdf5.c:149-184); read_field
!     reads with mdx_h5_read_field and uploads (io.f90:158-164).
!   * restart_from (no counterpart in the reference, whose driver never calls read_field) reads a
!     field file at start-up and takes its simulated time and time step.
module io
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type
    use :: mdx_c
    implicit none

    type :: output_workspace_type
        real(C_DOUBLE), allocatable :: uc(:,:,:), vc(:,:,:), wc(:,:,:)
    end type output_workspace_type

contains

logical function output_is_due(step, output_interval)
    integer, intent(in) :: step, output_interval
    output_is_due = output_interval > 0
    if (output_is_due) output_is_due = modulo(step, output_interval) == 0
end function output_is_due

subroutine init_output_workspace(out, g, output_interval)
    type(output_workspace_type), intent(inout) :: out
    type(grid_type), intent(in) :: g
    integer, intent(in) :: output_interval
    if (output_interval <= 0) return
    allocate(out%uc(1:g%nx,1:g%ny,1:g%nz), out%vc(1:g%nx,1:g%ny,1:g%nz), out%wc(1:g%nx,1:g%ny,1:g%nz))
    out%uc = 0.0d0
    out%vc = 0.0d0
    out%wc = 0.0d0
end subroutine init_output_workspace

subroutine destroy_output_workspace(out, g)
    type(output_workspace_type), intent(inout) :: out
    type(grid_type), intent(in) :: g
    if (allocated(out%uc)) deallocate(out%uc, out%vc, out%wc)
end subroutine destroy_output_workspace

subroutine maybe_write_vtk(out, f, g, step, output_interval, output_prefix)
    type(output_workspace_type), intent(inout) :: out
    type(field_type), intent(inout) :: f
    type(grid_type), intent(in) :: g
    integer, intent(in) :: step, output_interval
    character(len=*), intent(in) :: output_prefix
    if (output_is_due(step, output_interval)) call write_vtk(out, f, g, step, output_prefix)
end subroutine maybe_write_vtk

subroutine maybe_write_field(f, g, step, field_interval, field_prefix)
    type(field_type), intent(inout) :: f
    type(grid_type), intent(in) :: g
    integer, intent(in) :: step, field_interval
    character(len=*), intent(in) :: field_prefix
    if (output_is_due(step, field_interval)) call write_field(f, g, step, field_prefix)
end subroutine maybe_write_field

subroutine write_field(f, g, step, field_prefix)
    type(field_type), intent(inout) :: f
    type(grid_type), intent(in) :: g
    integer, intent(in) :: step
    character(len=*), intent(in) :: field_prefix
    character(len=256) :: file_name

    write(file_name,'(A,"_",I0,".h5")') trim(field_prefix), step
    print *, "current time step: ", step, "   field filename: ", trim(file_name)
    call mdx_check(mdx_download_fields(mdx_h, f%un, f%vn, f%wn, f%pn, f%us, f%vs, f%ws, &
                                       f%oldrhsu, f%oldrhsv, f%oldrhsw), "write_field")
    if (mdx_h5_write_field(to_c_string(file_name), int(g%nx, C_INT), int(g%ny, C_INT), int(g%nz, C_INT), &
            g%lx, g%ly, g%lz, g%dx, g%dy, g%dz, g%re, g%dt, g%t_current, f%un, f%vn, f%wn, f%pn) /= 0) then
        print *, "error: could not write HDF5 field file: ", trim(file_name), ": ", c_text(mdx_io_last_error())
        error stop
    end if
end subroutine write_field

subroutine read_field(f, g, file_name)
    type(field_type), intent(inout) :: f
    type(grid_type), intent(in) :: g
    character(len=*), intent(in) :: file_name

    if (mdx_h5_read_field(to_c_string(file_name), int(g%nx, C_INT), int(g%ny, C_INT), int(g%nz, C_INT), &
            f%un, f%vn, f%wn, f%pn) /= 0) then
        print *, "error: could not read HDF5 field file: ", trim(file_name), ": ", c_text(mdx_io_last_error())
        error stop
    end if
    call mdx_check(mdx_upload_fields(mdx_h, f%un, f%vn, f%wn, f%pn, f%us, f%vs, f%ws, &
                                     f%oldrhsu, f%oldrhsv, f%oldrhsw), "read_field")
end subroutine read_field

! Start from a field file: its un, vn, wn, pn (read_field), simulated time and time step.  The old
! right-hand sides are not in the file; a step boundary needs none (the first RK stage weights them
! by beta_1 = 0).
subroutine restart_from(f, g, file_name)
    type(field_type), intent(inout) :: f
    type(grid_type), intent(inout) :: g
    character(len=*), intent(in) :: file_name
    integer(C_INT32_T) :: nx, ny, nz
    real(C_DOUBLE) :: dt, t

    if (mdx_h5_read_field_info(to_c_string(file_name), nx, ny, nz, dt, t) /= 0) then
        print *, "error: could not read HDF5 field file: ", trim(file_name), ": ", c_text(mdx_io_last_error())
        error stop
    end if
    if (nx /= g%nx .or. ny /= g%ny .or. nz /= g%nz) then
        print *, "error: field file ", trim(file_name), " holds a ", nx, ny, nz, " grid"
        error stop
    end if
    call read_field(f, g, file_name)
    g%t_current = t
    g%dt = dt
    print *, "restart from ", trim(file_name), " at t = ", t, ", dt = ", dt
end subroutine restart_from

subroutine write_vtk(out, f, g, step, output_prefix)
    type(output_workspace_type), intent(inout) :: out
    type(field_type), intent(inout) :: f
    type(grid_type), intent(in) :: g
    integer, intent(in) :: step
    character(len=*), intent(in) :: output_prefix
    character(len=256) :: file_name

    write(file_name,'(A,"_",I0,".vtk")') trim(output_prefix), step
    print *, "current time step: ", step, "   filename: ", trim(file_name), "   cfl:", g%cfl*g%dt
    call center_vel(out, f, g)
    call data_output(out, g, file_name)
end subroutine write_vtk

subroutine center_vel(out, f, g)
    type(output_workspace_type), intent(inout) :: out
    type(field_type), intent(inout) :: f
    type(grid_type), intent(in) :: g
    call mdx_check(mdx_center_velocity(mdx_h, out%uc, out%vc, out%wc), "center_vel")
end subroutine center_vel

subroutine data_output(out, g, file_name)
    type(output_workspace_type), intent(in) :: out
    type(grid_type), intent(in) :: g
    character(len=*), intent(in) :: file_name
    integer :: u, i, j, k
    real(C_DOUBLE), allocatable :: yc(:), yv(:)

    open(newunit=u, file=trim(file_name), status="replace", action="write")
    write(u,'(A)') "# vtk DataFile Version 3.0"
    write(u,'(A)') "3D velocity field"
    write(u,'(A)') "ASCII"
    if (g%stretch_beta > 0.0d0) then
        ! faces of the cells: uniform in x and z, the stretched v-face positions in y
        allocate(yc(0:g%ny+1), yv(0:g%ny+1))
        call mdx_check(mdx_get_y_coordinates(mdx_h, yc, yv), "data_output")
        write(u,'(A)') "DATASET RECTILINEAR_GRID"
        write(u,'(A,3(1X,I0))') "DIMENSIONS", g%nx+1, g%ny+1, g%nz+1
        write(u,'(A,1X,I0,1X,A)') "X_COORDINATES", g%nx+1, "double"
        write(u,'(3ES20.12)') (g%dx*real(i, C_DOUBLE), i = 0, g%nx)
        write(u,'(A,1X,I0,1X,A)') "Y_COORDINATES", g%ny+1, "double"
        write(u,'(3ES20.12)') (yv(j), j = 1, g%ny+1)
        write(u,'(A,1X,I0,1X,A)') "Z_COORDINATES", g%nz+1, "double"
        write(u,'(3ES20.12)') (g%dz*real(k, C_DOUBLE), k = 0, g%nz)
    else
        write(u,'(A)') "DATASET STRUCTURED_POINTS"
        write(u,'(A,3(1X,I0))') "DIMENSIONS", g%nx+1, g%ny+1, g%nz+1
        write(u,'(A,1X,3ES20.12)') "ORIGIN", 0.0d0, 0.0d0, 0.0d0
        write(u,'(A,1X,3ES20.12)') "SPACING", g%dx, g%dy, g%dz
    end if
    write(u,'(A)') "FIELD FieldData 1"
    write(u,'(A)') "TIME 1 1 double"
    write(u,'(ES20.12)') g%t_current
    write(u,'(A,1X,I0)') "CELL_DATA", g%nx*g%ny*g%nz
    write(u,'(A)') "VECTORS U double"
    ! one record per cell (the format is reused every three values), x fastest
    write(u,'(3ES20.12)') (((out%uc(i,j,k), out%vc(i,j,k), out%wc(i,j,k), i = 1, g%nx), j = 1, g%ny), k = 1, g%nz)
    close(u)
end subroutine data_output

function to_c_string(text) result(c_text_out)
    character(len=*), intent(in) :: text
    character(kind=C_CHAR, len=:), allocatable :: c_text_out
    c_text_out = trim(text) // C_NULL_CHAR
end function to_c_string

end module io
