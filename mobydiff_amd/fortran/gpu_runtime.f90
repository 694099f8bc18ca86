! gpu_runtime.f90 -- device lifecycle with the reference interface (gpu_runtime.f90:11-48): the
! persistent device copies of field_type are the library handle's buffers, created and filled by
! enter_field_data and released by exit_field_data.
module gpu_runtime
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type
    use :: mdx_c
    implicit none

contains

    subroutine init_openmp_offload()
        print *, "MI355X time-step library: libmobydiff_hip (HIP, gfx950)"
    end subroutine init_openmp_offload

    subroutine enter_field_data(f, g)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        mdx_cfg%nx = g%nx; mdx_cfg%ny = g%ny; mdx_cfg%nz = g%nz
        mdx_cfg%lx = g%lx; mdx_cfg%ly = g%ly; mdx_cfg%lz = g%lz
        mdx_cfg%re = g%re
        mdx_cfg%forcing = [g%forcing_x, g%forcing_y, g%forcing_z]
        mdx_cfg%stretch_beta = g%stretch_beta
        mdx_cfg%flowrate = merge(1, 0, g%flow_rate)
        mdx_cfg%ub_target = g%ub_target
        mdx_cfg%flow_height = merge(g%flow_height, g%ly, g%flow_height > 0.0d0)
        call mdx_check(mdx_create(mdx_cfg, mdx_h), "enter_field_data")
        call mdx_check(mdx_upload_fields(mdx_h, f%un, f%vn, f%wn, f%pn, f%us, f%vs, f%ws, &
                                         f%oldrhsu, f%oldrhsv, f%oldrhsw), "enter_field_data")
    end subroutine enter_field_data

    ! bring the device fields back to the host arrays (`target update from`, io.f90:124-127)
    subroutine update_host_fields(f)
        type(field_type), intent(inout) :: f
        call mdx_check(mdx_download_fields(mdx_h, f%un, f%vn, f%wn, f%pn, f%us, f%vs, f%ws, &
                                           f%oldrhsu, f%oldrhsv, f%oldrhsw), "update_host_fields")
    end subroutine update_host_fields

    subroutine exit_field_data(f, g)
        type(field_type), intent(inout) :: f
        type(grid_type), intent(in) :: g
        if (c_associated(mdx_h)) then
            call update_host_fields(f)
            call mdx_check(mdx_destroy(mdx_h), "exit_field_data")
            mdx_h = c_null_ptr
        end if
    end subroutine exit_field_data

end module gpu_runtime
