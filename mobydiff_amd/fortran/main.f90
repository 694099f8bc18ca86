! main.f90 -- the reference driver (main.f90:10-120) with its setup order, RK3 stage loop, CFL/dt
! update, VTK/HDF5 output cadence and timing line, calling the GPU path through the drop-in modules.
! Differences: the pressure solver is chosen at run time ([pressure] solver) instead of by
! USE_REDBLACK, the step loop synchronises the device before the clock stops (the library is
! stream-asynchronous), an optional Poiseuille + perturbation initial field ([init] profile) or a
! field file ([init] restart, read by io's read_field) replaces the zero start, and an optional
! raw dump of the final fields ([output] dump_final) follows the loop.
program main
    use, intrinsic :: iso_fortran_env, only: int64
    use, intrinsic :: iso_c_binding
    use :: init
    use :: config
    use :: boundary
    use :: io
    use :: step
    use :: pressure_workspace
    use :: pressure_solver
    use :: gpu_runtime
    use :: ibmm
    use :: mdx_c
    implicit none

    integer :: i, arg_status, rkStage
    integer :: output_interval, field_interval
    integer(int64) :: sweeps
    logical :: need_cfl
    real(C_DOUBLE) :: dt_alpha, dt_beta, dt_gamma, div_inf
    real(C_DOUBLE) :: loop_seconds, seconds_per_step
    integer(int64) :: loop_clock_start, loop_clock_end, clock_rate
    character(len=256) :: input_file, output_prefix, field_prefix, field_prefix_check
    type(grid_type) :: g
    type(field_type) :: f
    type(boundary_type) :: bc
    type(pressure_solver_type) :: ps
    type(ibm_type) :: ibm
    type(output_workspace_type) :: out
    type(mdx_geometry), target :: geo

    call get_command_argument(1, input_file, status=arg_status)
    if (arg_status /= 0 .or. len_trim(input_file) == 0) input_file = "input.ini"

    print *, "initialising grid..."
    call init_grid(g)

    print *, "initialising BCs..."
    call init_bc(bc)
    call read_runtime_config(g, ps, bc, output_interval, output_prefix, field_interval, field_prefix, input_file)
    call get_command_argument(2, field_prefix_check, status=arg_status)
    if (arg_status == 0 .and. trim(field_prefix_check) == "--print-config") then
        call print_config()
        stop
    end if
    call init_openmp_offload()

    print *, "initialising fields..."
    call init_field(f, g)
    call enter_field_data(f, g)
    call init_output_workspace(out, g, output_interval)

    print *, "initialising pressure solver..."
    call init_pressure_solver(ps, g)

    print *, "initialising IBM..."
    call init_ibm(ibm, g)
    call set_ibm_coeff(g, ibm, ibm%coef_u, 1, 0, 0)
    call set_ibm_coeff(g, ibm, ibm%coef_v, 0, 1, 0)
    call set_ibm_coeff(g, ibm, ibm%coef_w, 0, 0, 1)
    call enter_ibm_data(ibm, g)

    if (mdx_init_poiseuille) then
        print *, "initial field: Poiseuille + seeded perturbation, projected"
        if (mdx_geo_enabled /= 0 .and. mdx_geo%kind == 1) then
            geo = mdx_geo
            call mdx_check(mdx_init_fields(mdx_h, int(mdx_init_seed, c_int32_t), mdx_init_amp, c_loc(geo), g%dt), &
                           "init_fields")
        else
            call mdx_check(mdx_init_fields(mdx_h, int(mdx_init_seed, c_int32_t), mdx_init_amp, c_null_ptr, g%dt), &
                           "init_fields")
        end if
    end if
    if (len_trim(mdx_restart_file) > 0) call restart_from(f, g, trim(mdx_restart_file))

    sweeps = 0
    print *, "main loop starting..."
    call system_clock(count_rate=clock_rate)
    call system_clock(count=loop_clock_start)
    do i = 1, g%nsteps
        g%t_current = g%t_current + g%dt

        do rkStage = 1, 3
            dt_alpha = g%dt*rk_alpha(rkStage)
            dt_beta  = g%dt*rk_beta(rkStage)
            dt_gamma = g%dt*rk_gamma(rkStage)

            call momentum(f, g, dt_alpha, dt_beta, dt_gamma, ibm, bc)
            if (ps%solver == MDX_SOLVER_FFT) call apply_bc(f, g)

            call pressure_projection(ps, f, g, dt_gamma, ibm, bc)
            sweeps = sweeps + ps%last%iterations
            if (ps%solver == MDX_SOLVER_FFT) then
                call apply_bc(f, g)
            else
                call apply_bc_redblack(f, g, bc)
            end if
        end do

        need_cfl = (g%cflmax > 0.0d0) .or. output_is_due(i, output_interval)
        if (need_cfl) g%cfl = get_cfl(f, g)

        if (g%cflmax > 0.0d0 .and. g%cfl > 0.0d0) then
            g%dt = min(g%cflmax/g%cfl, g%dtmax)
        end if

        if (output_interval > 0) call maybe_write_vtk(out, f, g, i, output_interval, output_prefix)
        if (field_interval > 0) call maybe_write_field(f, g, i, field_interval, field_prefix)
    end do
    call mdx_check(mdx_synchronize(mdx_h), "synchronize")
    call system_clock(count=loop_clock_end)
    loop_seconds = real(loop_clock_end - loop_clock_start, C_DOUBLE) / real(clock_rate, C_DOUBLE)
    seconds_per_step = loop_seconds / real(g%nsteps, C_DOUBLE)

    print *, "main loop ended..."
    write(*,'(A,1X,I0,1X,A,1X,ES16.8,1X,A,1X,ES16.8)') &
        "timing: nsteps", g%nsteps, "loop_seconds", loop_seconds, "seconds_per_step", seconds_per_step
    call mdx_check(mdx_div_norm(mdx_h, div_inf), "div_norm")
    write(*,'(A,1X,ES12.4,1X,A,1X,I0,1X,A,1X,F12.3)') "final: div_inf", div_inf, "sor_sweeps", sweeps, &
        "mcell_steps_per_s", real(g%nx, C_DOUBLE)*g%ny*g%nz/seconds_per_step/1.0d6

    call destroy_pressure_solver(ps)
    call exit_ibm_data(ibm, g)
    call destroy_output_workspace(out, g)
    call exit_field_data(f, g)
    if (len_trim(mdx_dump_final) > 0) call dump_fields(trim(mdx_dump_final))

contains

    ! the parsed run configuration as `key value` lines (second argument --print-config), no GPU use
    subroutine print_config()
        integer :: d, s, q
        write(*,'(A,3(1X,I0))') "grid", g%nx, g%ny, g%nz
        write(*,'(A,3(1X,ES24.16))') "lengths", g%lx, g%ly, g%lz
        write(*,'(A,1X,ES24.16)') "re", g%re
        write(*,'(A,3(1X,ES24.16))') "forcing", g%forcing_x, g%forcing_y, g%forcing_z
        write(*,'(A,1X,ES24.16)') "dt", g%dt
        write(*,'(A,1X,I0)') "nsteps", g%nsteps
        write(*,'(A,1X,ES24.16)') "t_final", g%t_final
        write(*,'(A,2(1X,ES24.16))') "cfl", g%cflmax, g%dtmax
        write(*,'(A,1X,ES24.16)') "stretch_beta", g%stretch_beta
        write(*,'(A,1X,L1,2(1X,ES24.16))') "flow_rate", g%flow_rate, g%ub_target, g%flow_height
        write(*,'(A,2(1X,I0),1X,ES24.16)') "pressure", ps%solver, ps%nIter, ps%sor
        write(*,'(A,1X,ES24.16,4(1X,I0))') "sor", ps%tol, ps%max_iter, ps%check_every, ps%precision, ps%inner_iters
        write(*,'(A,3(1X,L1))') "periodic", logical(bc%isPeriodic(1)), logical(bc%isPeriodic(2)), &
            logical(bc%isPeriodic(3))
        do q = 0, 3
            do s = 0, 1
                do d = 1, 3
                    write(*,'(A,3(1X,I0),1X,I0,1X,ES24.16)') "bc", d, s, q, bc%bcType(d,s,q), bc%bcValue(d,s,q)
                end do
            end do
        end do
        write(*,'(A,3(1X,I0),5(1X,ES24.16))') "ibm", mdx_geo_enabled, mdx_geo%kind, mdx_geo%n_wave_x, &
            mdx_geo%a, mdx_geo%lambda, mdx_geo%half_height, mdx_geo%amp_x, mdx_geo%phase_x
        write(*,'(A,1X,L1,1X,I0,1X,ES24.16)') "init", mdx_init_poiseuille, mdx_init_seed, mdx_init_amp
        write(*,'(A,1X,I0,1X,A,1X,I0,1X,A)') "output", output_interval, trim(output_prefix), field_interval, &
            trim(field_prefix)
        write(*,'(A,1X,A)') "dump_final", trim(mdx_dump_final)
        write(*,'(A,1X,A)') "restart", trim(mdx_restart_file)
    end subroutine print_config

    ! raw little-endian FP64 dump: header (nx, ny, nz as int32), then un, vn, wn, pn ghosted
    subroutine dump_fields(path)
        character(len=*), intent(in) :: path
        integer :: u
        open(newunit=u, file=path, access="stream", form="unformatted", status="replace", action="write")
        write(u) int(g%nx, c_int32_t), int(g%ny, c_int32_t), int(g%nz, c_int32_t)
        write(u) f%un, f%vn, f%wn, f%pn
        close(u)
        print *, "final fields written to ", path
    end subroutine dump_fields

end program main
