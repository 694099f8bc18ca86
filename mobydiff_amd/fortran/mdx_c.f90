! mdx_c.f90 -- iso_c_binding interface to libmobydiff_hip (include/mobydiff_hip.h) and the
! run-time state the drop-in modules share: the device handle, the solver settings read from
! input.ini, and the immersed-boundary geometry.  Every C call returns an int; mdx_check turns a
! non-zero code into `error stop` with the library's message, as the reference does for cuFFT/CUDA
! return codes (cufft_bindings.f90:56-74).
module mdx_c
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: MDX_SOLVER_FFT = 0, MDX_SOLVER_SOLA = 1, MDX_SOLVER_SOR = 2
    integer(c_int32_t), parameter :: MDX_PREC_FP64 = 0, MDX_PREC_MIXED = 1

    type, bind(C) :: mdx_config
        integer(c_int32_t) :: nx, ny, nz
        real(c_double) :: lx, ly, lz, re
        real(c_double) :: forcing(3)
        integer(c_int32_t) :: periodic(3)
        integer(c_int32_t) :: bc_type(24)
        real(c_double) :: bc_value(24)
        real(c_double) :: stretch_beta
        integer(c_int32_t) :: solver, niter
        real(c_double) :: omega, tol
        integer(c_int32_t) :: max_iter, check_every, precision, inner_iters, guess, flowrate
        real(c_double) :: ub_target, flow_height
        integer(c_int32_t) :: timing
    end type mdx_config

    type, bind(C) :: mdx_geometry
        integer(c_int32_t) :: kind, second_order, n_wave_x
        real(c_double) :: amp_x, phase_x, a, lambda, half_height
    end type mdx_geometry

    type, bind(C) :: mdx_solve_stats
        integer(c_int32_t) :: iterations, converged, fallback, checks
        real(c_double) :: div0_inf, div_inf, guess_div_inf
        integer(c_int32_t) :: lds_sweeps, guess_skipped, lds_restarts
    end type mdx_solve_stats

    ! shared run-time state of the drop-in modules
    type(c_ptr), save :: mdx_h = c_null_ptr
    type(mdx_config), save :: mdx_cfg
    type(mdx_geometry), save :: mdx_geo
    integer, save :: mdx_geo_enabled = 1          ! 0: no immersed boundary (coef = 0)
    logical, save :: mdx_init_poiseuille = .false.
    integer, save :: mdx_init_seed = 0
    real(c_double), save :: mdx_init_amp = 0.0d0
    character(len=256), save :: mdx_dump_final = ""
    character(len=256), save :: mdx_restart_file = ""   ! [init] restart: field file to start from

    interface
        function mdx_last_error() bind(C, name="mdx_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function mdx_create(cfg, h) bind(C, name="mdx_create") result(rc)
            import :: c_int, c_ptr, mdx_config
            type(mdx_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: h
            integer(c_int) :: rc
        end function
        function mdx_destroy(h) bind(C, name="mdx_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function
        function mdx_synchronize(h) bind(C, name="mdx_synchronize") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function
        function mdx_upload_fields(h, un, vn, wn, pn, us, vs, ws, oru, orv, orw) &
                bind(C, name="mdx_upload_fields") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: un(*), vn(*), wn(*), pn(*), us(*), vs(*), ws(*)
            real(c_double), intent(in) :: oru(*), orv(*), orw(*)
            integer(c_int) :: rc
        end function
        function mdx_download_fields(h, un, vn, wn, pn, us, vs, ws, oru, orv, orw) &
                bind(C, name="mdx_download_fields") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: un(*), vn(*), wn(*), pn(*), us(*), vs(*), ws(*)
            real(c_double), intent(inout) :: oru(*), orv(*), orw(*)
            integer(c_int) :: rc
        end function
        ! coefficient arrays by address; three null pointers switch the immersed boundary off
        function mdx_set_ibm(h, cu, cv, cw) bind(C, name="mdx_set_ibm") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h, cu, cv, cw
            integer(c_int) :: rc
        end function
        function mdx_ibm_coefficient(h, geo, dix, diy, diz, coeff) bind(C, name="mdx_ibm_coefficient") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t, mdx_geometry
            type(c_ptr), value :: h
            type(mdx_geometry), intent(in) :: geo
            integer(c_int32_t), value :: dix, diy, diz
            real(c_double), intent(inout) :: coeff(*)
            integer(c_int) :: rc
        end function
        function mdx_momentum(h, dta, dtb, dtg) bind(C, name="mdx_momentum") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: dta, dtb, dtg
            integer(c_int) :: rc
        end function
        function mdx_flowrate_shift(h, dtg, du) bind(C, name="mdx_flowrate_shift") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h, du
            real(c_double), value :: dtg
            integer(c_int) :: rc
        end function
        function mdx_apply_bc(h) bind(C, name="mdx_apply_bc") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function
        function mdx_apply_bc_redblack(h) bind(C, name="mdx_apply_bc_redblack") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function
        function mdx_pressure_projection(h, dtg, st) bind(C, name="mdx_pressure_projection") result(rc)
            import :: c_int, c_ptr, c_double, mdx_solve_stats
            type(c_ptr), value :: h
            real(c_double), value :: dtg
            type(mdx_solve_stats), intent(out) :: st
            integer(c_int) :: rc
        end function
        function mdx_get_cfl(h, cfl) bind(C, name="mdx_get_cfl") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: cfl
            integer(c_int) :: rc
        end function
        function mdx_div_norm(h, d) bind(C, name="mdx_div_norm") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: d
            integer(c_int) :: rc
        end function
        function mdx_init_fields(h, seed, amp, geo, dt) bind(C, name="mdx_init_fields") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h, geo
            integer(c_int32_t), value :: seed
            real(c_double), value :: amp, dt
            integer(c_int) :: rc
        end function
        function mdx_center_velocity(h, uc, vc, wc) bind(C, name="mdx_center_velocity") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: uc(*), vc(*), wc(*)
            integer(c_int) :: rc
        end function
        function mdx_get_y_coordinates(h, yc, yv) bind(C, name="mdx_get_y_coordinates") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(inout) :: yc(*), yv(*)
            integer(c_int) :: rc
        end function
        ! libmobydiff_io (include/mobydiff_io.h): HDF5 field files
        function mdx_io_last_error() bind(C, name="mdx_io_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function mdx_h5_write_field(path, nx, ny, nz, lx, ly, lz, dx, dy, dz, re, dt, t_current, &
                un, vn, wn, pn) bind(C, name="mdx_h5_write_field") result(rc)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: nx, ny, nz
            real(c_double), value :: lx, ly, lz, dx, dy, dz, re, dt, t_current
            real(c_double), intent(in) :: un(*), vn(*), wn(*), pn(*)
            integer(c_int) :: rc
        end function
        function mdx_h5_read_field(path, nx, ny, nz, un, vn, wn, pn) bind(C, name="mdx_h5_read_field") result(rc)
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: nx, ny, nz
            real(c_double), intent(inout) :: un(*), vn(*), wn(*), pn(*)
            integer(c_int) :: rc
        end function
        function mdx_h5_read_field_info(path, nx, ny, nz, dt, t_current) &
                bind(C, name="mdx_h5_read_field_info") result(rc)
            import :: c_int, c_char, c_double, c_int32_t
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int32_t), intent(out) :: nx, ny, nz
            real(c_double), intent(out) :: dt, t_current
            integer(c_int) :: rc
        end function
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

contains

    function last_error_text() result(msg)
        character(len=:), allocatable :: msg
        msg = c_text(mdx_last_error())
    end function last_error_text

    function c_text(p) result(msg)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function c_text

    subroutine mdx_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= 0) then
            print *, "mobydiff_hip error in ", trim(what), ": ", last_error_text()
            error stop "mobydiff_hip call failed"
        end if
    end subroutine mdx_check

end module mdx_c
