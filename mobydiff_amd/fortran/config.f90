! config.f90 -- input.ini reader with the reference's grammar and interface (config.f90:10-431):
! [section] headers, `key = value` or `key value`, `;`/`#` comments, case-insensitive keys, the
! same key aliases, the boundary key grammar <dir>_<side>_<var>_<type|value>, t_final -> nsteps,
! unparsable values keep the default with a warning, a missing file keeps all defaults.
! New keys (in new sections or as new names, so existing files parse unchanged):
!   [grid] stretch_beta            [flow] flow_rate, ub, flow_height
!   [pressure] solver = fft | sola | redblack | sor, tol, max_iter, check_every,
!              precision = fp64 | mixed, inner_iters
!   [ibm] geometry = bump | wavy | none, a, lambda, half_height, amp_x, phase_x, n_wave_x, second_order
!   [init] profile = zero | poiseuille, seed, amplitude, restart = <HDF5 field file>
!   [output] dump_final = <path>
module config
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, finalize_grid
    use :: pressure_workspace, only: pressure_solver_type
    use :: boundary, only: boundary_type
    use :: mdx_c
    implicit none

contains

subroutine read_runtime_config(g, ps, bc, output_interval, output_prefix, field_interval, field_prefix, input_file)
    type(grid_type), intent(inout) :: g
    type(pressure_solver_type), intent(inout) :: ps
    type(boundary_type), intent(inout) :: bc
    integer, intent(out) :: output_interval, field_interval
    character(len=*), intent(out) :: output_prefix, field_prefix
    character(len=*), intent(in) :: input_file
    integer :: unit, stat, line_no, cut, eq
    character(len=512) :: line, key, val
    character(len=64) :: section
    logical :: exists, nsteps_seen, t_final_seen

    output_interval = 1
    output_prefix = "data"
    field_interval = 0
    field_prefix = "field"
    section = ""
    nsteps_seen = .false.
    t_final_seen = .false.
    mdx_geo = mdx_geometry(0, 1, 1, 0.0d0, 0.0d0, 0.1d0, 1.0d0, 1.0d0)

    inquire(file=trim(input_file), exist=exists)
    if (exists) then
        open(newunit=unit, file=trim(input_file), status="old", action="read", iostat=stat)
        if (stat /= 0) then
            print *, "could not open input file; using defaults: ", trim(input_file)
            exists = .false.
        end if
    else
        print *, "input file not found; using defaults: ", trim(input_file)
    end if

    if (exists) then
        line_no = 0
        do
            read(unit, '(A)', iostat=stat) line
            if (stat /= 0) exit
            line_no = line_no + 1
            cut = scan(line, ";#")
            if (cut > 0) line(cut:) = ""
            line = adjustl(line)
            if (len_trim(line) == 0) cycle
            if (line(1:1) == "[") then
                cut = index(line, "]")
                if (cut > 2) section = to_lower(line(2:cut-1))
                cycle
            end if
            eq = index(line, "=")
            if (eq == 0) eq = scan(line, " "//char(9))
            if (eq > 0) then
                key = adjustl(line(:eq-1))
                val = adjustl(line(eq+1:))
            else
                key = line
                val = ""
            end if
            if (len_trim(key) == 0) cycle
            call apply_value(trim(section), to_lower(trim(key)), trim(val), line_no)
        end do
        close(unit)
    end if

    if (t_final_seen .and. .not. nsteps_seen) g%nsteps = max(1, ceiling(g%t_final/g%dt))
    call finalize_grid(g)

    ! settings the library needs at enter_field_data
    mdx_cfg%periodic = merge(1, 0, logical(bc%isPeriodic))
    mdx_cfg%bc_type = reshape(bc%bcType, [24])
    mdx_cfg%bc_value = reshape(bc%bcValue, [24])
    mdx_cfg%solver = ps%solver
    mdx_cfg%niter = ps%nIter
    mdx_cfg%omega = ps%sor
    mdx_cfg%tol = ps%tol
    mdx_cfg%max_iter = ps%max_iter
    mdx_cfg%check_every = ps%check_every
    mdx_cfg%precision = ps%precision
    mdx_cfg%inner_iters = ps%inner_iters
    mdx_cfg%guess = 0
    mdx_cfg%timing = 0

contains

    subroutine apply_value(sec, k, v, ln)
        character(len=*), intent(in) :: sec, k, v
        integer, intent(in) :: ln
        integer :: iv
        select case (sec)
        case ("grid")
            select case (k)
            case ("nx"); call get_int(v, g%nx, ln)
            case ("ny"); call get_int(v, g%ny, ln)
            case ("nz"); call get_int(v, g%nz, ln)
            case ("lx"); call get_real(v, g%lx, ln)
            case ("ly"); call get_real(v, g%ly, ln)
            case ("lz"); call get_real(v, g%lz, ln)
            case ("stretch_beta", "beta"); call get_real(v, g%stretch_beta, ln)
            end select
        case ("flow")
            select case (k)
            case ("re"); call get_real(v, g%re, ln)
            case ("forcing_x", "force_x", "body_force_x", "meanpx"); call get_real(v, g%forcing_x, ln)
            case ("forcing_y", "force_y", "body_force_y", "meanpy"); call get_real(v, g%forcing_y, ln)
            case ("forcing_z", "force_z", "body_force_z", "meanpz"); call get_real(v, g%forcing_z, ln)
            case ("flow_rate", "constant_flow_rate"); call get_bool(v, g%flow_rate, ln)
            case ("ub", "bulk_velocity"); call get_real(v, g%ub_target, ln)
            case ("flow_height"); call get_real(v, g%flow_height, ln)
            end select
        case ("time")
            select case (k)
            case ("dt"); call get_real(v, g%dt, ln)
            case ("nsteps", "steps"); call get_int(v, g%nsteps, ln); nsteps_seen = .true.
            case ("t_final"); call get_real(v, g%t_final, ln); t_final_seen = .true.
            case ("cflmax"); call get_real(v, g%cflmax, ln)
            case ("dtmax"); call get_real(v, g%dtmax, ln)
            end select
        case ("output")
            select case (k)
            case ("interval", "output_interval"); call get_int(v, output_interval, ln)
            case ("prefix"); output_prefix = unquote(v)
            case ("field_interval", "hdf5_interval", "h5_interval"); call get_int(v, field_interval, ln)
            case ("field_prefix", "hdf5_prefix", "h5_prefix"); field_prefix = unquote(v)
            case ("dump_final"); mdx_dump_final = unquote(v)
            end select
        case ("pressure")
            select case (k)
            case ("niter", "n_iter", "iterations")
                iv = ps%nIter
                call get_int(v, iv, ln)
                if (iv > 0) then
                    ps%nIter = int(iv, C_INT)
                else
                    print *, "warning: pressure nIter must be positive on input line", ln
                end if
            case ("sor", "omega"); call get_real(v, ps%sor, ln)
            case ("solver")
                select case (to_lower(unquote(v)))
                case ("fft"); ps%solver = 0
                case ("sola", "redblack", "red-black", "rb"); ps%solver = 1
                case ("sor", "sor_tol", "tolerance"); ps%solver = 2
                case default; print *, "warning: unknown pressure solver on input line", ln
                end select
            case ("tol", "tolerance"); call get_real(v, ps%tol, ln)
            case ("max_iter"); iv = ps%max_iter; call get_int(v, iv, ln); ps%max_iter = int(iv, C_INT)
            case ("check_every"); iv = ps%check_every; call get_int(v, iv, ln); ps%check_every = int(iv, C_INT)
            case ("inner_iters"); iv = ps%inner_iters; call get_int(v, iv, ln); ps%inner_iters = int(iv, C_INT)
            case ("precision")
                select case (to_lower(unquote(v)))
                case ("fp64", "double"); ps%precision = 0
                case ("mixed", "fp32"); ps%precision = 1
                case default; print *, "warning: precision must be fp64 or mixed on input line", ln
                end select
            end select
        case ("ibm")
            select case (k)
            case ("geometry")
                select case (to_lower(unquote(v)))
                case ("bump", "reference"); mdx_geo%kind = 0; mdx_geo_enabled = 1
                case ("wavy", "wavy_channel"); mdx_geo%kind = 1; mdx_geo_enabled = 1
                case ("none", "off", "flat"); mdx_geo_enabled = 0
                case default; print *, "warning: unknown ibm geometry on input line", ln
                end select
            case ("a", "amplitude"); call get_real(v, mdx_geo%a, ln)
            case ("lambda", "wavelength"); call get_real(v, mdx_geo%lambda, ln)
            case ("half_height"); call get_real(v, mdx_geo%half_height, ln)
            case ("amp_x"); call get_real(v, mdx_geo%amp_x, ln)
            case ("phase_x"); call get_real(v, mdx_geo%phase_x, ln)
            case ("n_wave_x"); iv = mdx_geo%n_wave_x; call get_int(v, iv, ln); mdx_geo%n_wave_x = iv
            case ("second_order")
                block
                    logical :: so
                    so = mdx_geo%second_order /= 0
                    call get_bool(v, so, ln)
                    mdx_geo%second_order = merge(1, 0, so)
                end block
            end select
        case ("init")
            select case (k)
            case ("profile"); mdx_init_poiseuille = to_lower(unquote(v)) == "poiseuille"
            case ("seed"); call get_int(v, mdx_init_seed, ln)
            case ("amplitude", "amp"); call get_real(v, mdx_init_amp, ln)
            case ("restart", "restart_file", "field_file"); mdx_restart_file = unquote(v)
            end select
        case ("boundary")
            call boundary_value(k, v, ln)
        end select
    end subroutine apply_value

    subroutine boundary_value(k, v, ln)
        character(len=*), intent(in) :: k, v
        integer, intent(in) :: ln
        integer :: p1, p2, p3, d, s, q
        character(len=64) :: part(4)
        select case (k)
        case ("periodic_x", "x_periodic", "isperiodic_x", "x_is_periodic"); call get_cbool(v, bc%isPeriodic(1), ln)
        case ("periodic_y", "y_periodic", "isperiodic_y", "y_is_periodic"); call get_cbool(v, bc%isPeriodic(2), ln)
        case ("periodic_z", "z_periodic", "isperiodic_z", "z_is_periodic"); call get_cbool(v, bc%isPeriodic(3), ln)
        case default
            d = 0; s = -1; q = -1
            p1 = index(k, "_")
            if (p1 > 1) then
                p2 = index(k(p1+1:), "_")
                if (p2 > 1) then
                    p2 = p1 + p2
                    p3 = index(k(p2+1:), "_")
                    if (p3 > 1) then
                        p3 = p2 + p3
                        part(1) = k(:p1-1); part(2) = k(p1+1:p2-1); part(3) = k(p2+1:p3-1); part(4) = k(p3+1:)
                        select case (trim(part(1)))
                        case ("x", "i", "streamwise"); d = 1
                        case ("y", "j", "wallnormal", "wall_normal"); d = 2
                        case ("z", "k", "spanwise"); d = 3
                        end select
                        select case (trim(part(2)))
                        case ("min", "lo", "low", "lower", "left", "bottom", "front"); s = 0
                        case ("max", "hi", "high", "upper", "right", "top", "back"); s = 1
                        end select
                        select case (trim(part(3)))
                        case ("p", "pressure"); q = 0
                        case ("u", "un"); q = 1
                        case ("v", "vn"); q = 2
                        case ("w", "wn"); q = 3
                        end select
                    end if
                end if
            end if
            if (d == 0 .or. s < 0 .or. q < 0) then
                print *, "warning: unknown boundary key on input line", ln, ": ", trim(k)
                return
            end if
            select case (trim(part(4)))
            case ("type"); call get_bctype(v, bc%bcType(d,s,q), ln)
            case ("value"); call get_real(v, bc%bcValue(d,s,q), ln)
            case default; print *, "warning: boundary key must end in _type or _value on input line", ln
            end select
        end select
    end subroutine boundary_value

end subroutine read_runtime_config

subroutine get_int(v, target, ln)
    character(len=*), intent(in) :: v
    integer, intent(inout) :: target
    integer, intent(in) :: ln
    integer :: stat, x
    read(v, *, iostat=stat) x
    if (stat == 0) then
        target = x
    else
        print *, "warning: could not parse integer on input line", ln
    end if
end subroutine get_int

subroutine get_real(v, target, ln)
    character(len=*), intent(in) :: v
    real(C_DOUBLE), intent(inout) :: target
    integer, intent(in) :: ln
    integer :: stat
    real(C_DOUBLE) :: x
    read(v, *, iostat=stat) x
    if (stat == 0) then
        target = x
    else
        print *, "warning: could not parse real value on input line", ln
    end if
end subroutine get_real

subroutine get_bool(v, target, ln)
    character(len=*), intent(in) :: v
    logical, intent(inout) :: target
    integer, intent(in) :: ln
    select case (to_lower(unquote(v)))
    case ("true", "t", ".true.", "yes", "y", "1", "periodic"); target = .true.
    case ("false", "f", ".false.", "no", "n", "0", "nonperiodic", "non-periodic"); target = .false.
    case default; print *, "warning: could not parse logical value on input line", ln
    end select
end subroutine get_bool

subroutine get_cbool(v, target, ln)
    character(len=*), intent(in) :: v
    logical(C_BOOL), intent(inout) :: target
    integer, intent(in) :: ln
    logical :: b
    b = target
    call get_bool(v, b, ln)
    target = logical(b, kind=C_BOOL)
end subroutine get_cbool

subroutine get_bctype(v, target, ln)
    character(len=*), intent(in) :: v
    integer(C_INT), intent(inout) :: target
    integer, intent(in) :: ln
    integer :: stat, x
    select case (to_lower(unquote(v)))
    case ("dirichlet", "d", "fixed", "value", "0"); target = 0
    case ("neumann", "n", "gradient", "normal_gradient", "1"); target = 1
    case default
        read(v, *, iostat=stat) x
        if (stat == 0 .and. (x == 0 .or. x == 1)) then
            target = int(x, C_INT)
        else
            print *, "warning: boundary type must be dirichlet/0 or neumann/1 on input line", ln
        end if
    end select
end subroutine get_bctype

function to_lower(s) result(o)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: o
    integer :: i, c
    do i = 1, len(s)
        c = iachar(s(i:i))
        if (c >= iachar("A") .and. c <= iachar("Z")) c = c + 32
        o(i:i) = achar(c)
    end do
end function to_lower

function unquote(s) result(o)
    character(len=*), intent(in) :: s
    character(len=:), allocatable :: o
    integer :: n
    o = trim(adjustl(s))
    n = len(o)
    if (n >= 2) then
        if ((o(1:1) == '"' .and. o(n:n) == '"') .or. (o(1:1) == "'" .and. o(n:n) == "'")) o = o(2:n-1)
    end if
end function unquote

end module config
