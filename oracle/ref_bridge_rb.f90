! ref_bridge_rb.f90 -- C-callable entry points into the UNMODIFIED reference modules of the
! red-black build (USE_REDBLACK, USE_IBM_SECONDORDER), compiled from their sources in place by
! oracle/Makefile.  Test infrastructure only: it lets the test suite run the reference's own
! momentum (step.f90:27), apply_bc_redblack (boundary.f90:173), pressure_projection_redblack
! (pressure_redblack.f90:27), set_ibm_coeff (ibm.f90:133) and get_cfl (step.f90:210) on
! arbitrary inputs and compare them with the C oracle, and run the time loop of main.f90:64-97
! (RK3 stages with the module's own rk_alpha/rk_beta/rk_gamma, the pressure_projection dispatch
! of pressure_solver.f90:48-63, and the CFL-driven dt update) over those procedures.
module ref_bridge_rb
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type, init_grid, finalize_grid
    use :: boundary, only: boundary_type, init_bc, apply_bc_redblack
    use :: step, only: momentum, get_cfl, rk_alpha, rk_beta, rk_gamma
    use :: pressure_solver, only: init_pressure_solver, pressure_projection, destroy_pressure_solver
    use :: ibmm, only: ibm_type, init_ibm, set_ibm_coeff
    use :: pressure_workspace, only: pressure_solver_type
    use :: pressure_redblack, only: pressure_projection_redblack
    implicit none
contains

    subroutine make_grid(g, nx, ny, nz, lx, ly, lz, re, forcing)
        type(grid_type), intent(inout) :: g
        integer(c_int), intent(in) :: nx, ny, nz
        real(c_double), intent(in) :: lx, ly, lz, re, forcing(3)
        call init_grid(g)
        g%nx = nx; g%ny = ny; g%nz = nz
        g%lx = lx; g%ly = ly; g%lz = lz
        g%re = re
        g%forcing_x = forcing(1); g%forcing_y = forcing(2); g%forcing_z = forcing(3)
        call finalize_grid(g)
    end subroutine make_grid

    subroutine make_bc(bc, periodic, bctype, bcvalue)
        type(boundary_type), intent(inout) :: bc
        integer(c_int), intent(in) :: periodic(3), bctype(3,0:1,0:3)
        real(c_double), intent(in) :: bcvalue(3,0:1,0:3)
        integer :: d
        call init_bc(bc)
        do d = 1, 3
            bc%isPeriodic(d) = logical(periodic(d) /= 0, kind=c_bool)
        end do
        bc%bcType = bctype
        bc%bcValue = bcvalue
    end subroutine make_bc

    subroutine alloc_fields(f, nx, ny, nz)
        type(field_type), intent(inout) :: f
        integer(c_int), intent(in) :: nx, ny, nz
        allocate(f%un(0:nx+1,0:ny+1,0:nz+1), f%us(0:nx+1,0:ny+1,0:nz+1), f%oldrhsu(1:nx,1:ny,1:nz))
        allocate(f%vn(0:nx+1,0:ny+1,0:nz+1), f%vs(0:nx+1,0:ny+1,0:nz+1), f%oldrhsv(1:nx,1:ny,1:nz))
        allocate(f%wn(0:nx+1,0:ny+1,0:nz+1), f%ws(0:nx+1,0:ny+1,0:nz+1), f%oldrhsw(1:nx,1:ny,1:nz))
        allocate(f%pn(0:nx+1,0:ny+1,0:nz+1))
    end subroutine alloc_fields

    subroutine ref_momentum(nx, ny, nz, lx, ly, lz, re, forcing, periodic, &
            un, vn, wn, pn, us, vs, ws, oru, orv, orw, cu, cv, cw, dta, dtb, dtg) bind(C, name="ref_momentum")
        integer(c_int), value :: nx, ny, nz
        real(c_double), value :: lx, ly, lz, re, dta, dtb, dtg
        real(c_double), intent(in) :: forcing(3)
        integer(c_int), intent(in) :: periodic(3)
        real(c_double), intent(in) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: wn(0:nx+1,0:ny+1,0:nz+1), pn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: us(0:nx+1,0:ny+1,0:nz+1), vs(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: ws(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: oru(nx,ny,nz), orv(nx,ny,nz), orw(nx,ny,nz)
        real(c_double), intent(in) :: cu(0:nx+1,0:ny+1,0:nz+1), cv(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: cw(0:nx+1,0:ny+1,0:nz+1)
        type(grid_type) :: g
        type(field_type) :: f
        type(ibm_type) :: ibm
        type(boundary_type) :: bc
        integer(c_int) :: bt(3,0:1,0:3)
        real(c_double) :: bv(3,0:1,0:3)
        bt = 0; bv = 0.0d0
        call make_grid(g, nx, ny, nz, lx, ly, lz, re, forcing)
        call make_bc(bc, periodic, bt, bv)
        call alloc_fields(f, nx, ny, nz)
        call init_ibm(ibm, g)
        f%un = un; f%vn = vn; f%wn = wn; f%pn = pn
        f%us = us; f%vs = vs; f%ws = ws
        f%oldrhsu = oru; f%oldrhsv = orv; f%oldrhsw = orw
        ibm%coef_u = cu; ibm%coef_v = cv; ibm%coef_w = cw
        call momentum(f, g, dta, dtb, dtg, ibm, bc)
        us = f%us; vs = f%vs; ws = f%ws
        oru = f%oldrhsu; orv = f%oldrhsv; orw = f%oldrhsw
    end subroutine ref_momentum

    real(c_double) function ref_get_cfl(nx, ny, nz, lx, ly, lz, un, vn, wn) bind(C, name="ref_get_cfl")
        integer(c_int), value :: nx, ny, nz
        real(c_double), value :: lx, ly, lz
        real(c_double), intent(in) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: wn(0:nx+1,0:ny+1,0:nz+1)
        type(grid_type) :: g
        type(field_type) :: f
        real(c_double) :: forcing(3)
        forcing = 0.0d0
        call make_grid(g, nx, ny, nz, lx, ly, lz, 1.0d0, forcing)
        call alloc_fields(f, nx, ny, nz)
        f%un = un; f%vn = vn; f%wn = wn
        ref_get_cfl = get_cfl(f, g)
    end function ref_get_cfl

    subroutine ref_apply_bc_redblack(nx, ny, nz, lx, ly, lz, periodic, bctype, bcvalue, &
            un, vn, wn, pn) bind(C, name="ref_apply_bc_redblack")
        integer(c_int), value :: nx, ny, nz
        real(c_double), value :: lx, ly, lz
        integer(c_int), intent(in) :: periodic(3), bctype(3,0:1,0:3)
        real(c_double), intent(in) :: bcvalue(3,0:1,0:3)
        real(c_double), intent(inout) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: wn(0:nx+1,0:ny+1,0:nz+1), pn(0:nx+1,0:ny+1,0:nz+1)
        type(grid_type) :: g
        type(field_type) :: f
        type(boundary_type) :: bc
        real(c_double) :: forcing(3)
        forcing = 0.0d0
        call make_grid(g, nx, ny, nz, lx, ly, lz, 1.0d0, forcing)
        call make_bc(bc, periodic, bctype, bcvalue)
        call alloc_fields(f, nx, ny, nz)
        f%un = un; f%vn = vn; f%wn = wn; f%pn = pn
        call apply_bc_redblack(f, g, bc)
        un = f%un; vn = f%vn; wn = f%wn; pn = f%pn
    end subroutine ref_apply_bc_redblack

    subroutine ref_projection_redblack(nx, ny, nz, lx, ly, lz, periodic, bctype, bcvalue, &
            niter, sor, dtg, cu, cv, cw, us, vs, ws, un, vn, wn, pn) bind(C, name="ref_projection_redblack")
        integer(c_int), value :: nx, ny, nz, niter
        real(c_double), value :: lx, ly, lz, sor, dtg
        integer(c_int), intent(in) :: periodic(3), bctype(3,0:1,0:3)
        real(c_double), intent(in) :: bcvalue(3,0:1,0:3)
        real(c_double), intent(in) :: cu(0:nx+1,0:ny+1,0:nz+1), cv(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: cw(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: us(0:nx+1,0:ny+1,0:nz+1), vs(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: ws(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: wn(0:nx+1,0:ny+1,0:nz+1), pn(0:nx+1,0:ny+1,0:nz+1)
        type(grid_type) :: g
        type(field_type) :: f
        type(boundary_type) :: bc
        type(ibm_type) :: ibm
        type(pressure_solver_type) :: ps
        real(c_double) :: forcing(3)
        forcing = 0.0d0
        call make_grid(g, nx, ny, nz, lx, ly, lz, 1.0d0, forcing)
        call make_bc(bc, periodic, bctype, bcvalue)
        call alloc_fields(f, nx, ny, nz)
        call init_ibm(ibm, g)
        ibm%coef_u = cu; ibm%coef_v = cv; ibm%coef_w = cw
        f%us = us; f%vs = vs; f%ws = ws
        f%un = un; f%vn = vn; f%wn = wn; f%pn = pn
        ps%nIter = niter
        ps%sor = sor
        call pressure_projection_redblack(ps, f, g, dtg, ibm, bc)
        un = f%un; vn = f%vn; wn = f%wn; pn = f%pn
    end subroutine ref_projection_redblack

    subroutine ref_set_ibm_coeff(nx, ny, nz, lx, ly, lz, re, amp_x, phase_x, n_wave_x, &
            dix, diy, diz, coeff) bind(C, name="ref_set_ibm_coeff")
        integer(c_int), value :: nx, ny, nz, n_wave_x, dix, diy, diz
        real(c_double), value :: lx, ly, lz, re, amp_x, phase_x
        real(c_double), intent(out) :: coeff(0:nx+1,0:ny+1,0:nz+1)
        type(grid_type) :: g
        type(ibm_type) :: ibm
        real(c_double) :: forcing(3)
        forcing = 0.0d0
        call make_grid(g, nx, ny, nz, lx, ly, lz, re, forcing)
        call init_ibm(ibm, g)
        if (amp_x > 0.0d0) ibm%amp_x = amp_x
        ibm%phase_x = phase_x
        if (n_wave_x > 0) ibm%n_wave_x = n_wave_x
        call set_ibm_coeff(g, ibm, ibm%coef_u, dix, diy, diz)
        coeff = ibm%coef_u
    end subroutine ref_set_ibm_coeff

    ! The time loop of main.f90:64-97 in the red-black build: per step t += dt, three RK stages of
    ! momentum -> pressure_projection -> apply_bc_redblack (main.f90:68-89, no apply_bc under
    ! USE_REDBLACK), then get_cfl and dt = min(cflmax/cfl, dtmax) when cflmax > 0 (main.f90:91-97).
    ! dts(n) receives the dt step n ran with, cfls(n) the CFL computed after it (0 when not due).
    subroutine ref_run_redblack(nx, ny, nz, lx, ly, lz, re, forcing, periodic, bctype, bcvalue, &
            niter, sor, nsteps, dt0, cflmax, dtmax, cu, cv, cw, &
            un, vn, wn, pn, us, vs, ws, oru, orv, orw, dts, cfls) bind(C, name="ref_run_redblack")
        integer(c_int), value :: nx, ny, nz, niter, nsteps
        real(c_double), value :: lx, ly, lz, re, sor, dt0, cflmax, dtmax
        real(c_double), intent(in) :: forcing(3)
        integer(c_int), intent(in) :: periodic(3), bctype(3,0:1,0:3)
        real(c_double), intent(in) :: bcvalue(3,0:1,0:3)
        real(c_double), intent(in) :: cu(0:nx+1,0:ny+1,0:nz+1), cv(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(in) :: cw(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: wn(0:nx+1,0:ny+1,0:nz+1), pn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: us(0:nx+1,0:ny+1,0:nz+1), vs(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: ws(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: oru(nx,ny,nz), orv(nx,ny,nz), orw(nx,ny,nz)
        real(c_double), intent(out) :: dts(nsteps), cfls(nsteps)
        type(grid_type) :: g
        type(field_type) :: f
        type(ibm_type) :: ibm
        type(boundary_type) :: bc
        type(pressure_solver_type) :: ps
        integer :: i, rkStage
        real(c_double) :: dt_alpha, dt_beta, dt_gamma
        call make_grid(g, nx, ny, nz, lx, ly, lz, re, forcing)
        g%dt = dt0
        g%cflmax = cflmax
        g%dtmax = dtmax
        call make_bc(bc, periodic, bctype, bcvalue)
        call alloc_fields(f, nx, ny, nz)
        call init_ibm(ibm, g)
        ibm%coef_u = cu; ibm%coef_v = cv; ibm%coef_w = cw
        f%un = un; f%vn = vn; f%wn = wn; f%pn = pn
        f%us = us; f%vs = vs; f%ws = ws
        f%oldrhsu = oru; f%oldrhsv = orv; f%oldrhsw = orw
        ps%nIter = niter
        ps%sor = sor
        call init_pressure_solver(ps, g)
        do i = 1, nsteps
            g%t_current = g%t_current + g%dt
            dts(i) = g%dt
            do rkStage = 1, 3
                dt_alpha = g%dt*rk_alpha(rkStage)
                dt_beta  = g%dt*rk_beta(rkStage)
                dt_gamma = g%dt*rk_gamma(rkStage)
                call momentum(f, g, dt_alpha, dt_beta, dt_gamma, ibm, bc)
                call pressure_projection(ps, f, g, dt_gamma, ibm, bc)
                call apply_bc_redblack(f, g, bc)
            end do
            cfls(i) = 0.0d0
            if (g%cflmax > 0.0d0) then
                g%cfl = get_cfl(f, g)
                cfls(i) = g%cfl
            end if
            if (g%cflmax > 0.0d0 .and. g%cfl > 0.0d0) then
                g%dt = min(g%cflmax/g%cfl, g%dtmax)
            end if
        end do
        call destroy_pressure_solver(ps)
        un = f%un; vn = f%vn; wn = f%wn; pn = f%pn
        us = f%us; vs = f%vs; ws = f%ws
        oru = f%oldrhsu; orv = f%oldrhsv; orw = f%oldrhsw
    end subroutine ref_run_redblack

end module ref_bridge_rb
