! ref_bridge_fft.f90 -- C-callable entry points into the UNMODIFIED reference modules of the
! FFT build, compiled from their sources in place by oracle/Makefile.  Test infrastructure only.
! The FFTW back-end (pressure_backend_fftw.f90) and therefore pressure_solver.f90 cannot be built
! in this image (no FFTW); what builds without it is exported here: apply_bc (boundary.f90:23,
! FFT variant) and the modified-wavenumber Thomas solver (pressure_fft.f90:20-148).
module ref_bridge_fft
    use, intrinsic :: iso_c_binding
    use :: init, only: grid_type, field_type, init_grid, finalize_grid
    use :: boundary, only: apply_bc
    use :: pressure_workspace, only: pressure_solver_type
    use :: pressure_fft, only: allocate_common_workspace, init_tridiag_coefficients, solve_tridiagonal_y
    implicit none
contains

    subroutine ref_apply_bc_fft(nx, ny, nz, un, vn, wn, us, vs, ws, pn) bind(C, name="ref_apply_bc_fft")
        integer(c_int), value :: nx, ny, nz
        real(c_double), intent(inout) :: un(0:nx+1,0:ny+1,0:nz+1), vn(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: wn(0:nx+1,0:ny+1,0:nz+1), us(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: vs(0:nx+1,0:ny+1,0:nz+1), ws(0:nx+1,0:ny+1,0:nz+1)
        real(c_double), intent(inout) :: pn(0:nx+1,0:ny+1,0:nz+1)
        type(grid_type) :: g
        type(field_type) :: f
        call init_grid(g)
        g%nx = nx; g%ny = ny; g%nz = nz
        call finalize_grid(g)
        allocate(f%un(0:nx+1,0:ny+1,0:nz+1), f%vn(0:nx+1,0:ny+1,0:nz+1), f%wn(0:nx+1,0:ny+1,0:nz+1))
        allocate(f%us(0:nx+1,0:ny+1,0:nz+1), f%vs(0:nx+1,0:ny+1,0:nz+1), f%ws(0:nx+1,0:ny+1,0:nz+1))
        allocate(f%pn(0:nx+1,0:ny+1,0:nz+1))
        f%un = un; f%vn = vn; f%wn = wn; f%us = us; f%vs = vs; f%ws = ws; f%pn = pn
        call apply_bc(f, g)
        un = f%un; vn = f%vn; wn = f%wn; us = f%us; vs = f%vs; ws = f%ws; pn = f%pn
    end subroutine ref_apply_bc_fft

    ! plane_hat(nxh,nz,ny) complex in/out; cp/den_inv(nxh,ny,nz) complex out
    subroutine ref_tridiag(nx, ny, nz, lx, ly, lz, plane_hat, cp, den_inv) bind(C, name="ref_tridiag")
        integer(c_int), value :: nx, ny, nz
        real(c_double), value :: lx, ly, lz
        complex(c_double_complex), intent(inout) :: plane_hat(nx/2+1, nz, ny)
        complex(c_double_complex), intent(out) :: cp(nx/2+1, ny, nz), den_inv(nx/2+1, ny, nz)
        type(grid_type) :: g
        type(pressure_solver_type) :: ps
        integer :: nxh
        call init_grid(g)
        g%nx = nx; g%ny = ny; g%nz = nz
        g%lx = lx; g%ly = ly; g%lz = lz
        call finalize_grid(g)
        nxh = nx/2 + 1
        call allocate_common_workspace(ps, nx, nxh, ny, nz)
        call init_tridiag_coefficients(ps, g)
        ps%plane_hat = plane_hat
        call solve_tridiagonal_y(ps, nxh, ny, nz, 1.0d0/g%dy**2)
        plane_hat = ps%plane_hat
        cp = ps%cp_hat
        den_inv = ps%den_inv_hat
    end subroutine ref_tridiag

end module ref_bridge_fft
