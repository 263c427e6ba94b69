! ------------------------------------------------------------------------------
! C-callable bridge onto the unmodified reference Fortran library
! (EdSantilli/MGPoisson2D, src/libBaseTools + src/libMGPoisson).
!
! TEST INFRASTRUCTURE ONLY.  This file is compiled together with the reference
! sources, where they lie, into oracle/_ref/libmgref*.so by oracle/Makefile.
! Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
! that library, and only as the checker / the timed CPU baseline.
!
! Conventions of every entry point (all arrays column-major, i fastest):
!   valid cells i = 0..nx-1, j = 0..ny-1 (the reference's tests use ilo=jlo=0,
!   tests/test.f90:29-32), grid spacing dx, dy.
!   J(0:nx-1,0:ny-1); Jxx,Jxy(0:nx,0:ny-1); Jyx,Jyy(0:nx-1,0:ny)  (no ghosts:
!   the reference never reads metric ghosts).
!   phi-like fields carry exactly one ghost layer: phi(-1:nx,-1:ny).
!   rhs / residual / invdiags fields carry none: f(0:nx-1,0:ny-1).
!   bct(4), bcm(4): BC type / mode for xlo, xhi, ylo, yhi with the reference's
!   integer codes (arrayutils.f90:27-49).  bd_* point at the face data: one
!   value for UNIFORM/LAH, ny values (x faces) or nx values (y faces) for
!   NONUNIFORM.
!   Logical flags are passed as C ints (0 = .false.).
! ------------------------------------------------------------------------------
module ref_bridge
    use iso_c_binding
    use ArrayUtils
    implicit none

contains

    subroutine mk_valid (valid, nx, ny, dx, dy)
        type(box), intent(out) :: valid
        integer, intent(in)    :: nx, ny
        real(dp), intent(in)   :: dx, dy
        call define_box (valid, 0, nx-1, 0, ny-1, dx, dy)
    end subroutine mk_valid

    subroutine mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        type(geo_data), intent(out) :: geo
        type(box), intent(in)       :: valid
        real(dp), intent(in)        :: J(0:valid%nx-1, 0:valid%ny-1)
        real(dp), intent(in)        :: Jxx(0:valid%nx, 0:valid%ny-1), Jxy(0:valid%nx, 0:valid%ny-1)
        real(dp), intent(in)        :: Jyx(0:valid%nx-1, 0:valid%ny), Jyy(0:valid%nx-1, 0:valid%ny)

        call define_box_data (geo%J, valid, 0, 0, BD_CELL, BD_CELL)
        call define_box_data (geo%Jgup_xx, valid, 0, 0, BD_NODE, BD_CELL)
        call define_box_data (geo%Jgup_xy, valid, 0, 0, BD_NODE, BD_CELL)
        call define_box_data (geo%Jgup_yx, valid, 0, 0, BD_CELL, BD_NODE)
        call define_box_data (geo%Jgup_yy, valid, 0, 0, BD_CELL, BD_NODE)
        geo%J%data = J
        geo%Jgup_xx%data = Jxx
        geo%Jgup_xy%data = Jxy
        geo%Jgup_yx%data = Jyx
        geo%Jgup_yy%data = Jyy
        geo%dx = valid%dx
        geo%dy = valid%dy
    end subroutine mk_geo

    subroutine free_geo (geo)
        type(geo_data), intent(inout) :: geo
        call undefine_box_data (geo%J)
        call undefine_box_data (geo%Jgup_xx)
        call undefine_box_data (geo%Jgup_xy)
        call undefine_box_data (geo%Jgup_yx)
        call undefine_box_data (geo%Jgup_yy)
    end subroutine free_geo

    subroutine copy_side (dst, mode, src, n)
        real(dp), dimension(:), intent(inout) :: dst
        integer, intent(in)                   :: mode, n
        type(c_ptr), intent(in)               :: src
        real(dp), pointer                     :: s(:)
        if (.not. c_associated(src)) then
            dst = zero
            return
        endif
        if (mode .eq. BCMODE_NONUNIFORM) then
            call c_f_pointer (src, s, [n])
            dst(1:n) = s(1:n)
        else
            call c_f_pointer (src, s, [1])
            dst(1) = s(1)
        endif
    end subroutine copy_side

    subroutine mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        type(bdry_data), intent(out) :: bc
        type(box), intent(in)        :: valid
        integer(c_int), intent(in)   :: bct(4), bcm(4)
        type(c_ptr), intent(in)      :: bd_xlo, bd_xhi, bd_ylo, bd_yhi

        call define_bdry_data (bc, valid, bct(1), bct(2), bct(3), bct(4), &
                               bcm(1), bcm(2), bcm(3), bcm(4))
        call copy_side (bc%data_xlo, bcm(1), bd_xlo, valid%ny)
        call copy_side (bc%data_xhi, bcm(2), bd_xhi, valid%ny)
        call copy_side (bc%data_ylo, bcm(3), bd_ylo, valid%nx)
        call copy_side (bc%data_yhi, bcm(4), bd_yhi, valid%nx)
    end subroutine mk_bc

    ! One-ghost cell-centred field.
    subroutine mk_cc1 (f, valid, a)
        type(box_data), intent(out) :: f
        type(box), intent(in)       :: valid
        real(dp), intent(in)        :: a(-1:valid%nx, -1:valid%ny)
        call define_box_data (f, valid, 1, 1, BD_CELL, BD_CELL)
        f%data = a
    end subroutine mk_cc1

    ! Ghost-free cell-centred field.
    subroutine mk_cc0 (f, valid, a)
        type(box_data), intent(out) :: f
        type(box), intent(in)       :: valid
        real(dp), intent(in)        :: a(0:valid%nx-1, 0:valid%ny-1)
        call define_box_data (f, valid, 0, 0, BD_CELL, BD_CELL)
        f%data = a
    end subroutine mk_cc0

    ! --------------------------------------------------------------------------
    ! Poisson2D::compute_laplacian (poisson2d.f90:886)
    ! --------------------------------------------------------------------------
    subroutine ref_laplacian (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                              bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, jscale, phi, lap) &
                              bind(C, name="ref_laplacian")
        use Poisson2D, only: compute_laplacian
        integer(c_int), value       :: nx, ny, homog, jscale
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)
        real(c_double), intent(out) :: lap(0:nx-1,0:ny-1)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, flap

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call define_box_data (flap, valid, 0, 0, BD_CELL, BD_CELL)
        flap%data = zero
        call compute_laplacian (flap, fphi, geo, bc, homog .ne. 0, jscale .ne. 0)
        phi = fphi%data
        lap = flap%data
        call undefine_box_data (fphi)
        call undefine_box_data (flap)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_laplacian

    ! --------------------------------------------------------------------------
    ! Poisson2D::compute_pd (poisson2d.f90:404): pd nodal in x (nodedir 1,
    ! pd(0:nx,0:ny-1)) or in y (nodedir 2, pd(0:nx-1,0:ny)); derivative along dir.
    ! --------------------------------------------------------------------------
    subroutine ref_compute_pd (nx, ny, dx, dy, phi, dir, nodedir, pd) bind(C, name="ref_compute_pd")
        use Poisson2D, only: compute_pd
        integer(c_int), value       :: nx, ny, dir, nodedir
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: phi(-1:nx,-1:ny)
        real(c_double), intent(out) :: pd(*)

        type(box) :: valid
        type(box_data) :: fphi, fpd

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_cc1 (fphi, valid, phi)
        if (nodedir .eq. 1) then
            call define_box_data (fpd, valid, 0, 0, BD_NODE, BD_CELL)
        else
            call define_box_data (fpd, valid, 0, 0, BD_CELL, BD_NODE)
        endif
        fpd%data = zero
        call compute_pd (fpd, fphi, dir)
        pd(1:size(fpd%data)) = reshape(fpd%data, [size(fpd%data)])
        call undefine_box_data (fphi)
        call undefine_box_data (fpd)
    end subroutine ref_compute_pd

    ! --------------------------------------------------------------------------
    ! Poisson2D::compute_grad (poisson2d.f90:809): fills phi's ghosts, returns
    ! J grad phi on x-faces xflux(0:nx,0:ny-1) and y-faces yflux(0:nx-1,0:ny).
    ! --------------------------------------------------------------------------
    subroutine ref_compute_grad (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                                 bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, phi, xflux, yflux) &
                                 bind(C, name="ref_compute_grad")
        use Poisson2D, only: compute_grad
        integer(c_int), value       :: nx, ny, homog
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)
        real(c_double), intent(out) :: xflux(0:nx,0:ny-1), yflux(0:nx-1,0:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, fx, fy, xwk, ywk

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call define_box_data (fx, valid, 0, 0, BD_NODE, BD_CELL)
        call define_box_data (fy, valid, 0, 0, BD_CELL, BD_NODE)
        call define_box_data (xwk, fx)
        call define_box_data (ywk, fy)
        fx%data = zero
        fy%data = zero
        xwk%data = zero
        ywk%data = zero
        call compute_grad (fx, fy, fphi, geo, bc, homog .ne. 0, xwk, ywk)
        phi = fphi%data
        xflux = fx%data
        yflux = fy%data
        call undefine_box_data (fphi)
        call undefine_box_data (fx)
        call undefine_box_data (fy)
        call undefine_box_data (xwk)
        call undefine_box_data (ywk)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_compute_grad

    ! --------------------------------------------------------------------------
    ! Poisson2D::compute_div (poisson2d.f90:860)
    ! --------------------------------------------------------------------------
    subroutine ref_compute_div (nx, ny, dx, dy, xflux, yflux, div) bind(C, name="ref_compute_div")
        use Poisson2D, only: compute_div
        integer(c_int), value       :: nx, ny
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: xflux(0:nx,0:ny-1), yflux(0:nx-1,0:ny)
        real(c_double), intent(out) :: div(0:nx-1,0:ny-1)

        type(box) :: valid
        type(box_data) :: fx, fy, fd

        call mk_valid (valid, nx, ny, dx, dy)
        call define_box_data (fx, valid, 0, 0, BD_NODE, BD_CELL)
        call define_box_data (fy, valid, 0, 0, BD_CELL, BD_NODE)
        call define_box_data (fd, valid, 0, 0, BD_CELL, BD_CELL)
        fx%data = xflux
        fy%data = yflux
        fd%data = zero
        call compute_div (fd, fx, fy)
        div = fd%data
        call undefine_box_data (fx)
        call undefine_box_data (fy)
        call undefine_box_data (fd)
    end subroutine ref_compute_div

    ! --------------------------------------------------------------------------
    ! ArrayUtils::fill_boundary_fluxes (arrayutils.f90:1571)
    ! --------------------------------------------------------------------------
    subroutine ref_fill_boundary_fluxes (nx, ny, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, xflux, yflux) &
                                         bind(C, name="ref_fill_boundary_fluxes")
        integer(c_int), value       :: nx, ny, homog
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(inout) :: xflux(0:nx,0:ny-1), yflux(0:nx-1,0:ny)

        type(box) :: valid
        type(bdry_data) :: bc
        type(box_data) :: fx, fy

        call mk_valid (valid, nx, ny, one, one)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call define_box_data (fx, valid, 0, 0, BD_NODE, BD_CELL)
        call define_box_data (fy, valid, 0, 0, BD_CELL, BD_NODE)
        fx%data = xflux
        fy%data = yflux
        call fill_boundary_fluxes (fx, fy, bc, homog .ne. 0)
        xflux = fx%data
        yflux = fy%data
        call undefine_box_data (fx)
        call undefine_box_data (fy)
        call undefine_bdry_data (bc)
    end subroutine ref_fill_boundary_fluxes

    ! --------------------------------------------------------------------------
    ! Poisson2D::compute_residual (poisson2d.f90:1327)
    ! --------------------------------------------------------------------------
    subroutine ref_residual (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                             bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, rhs, phi, res) &
                             bind(C, name="ref_residual")
        use Poisson2D, only: compute_residual
        integer(c_int), value       :: nx, ny, homog
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(in)  :: rhs(0:nx-1,0:ny-1)
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)
        real(c_double), intent(out) :: res(0:nx-1,0:ny-1)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs, fres

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        call define_box_data (fres, valid, 0, 0, BD_CELL, BD_CELL)
        call compute_residual (fres, frhs, fphi, geo, bc, homog .ne. 0)
        phi = fphi%data
        res = fres%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_box_data (fres)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_residual

    ! --------------------------------------------------------------------------
    ! Poisson2D::compute_invdiags (poisson2d.f90:16)
    ! --------------------------------------------------------------------------
    subroutine ref_invdiags (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                             bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, invd) &
                             bind(C, name="ref_invdiags")
        use Poisson2D, only: compute_invdiags
        integer(c_int), value       :: nx, ny, homog
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(out) :: invd(0:nx-1,0:ny-1)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: f

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call define_box_data (f, valid, 0, 0, BD_CELL, BD_CELL)
        call compute_invdiags (f, geo, bc, homog .ne. 0)
        invd = f%data
        call undefine_box_data (f)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_invdiags

    ! --------------------------------------------------------------------------
    ! Poisson2D::relax_gs (poisson2d.f90:1460), tol < 0 (fixed sweep count)
    ! --------------------------------------------------------------------------
    subroutine ref_relax_gs (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                             bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, invd, omega, maxiters, rhs, phi) &
                             bind(C, name="ref_relax_gs")
        use Poisson2D, only: relax_gs
        integer(c_int), value       :: nx, ny, homog, maxiters
        real(c_double), value       :: dx, dy, omega
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(in)  :: invd(0:nx-1,0:ny-1), rhs(0:nx-1,0:ny-1)
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs, finv

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        call mk_cc0 (finv, valid, invd)
        call relax_gs (fphi, frhs, geo, bc, homog .ne. 0, finv, omega, -one, maxiters, .false., 0)
        phi = fphi%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_box_data (finv)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_relax_gs

    ! --------------------------------------------------------------------------
    ! DeferredPoisson2D::compute_laplacian (MGPoisson2D_DD.f90:45)
    ! --------------------------------------------------------------------------
    subroutine ref_dc_laplacian (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                                 bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, phi, lap) &
                                 bind(C, name="ref_dc_laplacian")
        use DeferredPoisson2D, only: compute_laplacian
        integer(c_int), value       :: nx, ny, homog
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)
        real(c_double), intent(out) :: lap(0:nx-1,0:ny-1)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, flap

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call define_box_data (flap, valid, 0, 0, BD_CELL, BD_CELL)
        call compute_laplacian (flap, fphi, geo, bc, homog .ne. 0)
        phi = fphi%data
        lap = flap%data
        call undefine_box_data (fphi)
        call undefine_box_data (flap)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_dc_laplacian

    ! --------------------------------------------------------------------------
    ! DeferredPoisson2D::compute_invdiags (MGPoisson2D_DD.f90:16)
    ! --------------------------------------------------------------------------
    subroutine ref_dc_invdiags (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, invd) &
                                bind(C, name="ref_dc_invdiags")
        use DeferredPoisson2D, only: compute_invdiags
        integer(c_int), value       :: nx, ny
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        real(c_double), intent(out) :: invd(0:nx-1,0:ny-1)

        type(box) :: valid
        type(geo_data) :: geo
        type(box_data) :: f

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call define_box_data (f, valid, 0, 0, BD_CELL, BD_CELL)
        call compute_invdiags (f, geo)
        invd = f%data
        call undefine_box_data (f)
        call free_geo (geo)
    end subroutine ref_dc_invdiags

    ! --------------------------------------------------------------------------
    ! DeferredPoisson2D::relax_gs (MGPoisson2D_DD.f90:153), tol < 0
    ! --------------------------------------------------------------------------
    subroutine ref_dc_relax_gs (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                                bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, invd, omega, maxiters, rhs, phi) &
                                bind(C, name="ref_dc_relax_gs")
        use DeferredPoisson2D, only: relax_gs
        integer(c_int), value       :: nx, ny, homog, maxiters
        real(c_double), value       :: dx, dy, omega
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(in)  :: invd(0:nx-1,0:ny-1), rhs(0:nx-1,0:ny-1)
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs, finv

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        call mk_cc0 (finv, valid, invd)
        call relax_gs (fphi, frhs, geo, bc, homog .ne. 0, finv, omega, -one, maxiters, .false., 0)
        phi = fphi%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_box_data (finv)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_dc_relax_gs

    ! --------------------------------------------------------------------------
    ! ArrayUtils::fill_ghosts (arrayutils.f90:1277) and extrapolate_ghosts (:1497)
    ! --------------------------------------------------------------------------
    subroutine ref_fill_ghosts (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                                bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, do_neum, phi) &
                                bind(C, name="ref_fill_ghosts")
        integer(c_int), value       :: nx, ny, homog, do_neum
        real(c_double), value       :: dx, dy
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call fill_ghosts (fphi, bc, geo, homog .ne. 0, do_neum .ne. 0)
        phi = fphi%data
        call undefine_box_data (fphi)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_fill_ghosts

    subroutine ref_extrapolate_ghosts (nx, ny, order, phi) bind(C, name="ref_extrapolate_ghosts")
        integer(c_int), value         :: nx, ny, order
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)
        type(box) :: valid
        type(box_data) :: fphi
        call mk_valid (valid, nx, ny, one, one)
        call mk_cc1 (fphi, valid, phi)
        call extrapolate_ghosts (fphi, order)
        phi = fphi%data
        call undefine_box_data (fphi)
    end subroutine ref_extrapolate_ghosts

    ! --------------------------------------------------------------------------
    ! MGPoisson2D::restrict (MGPoisson2D.f90:135).  kind: 0 = cell-centred,
    ! 1 = x-face (nodal in x), 2 = y-face (nodal in y).
    ! fine/crse are passed as flat arrays of the matching staggered shape.
    ! --------------------------------------------------------------------------
    subroutine ref_restrict (kind, fnx, fny, fine, cnx, cny, crse) bind(C, name="ref_restrict")
        use MGPoisson2D, only: restrict
        integer(c_int), value         :: kind, fnx, fny, cnx, cny
        real(c_double), intent(in)    :: fine(*)
        real(c_double), intent(inout) :: crse(*)

        type(box) :: fv, cv
        type(box_data) :: ff, fc
        integer :: offi, offj, nf, nc

        offi = BD_CELL
        offj = BD_CELL
        if (kind .eq. 1) offi = BD_NODE
        if (kind .eq. 2) offj = BD_NODE
        call define_box (fv, 0, fnx-1, 0, fny-1, one/fnx, one/fny)
        call define_box (cv, 0, cnx-1, 0, cny-1, one/cnx, one/cny)
        call define_box_data (ff, fv, 0, 0, offi, offj)
        call define_box_data (fc, cv, 0, 0, offi, offj)
        nf = size(ff%data)
        nc = size(fc%data)
        ff%data = reshape (fine(1:nf), shape(ff%data))
        call restrict (ff, fc)
        crse(1:nc) = reshape (fc%data, [nc])
        call undefine_box_data (ff)
        call undefine_box_data (fc)
    end subroutine ref_restrict

    ! --------------------------------------------------------------------------
    ! MGPoisson2D::prolong (MGPoisson2D.f90:262).  crse bc is UNIFORM/zero
    ! (or LAH) with the given types, as the cycle drivers build it
    ! (MGPoisson2D.f90:516-522, MGPoisson2D_DD.f90:713-719).
    ! --------------------------------------------------------------------------
    subroutine ref_prolong (fnx, fny, fdx, fdy, fJ, fJxx, fJxy, fJyx, fJyy, &
                            cnx, cny, cdx, cdy, cJ, cJxx, cJxy, cJyx, cJyy, &
                            bct, bcm, order, fine, crse) bind(C, name="ref_prolong")
        use MGPoisson2D, only: prolong
        integer(c_int), value         :: fnx, fny, cnx, cny, order
        real(c_double), value         :: fdx, fdy, cdx, cdy
        real(c_double), intent(in)    :: fJ(0:fnx-1,0:fny-1), fJxx(0:fnx,0:fny-1), fJxy(0:fnx,0:fny-1)
        real(c_double), intent(in)    :: fJyx(0:fnx-1,0:fny), fJyy(0:fnx-1,0:fny)
        real(c_double), intent(in)    :: cJ(0:cnx-1,0:cny-1), cJxx(0:cnx,0:cny-1), cJxy(0:cnx,0:cny-1)
        real(c_double), intent(in)    :: cJyx(0:cnx-1,0:cny), cJyy(0:cnx-1,0:cny)
        integer(c_int), intent(in)    :: bct(4), bcm(4)
        real(c_double), intent(inout) :: fine(-1:fnx,-1:fny), crse(-1:cnx,-1:cny)

        type(box) :: fv, cv
        type(geo_data) :: fgeo, cgeo
        type(bdry_data) :: cbc
        type(box_data) :: ff, fc

        call mk_valid (fv, fnx, fny, fdx, fdy)
        call mk_valid (cv, cnx, cny, cdx, cdy)
        call mk_geo (fgeo, fv, fJ, fJxx, fJxy, fJyx, fJyy)
        call mk_geo (cgeo, cv, cJ, cJxx, cJxy, cJyx, cJyy)
        call define_bdry_data (cbc, cv, bct(1), bct(2), bct(3), bct(4), bcm(1), bcm(2), bcm(3), bcm(4))
        cbc%data_xlo = zero
        cbc%data_xhi = zero
        cbc%data_ylo = zero
        cbc%data_yhi = zero
        call mk_cc1 (ff, fv, fine)
        call mk_cc1 (fc, cv, crse)
        call prolong (ff, fc, fgeo, cgeo, cbc, order)
        fine = ff%data
        crse = fc%data
        call undefine_box_data (ff)
        call undefine_box_data (fc)
        call undefine_bdry_data (cbc)
        call free_geo (fgeo)
        call free_geo (cgeo)
    end subroutine ref_prolong

    ! --------------------------------------------------------------------------
    ! MGPoisson2D::estimate_maxdepth + create_mgschedule (MGPoisson2D.f90:38,67)
    ! refx/refy must hold at least maxdepth_in entries.
    ! --------------------------------------------------------------------------
    subroutine ref_schedule (nx, ny, dx, dy, maxdepth_in, maxdepth_out, refx, refy) &
                             bind(C, name="ref_schedule")
        use MGPoisson2D, only: estimate_maxdepth, create_mgschedule
        integer(c_int), value       :: nx, ny, maxdepth_in
        real(c_double), value       :: dx, dy
        integer(c_int), intent(out) :: maxdepth_out
        integer(c_int), intent(out) :: refx(*), refy(*)

        type(box) :: valid
        integer :: md
        integer, allocatable :: rx(:), ry(:)

        call mk_valid (valid, nx, ny, dx, dy)
        md = maxdepth_in
        if (md .lt. 0) md = 1000
        call estimate_maxdepth (md, valid)
        allocate (rx(0:md-1), ry(0:md-1))
        call create_mgschedule (rx, ry, md, valid)
        maxdepth_out = md
        if (md .gt. 0) then
            refx(1:md) = rx(0:md-1)
            refy(1:md) = ry(0:md-1)
        endif
        deallocate (rx, ry)
    end subroutine ref_schedule

    ! --------------------------------------------------------------------------
    ! MGPoisson2D::vcycle (MGPoisson2D.f90:400) when kind = 0,
    ! DeferredMGPoisson2D::vcycle (MGPoisson2D_DD.f90:589) when kind = 1, or
    ! MGPoisson2D::fmg (MGPoisson2D.f90:860) when kind = 2.
    ! --------------------------------------------------------------------------
    subroutine ref_vcycle (kind, nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                           bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, tol, maxiters, maxdepth, &
                           numcycles, sdown, sup, sbottom, zerophi, phi, rhs) &
                           bind(C, name="ref_vcycle")
        use MGPoisson2D, only: plain_vcycle => vcycle, fmg
        use DeferredMGPoisson2D, only: dc_vcycle => vcycle
        integer(c_int), value       :: kind, nx, ny, homog, maxiters, maxdepth, numcycles
        integer(c_int), value       :: sdown, sup, sbottom, zerophi
        real(c_double), value       :: dx, dy, tol
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny), rhs(0:nx-1,0:ny-1)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        if (kind .eq. 0) then
            call plain_vcycle (fphi, frhs, geo, bc, homog .ne. 0, 0, 0, tol, maxiters, maxdepth, &
                               numcycles, sdown, sup, sbottom, zerophi .ne. 0, 0)
        else if (kind .eq. 2) then
            call fmg (fphi, frhs, geo, bc, homog .ne. 0, 0, 0, tol, maxiters, maxdepth, &
                      numcycles, sdown, sup, sbottom, zerophi .ne. 0, 0)
        else
            call dc_vcycle (fphi, frhs, geo, bc, homog .ne. 0, 0, 0, tol, maxiters, maxdepth, &
                            numcycles, sdown, sup, sbottom, zerophi .ne. 0, 0)
        endif
        phi = fphi%data
        rhs = frhs%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_vcycle

    ! --------------------------------------------------------------------------
    ! Poisson2D::solve_bicgstab2 (poisson2d.f90:2253) when kind = 0,
    ! DeferredPoisson2D::solve_bicgstab2 (MGPoisson2D_DD.f90:339) when kind = 1.
    ! --------------------------------------------------------------------------
    subroutine ref_bicgstab2 (kind, nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                              bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, invd, tol, maxiters, &
                              maxrestarts, zerophi, rhs, phi) bind(C, name="ref_bicgstab2")
        use Poisson2D, only: plain_bicgstab2 => solve_bicgstab2
        use DeferredPoisson2D, only: dc_bicgstab2 => solve_bicgstab2
        integer(c_int), value       :: kind, nx, ny, homog, maxiters, maxrestarts, zerophi
        real(c_double), value       :: dx, dy, tol
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(in)  :: invd(0:nx-1,0:ny-1), rhs(0:nx-1,0:ny-1)
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs, finv

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        call mk_cc0 (finv, valid, invd)
        if (kind .eq. 0) then
            call plain_bicgstab2 (fphi, frhs, geo, bc, homog .ne. 0, finv, tol, maxiters, &
                                  maxrestarts, zerophi .ne. 0, 0)
        else
            call dc_bicgstab2 (fphi, frhs, geo, bc, homog .ne. 0, finv, tol, maxiters, &
                               maxrestarts, zerophi .ne. 0, 0)
        endif
        phi = fphi%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_box_data (finv)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_bicgstab2

    ! --------------------------------------------------------------------------
    ! Poisson2D::relax_jacobi (poisson2d.f90:1358) and Poisson2D::solve_bicgstab
    ! (poisson2d.f90:1993), verbosity 0: phi in, phi out.
    ! --------------------------------------------------------------------------
    subroutine ref_relax_jacobi (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                                 bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, invd, omega, tol, maxiters, &
                                 zerophi, rhs, phi) bind(C, name="ref_relax_jacobi")
        use Poisson2D, only: relax_jacobi
        integer(c_int), value       :: nx, ny, homog, maxiters, zerophi
        real(c_double), value       :: dx, dy, omega, tol
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(in)  :: invd(0:nx-1,0:ny-1), rhs(0:nx-1,0:ny-1)
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs, finv

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        call mk_cc0 (finv, valid, invd)
        call relax_jacobi (fphi, frhs, geo, bc, homog .ne. 0, finv, omega, tol, maxiters, zerophi .ne. 0, 0)
        phi = fphi%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_box_data (finv)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_relax_jacobi

    subroutine ref_bicgstab1 (nx, ny, dx, dy, J, Jxx, Jxy, Jyx, Jyy, bct, bcm, &
                              bd_xlo, bd_xhi, bd_ylo, bd_yhi, homog, invd, tol, maxiters, &
                              maxrestarts, zerophi, rhs, phi) bind(C, name="ref_bicgstab1")
        use Poisson2D, only: solve_bicgstab
        integer(c_int), value       :: nx, ny, homog, maxiters, maxrestarts, zerophi
        real(c_double), value       :: dx, dy, tol
        real(c_double), intent(in)  :: J(0:nx-1,0:ny-1), Jxx(0:nx,0:ny-1), Jxy(0:nx,0:ny-1)
        real(c_double), intent(in)  :: Jyx(0:nx-1,0:ny), Jyy(0:nx-1,0:ny)
        integer(c_int), intent(in)  :: bct(4), bcm(4)
        type(c_ptr), value          :: bd_xlo, bd_xhi, bd_ylo, bd_yhi
        real(c_double), intent(in)  :: invd(0:nx-1,0:ny-1), rhs(0:nx-1,0:ny-1)
        real(c_double), intent(inout) :: phi(-1:nx,-1:ny)

        type(box) :: valid
        type(geo_data) :: geo
        type(bdry_data) :: bc
        type(box_data) :: fphi, frhs, finv

        call mk_valid (valid, nx, ny, dx, dy)
        call mk_geo (geo, valid, J, Jxx, Jxy, Jyx, Jyy)
        call mk_bc (bc, valid, bct, bcm, bd_xlo, bd_xhi, bd_ylo, bd_yhi)
        call mk_cc1 (fphi, valid, phi)
        call mk_cc0 (frhs, valid, rhs)
        call mk_cc0 (finv, valid, invd)
        call solve_bicgstab (fphi, frhs, geo, bc, homog .ne. 0, finv, tol, maxiters, maxrestarts, &
                             zerophi .ne. 0, 0)
        phi = fphi%data
        call undefine_box_data (fphi)
        call undefine_box_data (frhs)
        call undefine_box_data (finv)
        call undefine_bdry_data (bc)
        call free_geo (geo)
    end subroutine ref_bicgstab1

    ! --------------------------------------------------------------------------
    ! ArrayUtils::pnorm (arrayutils.f90:753) over the valid box of a ghost-free
    ! cell-centred field.
    ! --------------------------------------------------------------------------
    function ref_pnorm (nx, ny, dx, dy, a, p) result (pn) bind(C, name="ref_pnorm")
        integer(c_int), value      :: nx, ny, p
        real(c_double), value      :: dx, dy
        real(c_double), intent(in) :: a(0:nx-1,0:ny-1)
        real(c_double)             :: pn
        type(box) :: valid
        type(box_data) :: f
        call mk_valid (valid, nx, ny, dx, dy)
        call mk_cc0 (f, valid, a)
        pn = pnorm (f, f%valid, p)
        call undefine_box_data (f)
    end function ref_pnorm

    ! --------------------------------------------------------------------------
    ! ArrayUtils::inner_prod (arrayutils.f90:637) and integrate2d (:670) over
    ! ghost-free cell-centred fields; jscale 0 passes opt_jscale = .false.
    ! --------------------------------------------------------------------------
    function ref_inner_prod (nx, ny, dx, dy, a, b) result (ip) bind(C, name="ref_inner_prod")
        integer(c_int), value      :: nx, ny
        real(c_double), value      :: dx, dy
        real(c_double), intent(in) :: a(0:nx-1,0:ny-1), b(0:nx-1,0:ny-1)
        real(c_double)             :: ip
        type(box) :: valid
        type(box_data) :: fa, fb
        call mk_valid (valid, nx, ny, dx, dy)
        call mk_cc0 (fa, valid, a)
        call mk_cc0 (fb, valid, b)
        ip = inner_prod (fa, fb)
        call undefine_box_data (fa)
        call undefine_box_data (fb)
    end function ref_inner_prod

    function ref_integrate2d (nx, ny, dx, dy, a, J, jscale) result (sm) bind(C, name="ref_integrate2d")
        integer(c_int), value      :: nx, ny, jscale
        real(c_double), value      :: dx, dy
        real(c_double), intent(in) :: a(0:nx-1,0:ny-1), J(0:nx-1,0:ny-1)
        real(c_double)             :: sm
        type(box) :: valid
        type(box_data) :: fa
        type(geo_data) :: geo
        call mk_valid (valid, nx, ny, dx, dy)
        call mk_cc0 (fa, valid, a)
        call mk_cc0 (geo%J, valid, J)
        geo%dx = dx
        geo%dy = dy
        sm = integrate2d (fa, valid, geo, jscale .ne. 0)
        call undefine_box_data (fa)
        call undefine_box_data (geo%J)
    end function ref_integrate2d

    ! --------------------------------------------------------------------------
    ! ArrayUtils::integrate2d_bdry (arrayutils.f90:721) and gpnorm (:802) on the
    ! box 0..nx-1 x 0..ny-1 (fluxes without ghosts, phi with one ghost layer).
    ! --------------------------------------------------------------------------
    function ref_integrate2d_bdry (nx, ny, dx, dy, xf, yf) result (sm) bind(C, name="ref_integrate2d_bdry")
        integer(c_int), value      :: nx, ny
        real(c_double), value      :: dx, dy
        real(c_double), intent(in) :: xf(0:nx,0:ny-1), yf(0:nx-1,0:ny)
        real(c_double)             :: sm
        type(box) :: valid
        type(box_data) :: xflux, yflux
        call mk_valid (valid, nx, ny, dx, dy)
        call define_box_data (xflux, valid, 0, 0, BD_NODE, BD_CELL)
        call define_box_data (yflux, valid, 0, 0, BD_CELL, BD_NODE)
        xflux%data = xf
        yflux%data = yf
        sm = integrate2d_bdry (xflux, yflux, valid)
        call undefine_box_data (xflux)
        call undefine_box_data (yflux)
    end function ref_integrate2d_bdry

    function ref_gpnorm (nx, ny, dx, dy, a, p) result (pn) bind(C, name="ref_gpnorm")
        integer(c_int), value      :: nx, ny, p
        real(c_double), value      :: dx, dy
        real(c_double), intent(in) :: a(-1:nx, -1:ny)
        real(c_double)             :: pn
        type(box) :: valid
        type(box_data) :: f
        call mk_valid (valid, nx, ny, dx, dy)
        call mk_cc1 (f, valid, a)
        pn = gpnorm (f, p)
        call undefine_box_data (f)
    end function ref_gpnorm

end module ref_bridge
