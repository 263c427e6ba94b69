! A caller written against the reference's libMGPoisson interface (module
! names, derived types and argument lists of EdSantilli/MGPoisson2D).  The
! same source is compiled once against the reference's modules and once
! against the drop-in layer (mgpoisson2d_amd/fortran); the parity test
! compares the two outputs.
!
!   dropin_caller <output file>
program dropin_caller
    use ArrayUtils
    use Poisson2D
    use MGPoisson2D, only: restrict, plain_vcycle => vcycle, fmg
    use DeferredMGPoisson2D, only: dc_vcycle => vcycle
    implicit none

    integer, parameter :: n = 64
    type(box)          :: valid, cvalid
    type(geo_data)     :: geo
    type(bdry_data)    :: bc
    type(box_data)     :: phi, rhs, lap, invd, cj, phi0
    type(box_data)     :: xflux, yflux, xwk, ywk
    integer            :: i, j, u
    character(len=256) :: fname
    real(dp)           :: x, y

    call get_command_argument (1, fname)
    call define_box (valid, 0, n-1, 0, n-1, one/n, one/n)
    call define_geometry (geo, valid)

    call define_bdry_data (bc, valid, BCTYPE_DIRI, BCTYPE_DIRI, BCTYPE_NEUM, BCTYPE_NEUM, &
                           BCMODE_UNIFORM, BCMODE_UNIFORM, BCMODE_UNIFORM, BCMODE_UNIFORM)
    bc%data_xlo = zero
    bc%data_xhi = zero
    bc%data_ylo = zero
    bc%data_yhi = zero

    call define_box_data (phi, valid, 1, 1, BD_CELL, BD_CELL)
    call define_box_data (phi0, valid, 1, 1, BD_CELL, BD_CELL)
    call define_box_data (rhs, valid, 0, 0, BD_CELL, BD_CELL)
    call define_box_data (lap, valid, 0, 0, BD_CELL, BD_CELL)
    call define_box_data (invd, valid, 0, 0, BD_CELL, BD_CELL)
    phi%data = zero
    do j = valid%jlo, valid%jhi
        do i = valid%ilo, valid%ihi
            x = (i + half) / n
            y = (j + half) / n
            phi%data(i,j) = sin(pi*x) * cos(two*pi*y) + 0.1_dp * x * y
            rhs%data(i,j) = cos(three*x) * sin(pi*y) * geo%J%data(i,j)
        enddo
    enddo
    phi0%data = phi%data

    open (newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')

    call compute_laplacian (lap, phi, geo, bc, .true.)
    write (u) lap%data
    call compute_invdiags (invd, geo, bc, .true.)
    write (u) invd%data
    call relax_gs (phi, rhs, geo, bc, .true., invd, one, -one, 3, .false., 0)
    write (u) phi%data(valid%ilo:valid%ihi, valid%jlo:valid%jhi)

    cvalid = valid
    call coarsen_box (cvalid, 2, 2)
    call define_box_data (cj, cvalid, 0, 0, BD_CELL, BD_CELL)
    call restrict (geo%J, cj)
    write (u) cj%data

    phi%data = zero
    call dc_vcycle (phi, rhs, geo, bc, .true., 0, 0, 1.0d-30, 4, -1, 1, 2, 2, 2, .true., 0)
    write (u) phi%data(valid%ilo:valid%ihi, valid%jlo:valid%jhi)
    write (u) rhs%data

    phi%data = phi0%data
    call plain_vcycle (phi, rhs, geo, bc, .true., 0, 0, 1.0d-30, 1, -1, 1, 2, 2, 2, .false., 0)
    write (u) phi%data(valid%ilo:valid%ihi, valid%jlo:valid%jhi)

    ! one full-multigrid cycle, kept (a tolerance it always meets)
    phi%data = phi0%data
    call fmg (phi, rhs, geo, bc, .true., 0, 0, 1.0d30, 1, -1, 1, 2, 2, 2, .false., 0)
    write (u) phi%data(valid%ilo:valid%ihi, valid%jlo:valid%jhi)

    ! single-level solvers: 12 Jacobi sweeps, then solve_bicgstab with restarts
    phi%data = phi0%data
    call relax_jacobi (phi, rhs, geo, bc, .true., invd, 0.8_dp, -one, 12, .false., 0)
    write (u) phi%data(valid%ilo:valid%ihi, valid%jlo:valid%jhi)
    phi%data = phi0%data
    call solve_bicgstab (phi, rhs, geo, bc, .true., invd, 1.0d-8, 20, 2, .true., 0)
    write (u) phi%data(valid%ilo:valid%ihi, valid%jlo:valid%jhi)

    ! projection tools with inhomogeneous data: J grad phi, J div of it, one
    ! transverse derivative, then the homogeneous boundary fluxes
    bc%data_xlo = 0.3_dp
    bc%data_ylo = 0.25_dp
    bc%data_yhi = -0.5_dp
    call define_box_data (xflux, valid, 0, 0, BD_NODE, BD_CELL)
    call define_box_data (yflux, valid, 0, 0, BD_CELL, BD_NODE)
    call define_box_data (xwk, xflux)
    call define_box_data (ywk, yflux)
    phi%data = phi0%data
    call compute_grad (xflux, yflux, phi, geo, bc, .false., xwk, ywk)
    write (u) xflux%data
    write (u) yflux%data
    call compute_div (lap, xflux, yflux)
    write (u) lap%data
    call compute_pd (ywk, phi, 1)
    write (u) ywk%data
    call fill_boundary_fluxes (xflux, yflux, bc, .true.)
    write (u) yflux%data

    ! reductions and ghost extrapolation
    write (u) pnorm (rhs, valid, 0), pnorm (rhs, valid, 1), pnorm (rhs, valid, 2), pnorm (rhs, valid, 3)
    write (u) inner_prod (rhs, lap), integrate2d (rhs, valid, geo), integrate2d (rhs, valid, geo, .false.)
    phi%data = phi0%data
    call extrapolate_ghosts (phi, 2)
    write (u) phi%data
    close (u)

contains

    ! x = xi + 0.08 sin(2 pi xi) sin(2 pi eta), y = (exp(2 eta) - 1)/(exp(2) - 1):
    ! analytic J and J g^ij at cell centres and faces, one ghost layer.
    subroutine define_geometry (g, vb)
        type(geo_data), intent(out) :: g
        type(box), intent(in)       :: vb
        call define_box_data (g%J, vb, 1, 1, BD_CELL, BD_CELL)
        call define_box_data (g%Jgup_xx, vb, 1, 1, BD_NODE, BD_CELL)
        call define_box_data (g%Jgup_xy, vb, 1, 1, BD_NODE, BD_CELL)
        call define_box_data (g%Jgup_yx, vb, 1, 1, BD_CELL, BD_NODE)
        call define_box_data (g%Jgup_yy, vb, 1, 1, BD_CELL, BD_NODE)
        call fill (g%J, 0)
        call fill (g%Jgup_xx, 1)
        call fill (g%Jgup_xy, 2)
        call fill (g%Jgup_yx, 2)
        call fill (g%Jgup_yy, 3)
        g%dx = vb%dx
        g%dy = vb%dy
    end subroutine define_geometry

    subroutine fill (f, what)
        type(box_data), intent(inout) :: f
        integer, intent(in)           :: what
        real(dp), parameter :: amp = 0.08_dp, beta = two
        real(dp) :: xi, eta, xx, xe, yx, ye, jac
        integer  :: i, j
        do j = f%bx%jlo, f%bx%jhi
            do i = f%bx%ilo, f%bx%ihi
                xi = (i + f%offx) * f%valid%dx
                eta = (j + f%offy) * f%valid%dy
                xx = one + amp * two * pi * cos(two*pi*xi) * sin(two*pi*eta)
                xe = amp * two * pi * sin(two*pi*xi) * cos(two*pi*eta)
                yx = zero
                ye = beta * exp(beta*eta) / (exp(beta) - one)
                jac = xx * ye - xe * yx
                select case (what)
                    case (0)
                        f%data(i,j) = jac
                    case (1)
                        f%data(i,j) = (xe*xe + ye*ye) / jac
                    case (2)
                        f%data(i,j) = -(xx*xe + yx*ye) / jac
                    case default
                        f%data(i,j) = (xx*xx + yx*yx) / jac
                end select
            enddo
        enddo
    end subroutine fill

end program dropin_caller
