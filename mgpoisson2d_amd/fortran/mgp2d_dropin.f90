! ------------------------------------------------------------------------------
! Fortran 90 drop-in layer for libMGPoisson (EdSantilli/MGPoisson2D).
!
! Modules precision, ArrayUtils, Poisson2D, MGPoisson2D, DeferredPoisson2D and
! DeferredMGPoisson2D keep the reference's names, derived types (component
! names, kinds, allocatable ranks, bounds) and procedure signatures, so a
! caller written against the reference recompiles unchanged against this
! layer.  The hot-path procedures marshal their arguments through
! iso_c_binding into the C ABI of include/mgp2d.h, which runs them on the
! MI355X; grid bookkeeping (define_*, setval_*, norms of host arrays) stays in
! Fortran.  Errors are reported the reference's way: print* then stop.
! ------------------------------------------------------------------------------

! precision.f90:4-53 -- kinds and named constants
module precision
    integer, parameter  :: dp = kind(0.d0)
    real(dp), parameter :: zero = 0.0_dp, one = 1.0_dp, two = 2.0_dp, three = 3.0_dp
    real(dp), parameter :: four = 4.0_dp, five = 5.0_dp, six = 6.0_dp, seven = 7.0_dp
    real(dp), parameter :: eight = 8.0_dp, nine = 9.0_dp, ten = 10.0_dp, eleven = 11.0_dp
    real(dp), parameter :: twelve = 12.0_dp, thirteen = 13.0_dp, fourteen = 14.0_dp
    real(dp), parameter :: fifteen = 15.0_dp, sixteen = 16.0_dp, seventeen = 17.0_dp
    real(dp), parameter :: eighteen = 18.0_dp, nineteen = 19.0_dp, twenty = 20.0_dp
    real(dp), parameter :: half = one / two, third = one / three, fourth = one / four
    real(dp), parameter :: fifth = one / five, sixth = one / six, seventh = one / seven
    real(dp), parameter :: eighth = one / eight, ninth = one / nine, tenth = one / ten
    real(dp), parameter :: eleventh = one / eleven, twelfth = one / twelve
    real(dp), parameter :: thirteenth = one / thirteen, fourteenth = one / fourteen
    real(dp), parameter :: fifteenth = one / fifteen, sixteenth = one / sixteen
    real(dp), parameter :: seventeenth = one / seventeen, eighteenth = one / eighteen
    real(dp), parameter :: nineteenth = one / nineteen, twentieth = one / twenty
    real(dp), parameter :: threehalves = three / two, threefourths = three / four
    real(dp), parameter :: pi = four * atan(one)
    real(dp), parameter :: bogus_val = 1.2345E300_dp
end module precision


! arrayutils.f90 -- grid containers and helpers
module ArrayUtils
    use precision
    use iso_c_binding
    implicit none

    type box
        integer  :: ilo, ihi
        integer  :: jlo, jhi
        integer  :: nx, ny
        real(dp) :: dx, dy
    end type box

    type(box), parameter :: empty_box = box(0, -1, 0, -1, 0, 0, zero, zero)

    integer, parameter :: BCTYPE_UNDEFINED = -2
    integer, parameter :: BCTYPE_NONE = -1
    integer, parameter :: BCTYPE_NEUM = 0
    integer, parameter :: BCTYPE_DIRI = 1
    integer, parameter :: BCTYPE_PERIODIC = 2
    integer, parameter :: BCTYPE_CF = 3
    integer, parameter :: BCTYPE_EXTRAP0 = 4
    integer, parameter :: BCTYPE_EXTRAP1 = 5
    integer, parameter :: BCTYPE_EXTRAP2 = 6

    integer, parameter :: BCMODE_UNIFORM = 0
    integer, parameter :: BCMODE_NONUNIFORM = 1
    integer, parameter :: BCMODE_LAH = 2

    ! Smoother orderings of the GPU layer (optional trailing argument of
    ! relax_gs / vcycle; ORDER_LEX reproduces the reference bit for bit).
    integer, parameter :: ORDER_LEX = 0
    integer, parameter :: ORDER_RB = 1
    integer, parameter :: ORDER_MC = 2

    type bdry_data
        type(box) :: valid
        integer :: type_xlo, type_xhi
        integer :: type_ylo, type_yhi
        integer :: mode_xlo, mode_xhi
        integer :: mode_ylo, mode_yhi
        real(dp), dimension(:), allocatable :: data_xlo, data_xhi
        real(dp), dimension(:), allocatable :: data_ylo, data_yhi
    end type bdry_data

    integer, parameter :: BD_NODE = 0
    integer, parameter :: BD_CELL = 1

    type box_data
        type(box) :: bx
        type(box) :: valid
        integer   :: ngx, ngy
        integer   :: offi, offj
        real(dp)  :: offx, offy
        real(dp)  :: L, H
        real(dp), dimension(:,:), allocatable :: data
    end type box_data

    integer, parameter :: SIDE_LO = 0
    integer, parameter :: SIDE_HI = 1

    type geo_data
        type(box_data) :: J
        type(box_data) :: Jgup_xx
        type(box_data) :: Jgup_xy
        type(box_data) :: Jgup_yx
        type(box_data) :: Jgup_yy
        real(dp)       :: dx, dy
    end type geo_data

    interface define_box_data
        module procedure define_box_data_itemize
        module procedure define_box_data_dup
    end interface

    ! C ABI (include/mgp2d.h) ---------------------------------------------------
    type, bind(C) :: mgp_geo_c
        integer(c_int) :: nx, ny
        real(c_double) :: dx, dy
        type(c_ptr)    :: J, Jxx, Jxy, Jyx, Jyy
    end type mgp_geo_c

    type, bind(C) :: mgp_bc_c
        integer(c_int) :: btype(4)
        integer(c_int) :: bmode(4)
        type(c_ptr)    :: bdata(4)
    end type mgp_bc_c

    ! Contiguous copies of a geo_data's valid faces (kept alive for a call).
    type geo_pack
        real(dp), allocatable :: J(:,:), Jxx(:,:), Jxy(:,:), Jyx(:,:), Jyy(:,:)
    end type geo_pack

    type bc_pack
        real(dp), allocatable :: d1(:), d2(:), d3(:), d4(:)
    end type bc_pack

    interface
        integer(c_int) function mgp_fill_ghosts_c(geo, bc, homog, do_neum, phi) bind(C, name="mgp_fill_ghosts")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog, do_neum
            real(c_double)              :: phi(*)
        end function
        integer(c_int) function mgp_pnorm_c(nx, ny, dx, dy, a, p, out) bind(C, name="mgp_pnorm")
            import :: c_int, c_double
            integer(c_int), value      :: nx, ny, p
            real(c_double), value      :: dx, dy
            real(c_double), intent(in) :: a(*)
            real(c_double)             :: out
        end function
        integer(c_int) function mgp_inner_prod_c(nx, ny, dx, dy, a, b, out) bind(C, name="mgp_inner_prod")
            import :: c_int, c_double
            integer(c_int), value      :: nx, ny
            real(c_double), value      :: dx, dy
            real(c_double), intent(in) :: a(*), b(*)
            real(c_double)             :: out
        end function
        integer(c_int) function mgp_integrate2d_c(nx, ny, dx, dy, a, J, out) bind(C, name="mgp_integrate2d")
            import :: c_int, c_double, c_ptr
            integer(c_int), value      :: nx, ny
            real(c_double), value      :: dx, dy
            real(c_double), intent(in) :: a(*)
            type(c_ptr), value         :: J
            real(c_double)             :: out
        end function
        integer(c_int) function mgp_extrapolate_ghosts_c(nx, ny, order, phi) bind(C, name="mgp_extrapolate_ghosts")
            import :: c_int, c_double
            integer(c_int), value      :: nx, ny, order
            real(c_double)             :: phi(*)
        end function
        integer(c_int) function mgp_fill_boundary_fluxes_c(nx, ny, bc, homog, xflux, yflux) &
                                                           bind(C, name="mgp_fill_boundary_fluxes")
            import :: mgp_bc_c, c_int, c_double
            integer(c_int), value      :: nx, ny, homog
            type(mgp_bc_c), intent(in) :: bc
            real(c_double)             :: xflux(*), yflux(*)
        end function
        integer(c_int) function mgp_integrate2d_bdry_c(nx, ny, dx, dy, xflux, yflux, out) &
                                                       bind(C, name="mgp_integrate2d_bdry")
            import :: c_int, c_double
            integer(c_int), value      :: nx, ny
            real(c_double), value      :: dx, dy
            real(c_double), intent(in) :: xflux(*), yflux(*)
            real(c_double)             :: out
        end function
        integer(c_int) function mgp_gpnorm_c(ilo, ihi, jlo, jhi, dx, dy, phi, p, out) bind(C, name="mgp_gpnorm")
            import :: c_int, c_double
            integer(c_int), value      :: ilo, ihi, jlo, jhi, p
            real(c_double), value      :: dx, dy
            real(c_double), intent(in) :: phi(*)
            real(c_double)             :: out
        end function
        type(c_ptr) function mgp_last_error_c() bind(C, name="mgp_last_error")
            import :: c_ptr
        end function
    end interface

contains

    ! Reports a failed C-ABI call the way the reference reports errors.
    subroutine mgp_check (status, where)
        integer(c_int), intent(in)   :: status
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=256) :: text
        integer :: k
        if (status .eq. 0) return
        text = ''
        call c_f_pointer (mgp_last_error_c(), msg, [256])
        do k = 1, 256
            if (msg(k) .eq. c_null_char) exit
            text(k:k) = msg(k)
        enddo
        print*, where, ': ', trim(text)
        stop
    end subroutine mgp_check

    pure subroutine define_box (bx, ilo, ihi, jlo, jhi, dx, dy)
        type(box), intent(out) :: bx
        integer, intent(in)    :: ilo, ihi, jlo, jhi
        real(dp), intent(in)   :: dx, dy
        bx%ilo = ilo
        bx%ihi = ihi
        bx%jlo = jlo
        bx%jhi = jhi
        bx%nx = ihi - ilo + 1
        bx%ny = jhi - jlo + 1
        bx%dx = dx
        bx%dy = dy
    end subroutine define_box

    pure subroutine coarsen_box (bx, refx, refy)
        type(box), intent(inout) :: bx
        integer, intent(in)      :: refx, refy
        bx%ilo = bx%ilo / refx
        bx%ihi = bx%ihi / refx
        bx%nx = bx%ihi - bx%ilo + 1
        bx%dx = bx%dx * refx
        bx%jlo = bx%jlo / refy
        bx%jhi = bx%jhi / refy
        bx%ny = bx%jhi - bx%jlo + 1
        bx%dy = bx%dy * refy
    end subroutine coarsen_box

    pure subroutine grow_box (bx, growx, growy)
        type(box), intent(inout) :: bx
        integer, intent(in)      :: growx, growy
        bx%ilo = bx%ilo - growx
        bx%ihi = bx%ihi + growx
        bx%jlo = bx%jlo - growy
        bx%jhi = bx%jhi + growy
    end subroutine grow_box

    function bdry_box (src, off, dir, side) result (bx)
        type(box), intent(in) :: src
        integer, intent(in)   :: off, dir, side
        type(box)             :: bx
        if (side .ne. SIDE_LO .and. side .ne. SIDE_HI) then
            print*, 'bdry_box: Bad side.'
            stop
        endif
        if (dir .eq. 1) then
            if (side .eq. SIDE_LO) then
                call define_box (bx, src%ilo, src%ilo, src%jlo, src%jhi, src%dx, src%dy)
            else
                call define_box (bx, src%ihi+off, src%ihi+off, src%jlo, src%jhi, src%dx, src%dy)
            endif
        else if (dir .eq. 2) then
            if (side .eq. SIDE_LO) then
                call define_box (bx, src%ilo, src%ihi, src%jlo, src%jlo, src%dx, src%dy)
            else
                call define_box (bx, src%ilo, src%ihi, src%jhi+off, src%jhi+off, src%dx, src%dy)
            endif
        else
            print*, 'bdry_box: Bad dir.'
            stop
        endif
    end function bdry_box

    subroutine define_bdry_data (bcd, valid, type_xlo, type_xhi, type_ylo, type_yhi, &
                                 mode_xlo, mode_xhi, mode_ylo, mode_yhi)
        type(bdry_data), intent(out) :: bcd
        type(box), intent(in)        :: valid
        integer, intent(in)          :: type_xlo, type_xhi, type_ylo, type_yhi
        integer, intent(in)          :: mode_xlo, mode_xhi, mode_ylo, mode_yhi
        bcd%valid = valid
        bcd%type_xlo = type_xlo
        bcd%type_xhi = type_xhi
        bcd%type_ylo = type_ylo
        bcd%type_yhi = type_yhi
        bcd%mode_xlo = mode_xlo
        bcd%mode_xhi = mode_xhi
        bcd%mode_ylo = mode_ylo
        bcd%mode_yhi = mode_yhi
        call alloc_side (bcd%data_xlo, mode_xlo, valid%jlo, valid%jhi)
        call alloc_side (bcd%data_xhi, mode_xhi, valid%jlo, valid%jhi)
        call alloc_side (bcd%data_ylo, mode_ylo, valid%ilo, valid%ihi)
        call alloc_side (bcd%data_yhi, mode_yhi, valid%ilo, valid%ihi)
    end subroutine define_bdry_data

    subroutine alloc_side (d, mode, lo, hi)
        real(dp), allocatable, intent(inout) :: d(:)
        integer, intent(in) :: mode, lo, hi
        integer :: ierr
        select case (mode)
            case (BCMODE_UNIFORM, BCMODE_LAH)
                allocate (d(1:1), stat=ierr)
            case (BCMODE_NONUNIFORM)
                allocate (d(lo:hi), stat=ierr)
            case default
                print*, 'define_bdry_data: Bad BCMODE'
                stop
        end select
        if (ierr .ne. 0) then
            print*, 'define_bdry_data: Out of memory'
            stop
        endif
    end subroutine alloc_side

    pure subroutine undefine_bdry_data (bcd)
        type(bdry_data), intent(inout) :: bcd
        bcd%valid = empty_box
        bcd%type_xlo = BCTYPE_UNDEFINED
        bcd%type_xhi = BCTYPE_UNDEFINED
        bcd%type_ylo = BCTYPE_UNDEFINED
        bcd%type_yhi = BCTYPE_UNDEFINED
        if (allocated(bcd%data_xlo)) deallocate(bcd%data_xlo)
        if (allocated(bcd%data_xhi)) deallocate(bcd%data_xhi)
        if (allocated(bcd%data_ylo)) deallocate(bcd%data_ylo)
        if (allocated(bcd%data_yhi)) deallocate(bcd%data_yhi)
    end subroutine undefine_bdry_data

    subroutine define_box_data_itemize (bd, cc_valid, ngx, ngy, offi, offj)
        type(box_data), intent(out) :: bd
        type(box), intent(in)       :: cc_valid
        integer, intent(in)         :: ngx, ngy, offi, offj
        type(box) :: bx, valid
        integer   :: ierr
        valid = cc_valid
        bd%offx = half
        bd%offy = half
        if (offi .eq. BD_NODE) then
            valid%ihi = valid%ihi + 1
            valid%nx = valid%nx + 1
            bd%offx = zero
        else if (offi .ne. BD_CELL) then
            print*, 'define_box_data: Bad offi'
        endif
        if (offj .eq. BD_NODE) then
            valid%jhi = valid%jhi + 1
            valid%ny = valid%ny + 1
            bd%offy = zero
        else if (offj .ne. BD_CELL) then
            print*, 'define_box_data: Bad offj'
        endif
        bx = valid
        bx%ilo = bx%ilo - ngx
        bx%ihi = bx%ihi + ngx
        bx%jlo = bx%jlo - ngy
        bx%jhi = bx%jhi + ngy
        bd%valid = valid
        bd%bx = bx
        bd%ngx = ngx
        bd%ngy = ngy
        bd%offi = offi
        bd%offj = offj
        bd%L = cc_valid%nx * cc_valid%dx
        bd%H = cc_valid%ny * cc_valid%dy
        allocate (bd%data(bx%ilo:bx%ihi, bx%jlo:bx%jhi), stat=ierr)
        if (ierr .ne. 0) then
            print*, 'define_box_data: Out of memory'
            stop
        endif
        bd%data = zero
    end subroutine define_box_data_itemize

    subroutine define_box_data_dup (dest, src)
        type(box_data), intent(inout) :: dest
        type(box_data), intent(in)    :: src
        type(box) :: cc_valid
        cc_valid = src%valid
        if (src%offi .eq. BD_NODE) then
            cc_valid%ihi = cc_valid%ihi - 1
            cc_valid%nx = cc_valid%nx - 1
        endif
        if (src%offj .eq. BD_NODE) then
            cc_valid%jhi = cc_valid%jhi - 1
            cc_valid%ny = cc_valid%ny - 1
        endif
        call define_box_data_itemize (dest, cc_valid, src%ngx, src%ngy, src%offi, src%offj)
    end subroutine define_box_data_dup

    pure subroutine undefine_box_data (bd)
        type(box_data), intent(inout) :: bd
        bd%valid = empty_box
        bd%bx = empty_box
        bd%ngx = 0
        bd%ngy = 0
        bd%offx = bogus_val
        bd%offy = bogus_val
        bd%offi = 0
        bd%offj = 0
        bd%L = bogus_val
        bd%H = bogus_val
        if (allocated(bd%data)) deallocate(bd%data)
    end subroutine undefine_box_data

    pure function stagger (min) result (mout)
        integer, intent(in) :: min
        integer             :: mout
        mout = 1 - min
    end function stagger

    pure subroutine setval_valid (bd, val)
        type(box_data), intent(inout) :: bd
        real(dp), intent(in)          :: val
        bd%data(bd%valid%ilo:bd%valid%ihi, bd%valid%jlo:bd%valid%jhi) = val
    end subroutine setval_valid

    ! arrayutils.f90:481 -- bd on the boundary faces of src (dir 1: x faces,
    ! dir 2: y faces), no ghosts
    subroutine define_box_data_bdry (bd, src, dir, side)
        type(box_data), intent(inout) :: bd
        type(box_data), intent(in)    :: src
        integer, intent(in)           :: dir, side
        integer :: ierr
        if (dir .eq. 1) then
            bd%valid = bdry_box (src%valid, src%offi, dir, side)
            bd%offi = BD_NODE
            bd%offx = zero
            bd%offj = src%offj
            bd%offy = src%offy
        else if (dir .eq. 2) then
            bd%valid = bdry_box (src%valid, src%offj, dir, side)
            bd%offi = src%offi
            bd%offx = src%offx
            bd%offj = BD_NODE
            bd%offy = zero
        else
            print*, 'define_box_data_bdry: Bad dir'
        endif
        bd%bx = bd%valid
        bd%ngx = 0
        bd%ngy = 0
        bd%L = src%L
        bd%H = src%H
        if (allocated(bd%data)) deallocate (bd%data)
        allocate (bd%data(bd%bx%ilo:bd%bx%ihi, bd%bx%jlo:bd%bx%jhi), stat=ierr)
        if (ierr .ne. 0) then
            print*, 'define_box_data: Out of memory'
            stop
        endif
        bd%data = zero
    end subroutine define_box_data_bdry

    ! arrayutils.f90:530 -- crse over fine's valid box divided by the ratios.
    ! The reference sets only the index bounds of the coarse box (its nx, ny,
    ! dx and dy stay undefined); here they follow from the indices and ratios.
    subroutine define_box_data_coarsened (crse, fine, refx, refy)
        type(box_data), intent(out) :: crse
        type(box_data), intent(in)  :: fine
        integer, intent(in)         :: refx, refy
        type(box) :: cv
        cv%ilo = fine%valid%ilo / refx
        cv%ihi = fine%valid%ihi / refx
        cv%jlo = fine%valid%jlo / refy
        cv%jhi = fine%valid%jhi / refy
        cv%nx = cv%ihi - cv%ilo + 1
        cv%ny = cv%jhi - cv%jlo + 1
        cv%dx = fine%valid%dx * refx
        cv%dy = fine%valid%dy * refy
        call define_box_data_itemize (crse, cv, fine%ngx, fine%ngy, fine%offi, fine%offj)
    end subroutine define_box_data_coarsened

    ! arrayutils.f90:583
    pure function compatible_boxes (bx1, bx2) result (is_same)
        type(box), intent(in) :: bx1, bx2
        logical :: is_same
        is_same = bx1%ilo .eq. bx2%ilo .and. bx1%ihi .eq. bx2%ihi .and. &
                  bx1%jlo .eq. bx2%jlo .and. bx1%jhi .eq. bx2%jhi .and. &
                  bx1%dx .eq. bx2%dx .and. bx1%dy .eq. bx2%dy
    end function compatible_boxes

    ! arrayutils.f90:615 -- every ghost (corners included) to val
    pure subroutine setval_ghosts (bd, val)
        type(box_data), intent(inout) :: bd
        real(dp), intent(in)          :: val
        if (bd%ngx .ge. 1) then
            bd%data(bd%bx%ilo:bd%valid%ilo-1, bd%bx%jlo:bd%bx%jhi) = val
            bd%data(bd%valid%ihi+1:bd%bx%ihi, bd%bx%jlo:bd%bx%jhi) = val
        endif
        if (bd%ngy .ge. 1) then
            bd%data(bd%bx%ilo:bd%bx%ihi, bd%bx%jlo:bd%valid%jlo-1) = val
            bd%data(bd%bx%ilo:bd%bx%ihi, bd%valid%jhi+1:bd%bx%jhi) = val
        endif
    end subroutine setval_ghosts

    ! arrayutils.f90:721 (runs on the GPU): net outward flux over bx
    function integrate2d_bdry (xflux, yflux, bx) result (sum)
        type(box_data), intent(in) :: xflux, yflux
        type(box), intent(in)      :: bx
        real(dp) :: sum
        real(dp), allocatable :: xf(:,:), yf(:,:)
        real(c_double) :: out
        allocate (xf(bx%ilo:bx%ihi+1, bx%jlo:bx%jhi), yf(bx%ilo:bx%ihi, bx%jlo:bx%jhi+1))
        xf = xflux%data(bx%ilo:bx%ihi+1, bx%jlo:bx%jhi)
        yf = yflux%data(bx%ilo:bx%ihi, bx%jlo:bx%jhi+1)
        call mgp_check (mgp_integrate2d_bdry_c(bx%nx, bx%ny, bx%dx, bx%dy, xf, yf, out), 'integrate2d_bdry')
        sum = out
    end function integrate2d_bdry

    ! arrayutils.f90:802 (runs on the GPU): norm over the ghost layer, with the
    ! reference's index pattern (see mgp_gpnorm)
    function gpnorm (bd, p) result (pn)
        type(box_data), intent(in) :: bd
        integer, intent(in)        :: p
        real(dp) :: pn
        real(dp), allocatable :: a(:,:)
        real(c_double) :: out
        type(box) :: v
        if ((bd%offi .ne. BD_CELL) .or. (bd%offj .ne. BD_CELL)) then
            print*, 'gpnorm: Only works with cell-centered data for now.'
            stop
        endif
        v = bd%valid
        allocate (a(v%ilo-1:v%ihi+1, v%jlo-1:v%jhi+1))
        a = bd%data(v%ilo-1:v%ihi+1, v%jlo-1:v%jhi+1)
        call mgp_check (mgp_gpnorm_c(v%ilo, v%ihi, v%jlo, v%jhi, v%dx, v%dy, a, p, out), 'gpnorm')
        pn = out
    end function gpnorm

    ! arrayutils.f90:637 (runs on the GPU)
    function inner_prod (bd1, bd2) result (ip)
        type(box_data), intent(in) :: bd1, bd2
        real(dp) :: ip
        real(dp), allocatable :: a(:,:), b(:,:)
        real(c_double) :: out
        integer  :: ilo, ihi, jlo, jhi
        ilo = max(bd1%valid%ilo, bd2%valid%ilo)
        ihi = min(bd1%valid%ihi, bd2%valid%ihi)
        jlo = max(bd1%valid%jlo, bd2%valid%jlo)
        jhi = min(bd1%valid%jhi, bd2%valid%jhi)
        if ((ilo .gt. ihi) .or. (jlo .gt. jhi)) then
            print*, 'ERROR: inner_prod_box_data: given box_data do not overlap.'
            stop
        endif
        allocate (a(ilo:ihi, jlo:jhi), b(ilo:ihi, jlo:jhi))
        a = bd1%data(ilo:ihi, jlo:jhi)
        b = bd2%data(ilo:ihi, jlo:jhi)
        call mgp_check (mgp_inner_prod_c(ihi - ilo + 1, jhi - jlo + 1, bd1%valid%dx, bd1%valid%dy, a, b, out), &
                        'inner_prod')
        ip = out
    end function inner_prod

    ! arrayutils.f90:753 (runs on the GPU)
    function pnorm (bd, bx, p) result (pn)
        type(box_data), intent(in) :: bd
        type(box), intent(in)      :: bx
        integer, intent(in)        :: p
        real(dp)                   :: pn
        real(dp), allocatable      :: a(:,:)
        real(c_double)             :: out
        allocate (a(bx%ilo:bx%ihi, bx%jlo:bx%jhi))
        a = bd%data(bx%ilo:bx%ihi, bx%jlo:bx%jhi)
        call mgp_check (mgp_pnorm_c(bx%nx, bx%ny, bd%bx%dx, bd%bx%dy, a, p, out), 'pnorm')
        pn = out
    end function pnorm

    ! arrayutils.f90:670 (runs on the GPU)
    function integrate2d (bd, bx, geo, opt_jscale) result (sum)
        type(box_data), intent(in) :: bd
        type(box), intent(in)      :: bx
        type(geo_data), intent(in) :: geo
        logical, optional          :: opt_jscale
        real(dp) :: sum
        real(dp), allocatable, target :: a(:,:), jac(:,:)
        real(c_double) :: out
        logical  :: jscale
        if ((bd%offi .ne. BD_CELL) .or. (bd%offj .ne. BD_CELL)) then
            print*, 'pnorm: Only works with cell-centered data for now.'
            stop
        endif
        jscale = .true.
        if (present(opt_jscale)) jscale = opt_jscale
        allocate (a(bx%ilo:bx%ihi, bx%jlo:bx%jhi))
        a = bd%data(bx%ilo:bx%ihi, bx%jlo:bx%jhi)
        if (jscale) then
            allocate (jac(bx%ilo:bx%ihi, bx%jlo:bx%jhi))
            jac = geo%J%data(bx%ilo:bx%ihi, bx%jlo:bx%jhi)
            call mgp_check (mgp_integrate2d_c(bx%nx, bx%ny, geo%dx, geo%dy, a, c_loc(jac), out), 'integrate2d')
        else
            call mgp_check (mgp_integrate2d_c(bx%nx, bx%ny, geo%dx, geo%dy, a, c_null_ptr, out), 'integrate2d')
        endif
        sum = out
    end function integrate2d

    ! arrayutils.f90:1497 (orders 0..3, runs on the GPU)
    subroutine extrapolate_ghosts (phi, order)
        type(box_data), intent(inout) :: phi
        integer, intent(in)           :: order
        real(dp), allocatable :: a(:,:)
        if (order .lt. 0 .or. order .gt. 3) then
            print*, 'extrapolate_ghosts: order = ', order, ' is not supported.'
            stop
        endif
        call pack_phi (a, phi)
        call mgp_check (mgp_extrapolate_ghosts_c(phi%valid%nx, phi%valid%ny, order, a), 'extrapolate_ghosts')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
    end subroutine extrapolate_ghosts

    ! ---- marshalling helpers -------------------------------------------------
    subroutine pack_geo (gp, gc, geo, valid)
        type(geo_pack), target, intent(inout) :: gp
        type(mgp_geo_c), intent(out)          :: gc
        type(geo_data), intent(in)            :: geo
        type(box), intent(in)                 :: valid
        integer :: ilo, ihi, jlo, jhi
        ilo = valid%ilo
        ihi = valid%ihi
        jlo = valid%jlo
        jhi = valid%jhi
        allocate (gp%J(ilo:ihi, jlo:jhi), gp%Jxx(ilo:ihi+1, jlo:jhi), gp%Jxy(ilo:ihi+1, jlo:jhi))
        allocate (gp%Jyx(ilo:ihi, jlo:jhi+1), gp%Jyy(ilo:ihi, jlo:jhi+1))
        gp%J = geo%J%data(ilo:ihi, jlo:jhi)
        gp%Jxx = geo%Jgup_xx%data(ilo:ihi+1, jlo:jhi)
        gp%Jxy = geo%Jgup_xy%data(ilo:ihi+1, jlo:jhi)
        gp%Jyx = geo%Jgup_yx%data(ilo:ihi, jlo:jhi+1)
        gp%Jyy = geo%Jgup_yy%data(ilo:ihi, jlo:jhi+1)
        gc%nx = valid%nx
        gc%ny = valid%ny
        gc%dx = geo%dx
        gc%dy = geo%dy
        gc%J = c_loc(gp%J)
        gc%Jxx = c_loc(gp%Jxx)
        gc%Jxy = c_loc(gp%Jxy)
        gc%Jyx = c_loc(gp%Jyx)
        gc%Jyy = c_loc(gp%Jyy)
    end subroutine pack_geo

    subroutine pack_side (dst, src)
        real(dp), allocatable, target, intent(inout) :: dst(:)
        real(dp), allocatable, intent(in)            :: src(:)
        if (allocated(src)) then
            allocate (dst(size(src)))
            dst = src
        else
            allocate (dst(1))
            dst = zero
        endif
    end subroutine pack_side

    subroutine pack_bc (bp, bcc, bc)
        type(bc_pack), target, intent(inout) :: bp
        type(mgp_bc_c), intent(out)          :: bcc
        type(bdry_data), intent(in)          :: bc
        bcc%btype = [bc%type_xlo, bc%type_xhi, bc%type_ylo, bc%type_yhi]
        bcc%bmode = [bc%mode_xlo, bc%mode_xhi, bc%mode_ylo, bc%mode_yhi]
        call pack_side (bp%d1, bc%data_xlo)
        call pack_side (bp%d2, bc%data_xhi)
        call pack_side (bp%d3, bc%data_ylo)
        call pack_side (bp%d4, bc%data_yhi)
        bcc%bdata(1) = c_loc(bp%d1)
        bcc%bdata(2) = c_loc(bp%d2)
        bcc%bdata(3) = c_loc(bp%d3)
        bcc%bdata(4) = c_loc(bp%d4)
    end subroutine pack_bc

    ! one-ghost copy of a cell field's valid region and ghost ring
    subroutine pack_phi (a, f)
        real(dp), allocatable, intent(out) :: a(:,:)
        type(box_data), intent(in)         :: f
        if (f%ngx .ne. 1 .or. f%ngy .ne. 1) then
            print*, 'Only one ghost layer is supported for this field'
            stop
        endif
        allocate (a(f%valid%ilo-1:f%valid%ihi+1, f%valid%jlo-1:f%valid%jhi+1))
        a = f%data(f%valid%ilo-1:f%valid%ihi+1, f%valid%jlo-1:f%valid%jhi+1)
    end subroutine pack_phi

    subroutine pack_cc (a, f)
        real(dp), allocatable, intent(out) :: a(:,:)
        type(box_data), intent(in)         :: f
        allocate (a(f%valid%ilo:f%valid%ihi, f%valid%jlo:f%valid%jhi))
        a = f%data(f%valid%ilo:f%valid%ihi, f%valid%jlo:f%valid%jhi)
    end subroutine pack_cc

    ! arrayutils.f90:1277 fill_ghosts (runs on the GPU)
    subroutine fill_ghosts (phi, bc, geo, homog, do_neum_opt)
        type(box_data), intent(inout) :: phi
        type(bdry_data), intent(in)   :: bc
        type(geo_data), intent(in)    :: geo
        logical, intent(in)           :: homog
        logical, intent(in), optional :: do_neum_opt
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:)
        logical :: do_neum
        do_neum = .true.
        if (present(do_neum_opt)) do_neum = do_neum_opt
        call pack_geo (gp, gc, geo, phi%valid)
        call pack_bc (bp, bcc, bc)
        call pack_phi (a, phi)
        call mgp_check (mgp_fill_ghosts_c(gc, bcc, merge(1, 0, homog), merge(1, 0, do_neum), a), 'fill_ghosts')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
    end subroutine fill_ghosts

    ! arrayutils.f90:1571 fill_boundary_fluxes (runs on the GPU).  Neumann
    ! faces of the valid region take the BC flux; other faces are untouched.
    subroutine fill_boundary_fluxes (xflux, yflux, bc, homog)
        type(box_data), intent(inout) :: xflux, yflux
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        type(bc_pack) :: bp
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: xf(:,:), yf(:,:)
        call pack_bc (bp, bcc, bc)
        call pack_cc (xf, xflux)
        call pack_cc (yf, yflux)
        call mgp_check (mgp_fill_boundary_fluxes_c(yflux%valid%nx, xflux%valid%ny, bcc, merge(1, 0, homog), &
                                                   xf, yf), 'fill_boundary_fluxes')
        xflux%data(xflux%valid%ilo:xflux%valid%ihi, xflux%valid%jlo:xflux%valid%jhi) = xf
        yflux%data(yflux%valid%ilo:yflux%valid%ihi, yflux%valid%jlo:yflux%valid%jhi) = yf
    end subroutine fill_boundary_fluxes

end module ArrayUtils


! poisson2d.f90 -- operator, residual, inverse diagonal, GS, BiCGStab2 (GPU)
module Poisson2D
    use ArrayUtils
    implicit none

    interface
        integer(c_int) function c_laplacian(geo, bc, homog, jscale, phi, lap) bind(C, name="mgp_compute_laplacian")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog, jscale
            real(c_double)              :: phi(*), lap(*)
        end function
        integer(c_int) function c_residual(geo, bc, homog, rhs, phi, res) bind(C, name="mgp_compute_residual")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog
            real(c_double), intent(in)  :: rhs(*)
            real(c_double)              :: phi(*), res(*)
        end function
        integer(c_int) function c_invdiags(geo, bc, homog, invd) bind(C, name="mgp_compute_invdiags")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog
            real(c_double)              :: invd(*)
        end function
        integer(c_int) function c_relax_gs(geo, bc, homog, invd, omega, maxiters, rhs, phi, ordering) &
                                           bind(C, name="mgp_relax_gs")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog, maxiters, ordering
            real(c_double), value       :: omega
            real(c_double), intent(in)  :: invd(*), rhs(*)
            real(c_double)              :: phi(*)
        end function
        integer(c_int) function c_pd(nx, ny, dx, dy, phi, dir, nodedir, pd) bind(C, name="mgp_compute_pd")
            import :: c_int, c_double
            integer(c_int), value       :: nx, ny, dir, nodedir
            real(c_double), value       :: dx, dy
            real(c_double), intent(in)  :: phi(*)
            real(c_double)              :: pd(*)
        end function
        integer(c_int) function c_grad(geo, bc, homog, phi, xflux, yflux) bind(C, name="mgp_compute_grad")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog
            real(c_double)              :: phi(*), xflux(*), yflux(*)
        end function
        integer(c_int) function c_div(nx, ny, dx, dy, xflux, yflux, div) bind(C, name="mgp_compute_div")
            import :: c_int, c_double
            integer(c_int), value       :: nx, ny
            real(c_double), value       :: dx, dy
            real(c_double), intent(in)  :: xflux(*), yflux(*)
            real(c_double)              :: div(*)
        end function
        integer(c_int) function c_jacobi(geo, bc, homog, invd, omega, tol, maxiters, zerophi, rhs, phi, ordering, &
                                         relres, iters) bind(C, name="mgp_relax_jacobi")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog, maxiters, zerophi, ordering
            real(c_double), value       :: omega, tol
            real(c_double), intent(in)  :: invd(*), rhs(*)
            real(c_double)              :: phi(*), relres(*)
            integer(c_int)              :: iters
        end function
        integer(c_int) function c_bicgstab1(geo, bc, homog, invd, tol, maxiters, maxrestarts, zerophi, rhs, phi, &
                                            ordering, relres, last) bind(C, name="mgp_solve_bicgstab")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog, maxiters, maxrestarts, zerophi, ordering
            real(c_double), value       :: tol
            real(c_double), intent(in)  :: invd(*), rhs(*)
            real(c_double)              :: phi(*), relres(*)
            integer(c_int)              :: last
        end function
        integer(c_int) function c_bicgstab2(kind, geo, bc, homog, invd, tol, maxiters, maxrestarts, zerophi, &
                                            rhs, phi, ordering) bind(C, name="mgp_solve_bicgstab2")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            integer(c_int), value       :: kind, homog, maxiters, maxrestarts, zerophi, ordering
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            real(c_double), value       :: tol
            real(c_double), intent(in)  :: invd(*), rhs(*)
            real(c_double)              :: phi(*)
        end function
    end interface

contains

    ! poisson2d.f90:16
    subroutine compute_invdiags (invdiags, geo, bc, homog)
        type(box_data), intent(inout) :: invdiags
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:)
        call pack_geo (gp, gc, geo, invdiags%valid)
        call pack_bc (bp, bcc, bc)
        call pack_cc (a, invdiags)
        call mgp_check (c_invdiags(gc, bcc, merge(1, 0, homog), a), 'Poisson2d::compute_invdiags')
        invdiags%data(invdiags%valid%ilo:invdiags%valid%ihi, invdiags%valid%jlo:invdiags%valid%jhi) = a
    end subroutine compute_invdiags

    ! poisson2d.f90:404 -- pd must be nodal in exactly one direction; phi's
    ! ghosts are read as they are.
    subroutine compute_pd (pd, phi, dir)
        type(box_data), intent(inout) :: pd
        type(box_data), intent(in)    :: phi
        integer, intent(in)           :: dir
        real(dp), allocatable :: a(:,:), d(:,:)
        integer :: nodedir, ilo, ihi, jlo, jhi
        if ((pd%offi .ne. 0) .and. (pd%offj .ne. 0)) then
            print*, 'compute_pd cannot handle totally cell-centered data.'
            stop
        endif
        if ((pd%offi .eq. 0) .and. (pd%offj .eq. 0)) then
            print*, 'compute_pd cannot handle totally nodal-centered data.'
            stop
        endif
        nodedir = merge(1, 2, pd%offi .eq. 0)
        ilo = phi%valid%ilo
        ihi = phi%valid%ihi + merge(1, 0, nodedir .eq. 1)
        jlo = phi%valid%jlo
        jhi = phi%valid%jhi + merge(1, 0, nodedir .eq. 2)
        call pack_phi (a, phi)
        allocate (d(ilo:ihi, jlo:jhi))
        call mgp_check (c_pd(phi%valid%nx, phi%valid%ny, phi%valid%dx, phi%valid%dy, a, dir, nodedir, d), &
                        'Poisson2d::compute_pd')
        pd%data(ilo:ihi, jlo:jhi) = d
    end subroutine compute_pd

    ! poisson2d.f90:809 -- xwk and ywk are accepted for the reference's
    ! argument list; the GPU path needs no workspace from the caller.
    subroutine compute_grad (xflux, yflux, phi, geo, bc, homog, xwk, ywk)
        type(box_data), intent(inout) :: xflux, yflux
        type(box_data), intent(inout) :: phi
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        type(box_data), intent(inout) :: xwk, ywk
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), xf(:,:), yf(:,:)
        integer :: ilo, ihi, jlo, jhi
        ilo = phi%valid%ilo
        ihi = phi%valid%ihi
        jlo = phi%valid%jlo
        jhi = phi%valid%jhi
        call pack_geo (gp, gc, geo, phi%valid)
        call pack_bc (bp, bcc, bc)
        call pack_phi (a, phi)
        allocate (xf(ilo:ihi+1, jlo:jhi), yf(ilo:ihi, jlo:jhi+1))
        call mgp_check (c_grad(gc, bcc, merge(1, 0, homog), a, xf, yf), 'Poisson2d::compute_grad')
        phi%data(ilo-1:ihi+1, jlo-1:jhi+1) = a
        xflux%data(ilo:ihi+1, jlo:jhi) = xf
        yflux%data(ilo:ihi, jlo:jhi+1) = yf
    end subroutine compute_grad

    ! poisson2d.f90:860
    subroutine compute_div (div, xflux, yflux)
        type(box_data), intent(inout) :: div
        type(box_data), intent(in)    :: xflux, yflux
        real(dp), allocatable :: xf(:,:), yf(:,:), d(:,:)
        integer :: ilo, ihi, jlo, jhi
        ilo = div%valid%ilo
        ihi = div%valid%ihi
        jlo = div%valid%jlo
        jhi = div%valid%jhi
        allocate (xf(ilo:ihi+1, jlo:jhi), yf(ilo:ihi, jlo:jhi+1), d(ilo:ihi, jlo:jhi))
        xf = xflux%data(ilo:ihi+1, jlo:jhi)
        yf = yflux%data(ilo:ihi, jlo:jhi+1)
        call mgp_check (c_div(div%valid%nx, div%valid%ny, div%valid%dx, div%valid%dy, xf, yf, d), &
                        'Poisson2d::compute_div')
        div%data(ilo:ihi, jlo:jhi) = d
    end subroutine compute_div

    ! poisson2d.f90:886
    subroutine compute_laplacian (lap, phi, geo, bc, homog, opt_jscale)
        type(box_data), intent(inout) :: lap
        type(box_data), intent(inout) :: phi
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        logical, intent(in), optional :: opt_jscale
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), l(:,:)
        integer :: js
        js = 0
        if (present(opt_jscale)) js = merge(1, 0, opt_jscale)
        call pack_geo (gp, gc, geo, lap%valid)
        call pack_bc (bp, bcc, bc)
        call pack_phi (a, phi)
        call pack_cc (l, lap)
        call mgp_check (c_laplacian(gc, bcc, merge(1, 0, homog), js, a, l), 'Poisson2d::compute_laplacian')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
        lap%data(lap%valid%ilo:lap%valid%ihi, lap%valid%jlo:lap%valid%jhi) = l
    end subroutine compute_laplacian

    ! poisson2d.f90:1327
    subroutine compute_residual (res, rhs, phi, geo, bc, homog, opt_jscale)
        type(box_data), intent(inout) :: res, phi
        type(box_data), intent(in)    :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        logical, intent(in), optional :: opt_jscale
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), r(:,:), f(:,:)
        logical :: jscale
        integer :: ilo, ihi, jlo, jhi
        jscale = .false.
        if (present(opt_jscale)) jscale = opt_jscale
        ilo = rhs%valid%ilo
        ihi = rhs%valid%ihi
        jlo = rhs%valid%jlo
        jhi = rhs%valid%jhi
        if (jscale) then
            call compute_laplacian (res, phi, geo, bc, homog, .true.)
            res%data(ilo:ihi,jlo:jhi) = rhs%data(ilo:ihi,jlo:jhi) - res%data(ilo:ihi,jlo:jhi)
            return
        endif
        call pack_geo (gp, gc, geo, res%valid)
        call pack_bc (bp, bcc, bc)
        call pack_phi (a, phi)
        call pack_cc (f, rhs)
        call pack_cc (r, res)
        call mgp_check (c_residual(gc, bcc, merge(1, 0, homog), f, a, r), 'Poisson2d::compute_residual')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
        res%data(ilo:ihi,jlo:jhi) = r
    end subroutine compute_residual

    ! poisson2d.f90:1460 -- lexicographic order as in the reference.  A
    ! positive tol monitors the residual after each sweep on the host side.
    subroutine relax_gs (phi, rhs, geo, bc, homog, invdiags, omega, tol, maxiters, zerophi, verbosity, ordering)
        type(box_data), intent(inout)   :: phi
        type(box_data), intent(in)      :: rhs
        type(geo_data), intent(in)      :: geo
        type(bdry_data), intent(in)     :: bc
        logical, intent(in)             :: homog
        type(box_data), intent(in)      :: invdiags
        real(dp), intent(in)            :: omega, tol
        integer, intent(in)             :: maxiters
        logical, intent(in)             :: zerophi
        integer, intent(in)             :: verbosity
        integer, intent(in), optional   :: ordering
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), f(:,:), iv(:,:)
        type(box_data) :: r
        real(dp) :: rscale, rel
        integer :: iter, ord
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (maxiters .eq. 0) return
        if (zerophi) phi%data = zero
        call pack_geo (gp, gc, geo, rhs%valid)
        call pack_bc (bp, bcc, bc)
        call pack_cc (f, rhs)
        call pack_cc (iv, invdiags)
        if (tol .le. zero) then
            call pack_phi (a, phi)
            call mgp_check (c_relax_gs(gc, bcc, merge(1, 0, homog), iv, omega, maxiters, f, a, ord), &
                            'Poisson2d::relax_gs')
            phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
            return
        endif
        call define_box_data (r, rhs)
        call compute_residual (r, rhs, phi, geo, bc, homog)
        rscale = pnorm (r, r%valid, 2)
        if (rscale .eq. zero) rscale = one
        do iter = 1, maxiters
            call pack_phi (a, phi)
            call mgp_check (c_relax_gs(gc, bcc, merge(1, 0, homog), iv, omega, 1, f, a, ord), 'Poisson2d::relax_gs')
            phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
            call compute_residual (r, rhs, phi, geo, bc, homog)
            rel = pnorm (r, r%valid, 2) / rscale
            if (verbosity .ge. 3) print*, 'iter ', iter, ': rel |res| = ', rel
            if (rel .le. tol) exit
        enddo
        call undefine_box_data (r)
    end subroutine relax_gs

    ! poisson2d.f90:1358 -- verbosity >= 3 prints the relres history as the
    ! reference does; the optional ordering selects the reduction order.
    subroutine relax_jacobi (phi, rhs, geo, bc, homog, invdiags, omega, tol, maxiters, zerophi, verbosity, &
                             ordering)
        type(box_data), intent(inout)   :: phi
        type(box_data), intent(in)      :: rhs
        type(geo_data), intent(in)      :: geo
        type(bdry_data), intent(in)     :: bc
        logical, intent(in)             :: homog
        type(box_data), intent(in)      :: invdiags
        real(dp), intent(in)            :: omega, tol
        integer, intent(in)             :: maxiters
        logical, intent(in)             :: zerophi
        integer, intent(in)             :: verbosity
        integer, intent(in), optional   :: ordering
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), f(:,:), iv(:,:)
        real(dp) :: relres(0:max(maxiters, 0))
        integer :: ord, n, k
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (maxiters .eq. 0) return
        call pack_geo (gp, gc, geo, rhs%valid)
        call pack_bc (bp, bcc, bc)
        call pack_cc (f, rhs)
        call pack_cc (iv, invdiags)
        call pack_phi (a, phi)
        relres = zero
        n = 0
        call mgp_check (c_jacobi(gc, bcc, merge(1, 0, homog), iv, omega, tol, maxiters, merge(1, 0, zerophi), f, a, &
                                 ord, relres, n), 'Poisson2d::relax_jacobi')
        if (zerophi) phi%data = zero
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
        if (verbosity .ge. 3 .and. tol .gt. zero) then
            do k = 0, n
                print*, 'iter ', k, ': rel |res| = ', relres(k)
            enddo
            if (relres(n) .le. tol) print*, "Converged."
        endif
    end subroutine relax_jacobi

    ! poisson2d.f90:1993 -- the first BiCGStab version; verbosity >= 3 prints
    ! its relres history.  The optional ordering: ORDER_LEX (the reference's
    ! order, bit-identical) or ORDER_MC (grid-wide, for large grids).
    subroutine solve_bicgstab (phi, rhs, geo, bc, homog, invdiags, tol, max_iters, max_restarts, zerophi, &
                               verbosity, ordering)
        type(box_data), intent(inout) :: phi
        type(box_data), intent(in)    :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        type(box_data), intent(in)    :: invdiags
        real(dp), intent(in)          :: tol
        integer, intent(in)           :: max_iters, max_restarts
        logical, intent(in)           :: zerophi
        integer, intent(in)           :: verbosity
        integer, intent(in), optional :: ordering
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), f(:,:), iv(:,:)
        real(dp) :: relres(0:max(max_iters, 0) + max(max_restarts, 0))
        integer :: ord, last, k
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (max_iters .eq. 0) return
        call pack_geo (gp, gc, geo, rhs%valid)
        call pack_bc (bp, bcc, bc)
        call pack_cc (f, rhs)
        call pack_cc (iv, invdiags)
        call pack_phi (a, phi)
        relres = zero
        last = 0
        call mgp_check (c_bicgstab1(gc, bcc, merge(1, 0, homog), iv, tol, max_iters, max_restarts, &
                                    merge(1, 0, zerophi), f, a, ord, relres, last), 'solve_bicgstab')
        if (zerophi) phi%data = zero
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
        if (verbosity .ge. 3) then
            do k = 0, last
                print*, 'iter ', k, ': rel |res| = ', relres(k)
            enddo
        endif
    end subroutine solve_bicgstab

    ! poisson2d.f90:2253
    subroutine solve_bicgstab2 (phi, rhs, geo, bc, homog, invdiags, tol, max_iters, max_restarts, &
                                zerophi, verbosity)
        type(box_data), intent(inout) :: phi
        type(box_data), intent(in)    :: rhs
        type(box_data), intent(in)    :: invdiags
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        real(dp), intent(in)          :: tol
        integer, intent(in)           :: max_iters, max_restarts
        logical, intent(in)           :: zerophi
        integer, intent(in)           :: verbosity
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), f(:,:), iv(:,:)
        if (verbosity .lt. -100) print*, verbosity
        call pack_geo (gp, gc, geo, rhs%valid)
        call pack_bc (bp, bcc, bc)
        call pack_cc (f, rhs)
        call pack_cc (iv, invdiags)
        call pack_phi (a, phi)
        call mgp_check (c_bicgstab2(0, gc, bcc, merge(1, 0, homog), iv, tol, max_iters, max_restarts, &
                                    merge(1, 0, zerophi), f, a, 0), 'Poisson2d::solve_bicgstab2')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
    end subroutine solve_bicgstab2

end module Poisson2D


! MGPoisson2D.f90 -- schedule, transfers, V-cycle (GPU)
module MGPoisson2D
    use Poisson2D
    implicit none

    interface
        integer(c_int) function c_schedule(nx, ny, dx, dy, md_in, md_out, refx, refy) &
                                           bind(C, name="mgp_create_mgschedule")
            import :: c_int, c_double
            integer(c_int), value :: nx, ny, md_in
            real(c_double), value :: dx, dy
            integer(c_int)        :: md_out, refx(*), refy(*)
        end function
        integer(c_int) function c_restrict(kind, fnx, fny, fine, cnx, cny, crse) bind(C, name="mgp_restrict")
            import :: c_int, c_double
            integer(c_int), value      :: kind, fnx, fny, cnx, cny
            real(c_double), intent(in) :: fine(*)
            real(c_double)             :: crse(*)
        end function
        integer(c_int) function c_prolong(fgeo, cgeo, btypes, bmodes, order, fine, crse) bind(C, name="mgp_prolong")
            import :: mgp_geo_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: fgeo, cgeo
            integer(c_int), intent(in)  :: btypes(4), bmodes(4)
            integer(c_int), value       :: order
            real(c_double)              :: fine(*), crse(*)
        end function
        integer(c_int) function c_vcycle(kind, geo, bc, homog, tol, maxiters, maxdepth, numcycles, sdown, sup, &
                                         sbottom, zerophi, phi, rhs, ordering, relres, iters, outcome) &
                                         bind(C, name="mgp_vcycle")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            integer(c_int), value       :: kind, homog, maxiters, maxdepth, numcycles, sdown, sup, sbottom
            integer(c_int), value       :: zerophi, ordering
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            real(c_double), value       :: tol
            real(c_double)              :: phi(*), rhs(*), relres(*)
            integer(c_int)              :: iters, outcome
        end function
    end interface

contains

    ! MGPoisson2D.f90:38 -- this overshoots, exactly like the reference
    subroutine estimate_maxdepth (maxdepth, valid)
        integer, intent(inout) :: maxdepth
        type(box), intent(in)  :: valid
        integer :: mgnx, mgny, d
        mgnx = valid%nx
        mgny = valid%ny
        d = 0
        do while (d .lt. maxdepth)
            if (mod(mgnx,2) .eq. 0) then
                mgnx = mgnx / 2
            else if (mod(mgny,2) .eq. 0) then
                mgny = mgny / 2
            else
                exit
            endif
            d = d + 1
        enddo
        maxdepth = min(maxdepth, d+1)
    end subroutine estimate_maxdepth

    ! MGPoisson2D.f90:67
    subroutine create_mgschedule (refx, refy, maxdepth, valid)
        integer, intent(inout)                        :: maxdepth
        integer, intent(out), dimension(0:maxdepth-1) :: refx, refy
        type(box), intent(in)                         :: valid
        integer(c_int) :: rx(64), ry(64), md
        refx = 0
        refy = 0
        call mgp_check (c_schedule(valid%nx, valid%ny, valid%dx, valid%dy, maxdepth, md, rx, ry), &
                        'create_mgschedule')
        maxdepth = md
        if (md .gt. 0) then
            refx(0:md-1) = rx(1:md)
            refy(0:md-1) = ry(1:md)
        endif
    end subroutine create_mgschedule

    ! MGPoisson2D.f90:135
    subroutine restrict (fine, crse)
        type(box_data), intent(in)    :: fine
        type(box_data), intent(inout) :: crse
        real(dp), allocatable :: f(:,:), c(:,:)
        integer :: kind, fnx, fny, cnx, cny
        kind = 0
        if (fine%offi .eq. BD_NODE .and. fine%offj .eq. BD_CELL) kind = 1
        if (fine%offi .eq. BD_CELL .and. fine%offj .eq. BD_NODE) kind = 2
        if (fine%offi .eq. BD_NODE .and. fine%offj .eq. BD_NODE) then
            print*, 'restrict: nodal data is not supported'
            stop
        endif
        fnx = fine%valid%nx
        fny = fine%valid%ny
        cnx = crse%valid%nx
        cny = crse%valid%ny
        if (kind .eq. 1) then
            fnx = fnx - 1
            cnx = cnx - 1
        else if (kind .eq. 2) then
            fny = fny - 1
            cny = cny - 1
        endif
        call pack_cc (f, fine)
        call pack_cc (c, crse)
        call mgp_check (c_restrict(kind, fnx, fny, f, cnx, cny, c), 'restrict')
        crse%data = zero
        crse%data(crse%valid%ilo:crse%valid%ihi, crse%valid%jlo:crse%valid%jhi) = c
    end subroutine restrict

    ! MGPoisson2D.f90:262
    subroutine prolong (fine, crse, fine_geo, crse_geo, crse_bc, order)
        type(box_data), intent(inout) :: fine
        type(box_data), intent(inout) :: crse
        type(geo_data), intent(in)    :: fine_geo
        type(geo_data), intent(in)    :: crse_geo
        type(bdry_data), intent(in)   :: crse_bc
        integer, intent(in)           :: order
        type(geo_pack) :: fp, cp
        type(mgp_geo_c) :: fg, cg
        real(dp), allocatable :: f(:,:), c(:,:)
        integer(c_int) :: t(4), m(4)
        call pack_geo (fp, fg, fine_geo, fine%valid)
        call pack_geo (cp, cg, crse_geo, crse%valid)
        t = [crse_bc%type_xlo, crse_bc%type_xhi, crse_bc%type_ylo, crse_bc%type_yhi]
        m = [crse_bc%mode_xlo, crse_bc%mode_xhi, crse_bc%mode_ylo, crse_bc%mode_yhi]
        call pack_phi (f, fine)
        call pack_phi (c, crse)
        call mgp_check (c_prolong(fg, cg, t, m, order, f, c), 'prolong')
        fine%data(fine%valid%ilo-1:fine%valid%ihi+1, fine%valid%jlo-1:fine%valid%jhi+1) = f
        crse%data(crse%valid%ilo-1:crse%valid%ihi+1, crse%valid%jlo-1:crse%valid%jhi+1) = c
    end subroutine prolong

    ! Shared body of MGPoisson2D::vcycle (kind 0), DeferredMGPoisson2D::vcycle (kind 1)
    ! and MGPoisson2D::fmg (kind 2).
    subroutine dropin_vcycle (kind, phi, rhs, geo, bc, homog, tol, maxiters, maxdepth_user, numcycles, &
                              smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ordering)
        integer, intent(in)           :: kind
        type(box_data), intent(inout) :: phi, rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog, zerophi
        real(dp), intent(in)          :: tol
        integer, intent(in)           :: maxiters, maxdepth_user, numcycles
        integer, intent(in)           :: smooth_down, smooth_up, smooth_bottom, verbosity, ordering
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), f(:,:), relres(:)
        integer(c_int) :: iters, outcome
        integer :: k
        call pack_geo (gp, gc, geo, rhs%valid)
        call pack_bc (bp, bcc, bc)
        call pack_phi (a, phi)
        call pack_cc (f, rhs)
        allocate (relres(0:max(maxiters, 0)))
        call mgp_check (c_vcycle(kind, gc, bcc, merge(1, 0, homog), tol, maxiters, maxdepth_user, numcycles, &
                                 smooth_down, smooth_up, smooth_bottom, merge(1, 0, zerophi), a, f, ordering, relres, &
                                 iters, outcome), 'vcycle')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
        rhs%data(rhs%valid%ilo:rhs%valid%ihi, rhs%valid%jlo:rhs%valid%jhi) = f
        if (verbosity .ge. 1) then
            do k = 0, iters
                if (kind .eq. 2) then
                    print*, 'FMG-Cycle iter ', k, ': rel |res| = ', relres(k)
                else
                    print*, 'V-Cycle iter ', k, ': rel |res| = ', relres(k)
                endif
            enddo
            if (outcome .eq. 0) print*, 'Converged.'
            if (outcome .eq. 1) print*, 'Diverging.'
        endif
    end subroutine dropin_vcycle

    ! MGPoisson2D.f90:400
    subroutine vcycle (phi, rhs, geo, bc, homog, amrrefx, amrrefy, tol, maxiters, maxdepth_user, numcycles, &
                       smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ordering)
        type(box_data), intent(inout) :: phi
        type(box_data), intent(inout) :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        integer, intent(in)           :: amrrefx, amrrefy
        real(dp), intent(in)          :: tol
        integer, intent(in)           :: maxiters, maxdepth_user, numcycles
        integer, intent(in)           :: smooth_down, smooth_up, smooth_bottom
        logical, intent(in)           :: zerophi
        integer, intent(in)           :: verbosity
        integer, intent(in), optional :: ordering
        integer :: ord
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (amrrefx .lt. -1000000 .or. amrrefy .lt. -1000000) print*, 'unused'
        call dropin_vcycle (0, phi, rhs, geo, bc, homog, tol, maxiters, maxdepth_user, numcycles, &
                            smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ord)
    end subroutine vcycle

    ! MGPoisson2D.f90:860 -- full multigrid; rhs is not rescaled (the reference's
    ! J scaling is commented out there)
    subroutine fmg (phi, rhs, geo, bc, homog, amrrefx, amrrefy, tol, maxiters, maxdepth_user, numcycles, &
                    smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ordering)
        type(box_data), intent(inout) :: phi
        type(box_data), intent(inout) :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        integer, intent(in)           :: amrrefx, amrrefy
        real(dp), intent(in)          :: tol
        integer, intent(in)           :: maxiters, maxdepth_user, numcycles
        integer, intent(in)           :: smooth_down, smooth_up, smooth_bottom
        logical, intent(in)           :: zerophi
        integer, intent(in)           :: verbosity
        integer, intent(in), optional :: ordering
        integer :: ord
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (amrrefx .lt. -1000000 .or. amrrefy .lt. -1000000) print*, 'unused'
        call dropin_vcycle (2, phi, rhs, geo, bc, homog, tol, maxiters, maxdepth_user, numcycles, &
                            smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ord)
    end subroutine fmg

end module MGPoisson2D


! MGPoisson2D_DD.f90:3 -- low-order deferred-correction operator (GPU)
module DeferredPoisson2D
    use ArrayUtils
    implicit none

    interface
        integer(c_int) function c_dc_laplacian(geo, bc, homog, phi, lap) bind(C, name="mgp_dc_compute_laplacian")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog
            real(c_double)              :: phi(*), lap(*)
        end function
        integer(c_int) function c_dc_invdiags(geo, invd) bind(C, name="mgp_dc_compute_invdiags")
            import :: mgp_geo_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            real(c_double)              :: invd(*)
        end function
        integer(c_int) function c_dc_relax_gs(geo, bc, homog, invd, omega, maxiters, rhs, phi, ordering) &
                                              bind(C, name="mgp_dc_relax_gs")
            import :: mgp_geo_c, mgp_bc_c, c_int, c_double
            type(mgp_geo_c), intent(in) :: geo
            type(mgp_bc_c), intent(in)  :: bc
            integer(c_int), value       :: homog, maxiters, ordering
            real(c_double), value       :: omega
            real(c_double), intent(in)  :: invd(*), rhs(*)
            real(c_double)              :: phi(*)
        end function
    end interface

contains

    ! MGPoisson2D_DD.f90:16
    subroutine compute_invdiags (invdiags, geo)
        type(box_data), intent(inout) :: invdiags
        type(geo_data), intent(in)    :: geo
        type(geo_pack) :: gp
        type(mgp_geo_c) :: gc
        real(dp), allocatable :: a(:,:)
        call pack_geo (gp, gc, geo, invdiags%valid)
        call pack_cc (a, invdiags)
        call mgp_check (c_dc_invdiags(gc, a), 'DeferredPoisson2D::compute_invdiags')
        invdiags%data(invdiags%valid%ilo:invdiags%valid%ihi, invdiags%valid%jlo:invdiags%valid%jhi) = a
    end subroutine compute_invdiags

    ! MGPoisson2D_DD.f90:45
    subroutine compute_laplacian (lap, phi, geo, bc, homog)
        type(box_data), intent(inout) :: lap, phi
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), l(:,:)
        call pack_geo (gp, gc, geo, lap%valid)
        call pack_bc (bp, bcc, bc)
        call pack_phi (a, phi)
        call pack_cc (l, lap)
        call mgp_check (c_dc_laplacian(gc, bcc, merge(1, 0, homog), a, l), 'DeferredPoisson2D::compute_laplacian')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
        lap%data(lap%valid%ilo:lap%valid%ihi, lap%valid%jlo:lap%valid%jhi) = l
    end subroutine compute_laplacian

    ! MGPoisson2D_DD.f90:127
    subroutine compute_residual (res, rhs, phi, geo, bc, homog)
        type(box_data), intent(inout) :: res, phi
        type(box_data), intent(in)    :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        integer :: ilo, ihi, jlo, jhi
        call compute_laplacian (res, phi, geo, bc, homog)
        ilo = rhs%valid%ilo
        ihi = rhs%valid%ihi
        jlo = rhs%valid%jlo
        jhi = rhs%valid%jhi
        res%data(ilo:ihi,jlo:jhi) = rhs%data(ilo:ihi,jlo:jhi) - res%data(ilo:ihi,jlo:jhi)
    end subroutine compute_residual

    ! MGPoisson2D_DD.f90:153 (lexicographic, as compiled in the reference)
    subroutine relax_gs (phi, rhs, geo, bc, homog, invdiags, omega, tol, maxiters, zerophi, verbosity, ordering)
        type(box_data), intent(inout) :: phi
        type(box_data), intent(in)    :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        type(box_data), intent(in)    :: invdiags
        real(dp), intent(in)          :: omega, tol
        integer, intent(in)           :: maxiters
        logical, intent(in)           :: zerophi
        integer, intent(in)           :: verbosity
        integer, intent(in), optional :: ordering
        type(geo_pack) :: gp
        type(bc_pack) :: bp
        type(mgp_geo_c) :: gc
        type(mgp_bc_c) :: bcc
        real(dp), allocatable :: a(:,:), f(:,:), iv(:,:)
        integer :: ord
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (maxiters .eq. 0) return
        if (tol .gt. zero .and. verbosity .ge. 3) print*, 'relax_gs: residual monitoring is not implemented'
        if (zerophi) phi%data = zero
        call pack_geo (gp, gc, geo, rhs%valid)
        call pack_bc (bp, bcc, bc)
        call pack_cc (f, rhs)
        call pack_cc (iv, invdiags)
        call pack_phi (a, phi)
        call mgp_check (c_dc_relax_gs(gc, bcc, merge(1, 0, homog), iv, omega, maxiters, f, a, ord), &
                        'DeferredPoisson2D::relax_gs')
        phi%data(phi%valid%ilo-1:phi%valid%ihi+1, phi%valid%jlo-1:phi%valid%jhi+1) = a
    end subroutine relax_gs

end module DeferredPoisson2D


! MGPoisson2D_DD.f90:574 -- deferred-correction V-cycle (GPU)
module DeferredMGPoisson2D
    use ArrayUtils
    use MGPoisson2D, only: estimate_maxdepth, create_mgschedule, restrict, prolong, dropin_vcycle
    implicit none

contains

    ! MGPoisson2D_DD.f90:589
    subroutine vcycle (phi, rhs, geo, bc, homog, amrrefx, amrrefy, tol, maxiters, maxdepth_user, numcycles, &
                       smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ordering)
        type(box_data), intent(inout) :: phi
        type(box_data), intent(inout) :: rhs
        type(geo_data), intent(in)    :: geo
        type(bdry_data), intent(in)   :: bc
        logical, intent(in)           :: homog
        integer, intent(in)           :: amrrefx, amrrefy
        real(dp), intent(in)          :: tol
        integer, intent(in)           :: maxiters, maxdepth_user, numcycles
        integer, intent(in)           :: smooth_down, smooth_up, smooth_bottom
        logical, intent(in)           :: zerophi
        integer, intent(in)           :: verbosity
        integer, intent(in), optional :: ordering
        integer :: ord
        ord = ORDER_LEX
        if (present(ordering)) ord = ordering
        if (amrrefx .lt. -1000000 .or. amrrefy .lt. -1000000) print*, 'unused'
        call dropin_vcycle (1, phi, rhs, geo, bc, homog, tol, maxiters, maxdepth_user, numcycles, &
                            smooth_down, smooth_up, smooth_bottom, zerophi, verbosity, ord)
    end subroutine vcycle

end module DeferredMGPoisson2D
