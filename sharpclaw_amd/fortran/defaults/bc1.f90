! Default boundary conditions of a 1D grid on the host (the rules of
! src/1d/bc1.f90:32-107): 1 zero-order extrapolation, 2 periodic, 3 solid
! wall (mirror, component 2 negated); 0 asks for the application's own bc,
! which it links in place of this one.  Called by the driver's host-side
! step only (the device stages synthesise the ghost cells themselves).
subroutine bc(maxmx, meqn, mbc, mx, xlower, dx, q, maux, aux, t, dt, mthbc)
    implicit none
    integer :: maxmx, meqn, mbc, mx, maux
    integer :: mthbc(2)
    double precision :: xlower, dx, t, dt
    double precision :: q(1-mbc:mx+mbc, meqn), aux(*)
    integer :: side, ibc, m, ig, is
    do side = 1, 2
        if (mthbc(side) < 1 .or. mthbc(side) > 3) then
            write(6,*) '*** ERROR *** mthbc(', side, ')=', mthbc(side), ' and no BCs specified in bc1'
            stop
        endif
        do ibc = 1, mbc
            ! ghost cell ig takes interior cell is
            if (side == 1) then
                ig = 1 - ibc
                select case (mthbc(1))
                case (1)
                    is = 1
                case (2)
                    is = mx + 1 - ibc
                case default
                    is = ibc
                end select
            else
                ig = mx + ibc
                select case (mthbc(2))
                case (1)
                    is = mx
                case (2)
                    is = ibc
                case default
                    is = mx + 1 - ibc
                end select
            endif
            do m = 1, meqn
                q(ig,m) = q(is,m)
            enddo
            if (mthbc(side) == 3) q(ig,2) = -q(is,2)
        enddo
    enddo
end subroutine bc
