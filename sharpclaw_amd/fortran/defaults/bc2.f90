! Default boundary conditions of a 2D grid on the host (the rules of
! src/2d/bc2.f90:31-206): the x ghosts of every row (ghost rows included)
! first, then the y ghosts of every column (ghost columns included), so the
! corners come from the y rule; 1 zero-order extrapolation, 2 periodic,
! 3 solid wall (mirror, the normal momentum -- component 2 in x, 3 in y --
! negated); 0 asks for the application's own bc, which it links in place of
! this one.  Called by the driver's host-side step only.
subroutine bc(meqn, mbc, nx, xlower, dx, q, maux, aux, t, dt, mthbc)
    implicit none
    integer :: meqn, mbc, maux
    integer :: nx(2), mthbc(4)
    double precision :: xlower(2), dx(2), t, dt
    double precision :: q(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, meqn), aux(*)
    integer :: side, ibc, m, i, j, ig, is, n, k
    do side = 1, 4
        if (mthbc(side) < 1 .or. mthbc(side) > 3) then
            write(6,*) '*** ERROR *** mthbc(', side, ')=', mthbc(side), ' and no BCs specified in bc2'
            stop
        endif
    enddo
    do side = 1, 4
        k = mthbc(side)
        if (side <= 2) then
            n = nx(1)
        else
            n = nx(2)
        endif
        do ibc = 1, mbc
            if (mod(side, 2) == 1) then
                ig = 1 - ibc
                if (k == 1) is = 1
                if (k == 2) is = n + 1 - ibc
                if (k == 3) is = ibc
            else
                ig = n + ibc
                if (k == 1) is = n
                if (k == 2) is = ibc
                if (k == 3) is = n + 1 - ibc
            endif
            if (side <= 2) then
                do m = 1, meqn
                    do j = 1-mbc, nx(2)+mbc
                        q(ig,j,m) = q(is,j,m)
                    enddo
                enddo
                if (k == 3) then
                    do j = 1-mbc, nx(2)+mbc
                        q(ig,j,2) = -q(is,j,2)
                    enddo
                endif
            else
                do m = 1, meqn
                    do i = 1-mbc, nx(1)+mbc
                        q(i,ig,m) = q(i,is,m)
                    enddo
                enddo
                if (k == 3) then
                    do i = 1-mbc, nx(1)+mbc
                        q(i,ig,3) = -q(i,is,3)
                    enddo
                endif
            endif
        enddo
    enddo
end subroutine bc
