! Sample initial conditions of the benchmark configurations (user hook).
! Project-written 2D initial conditions, argument list as called by
! src/main.f90:169.  qinit.data (list-directed): kind, then parameters.
!   1 Euler four quadrants: x0 y0, then rho u v p for NE, NW, SW, SE
!   2 shallow water pulse: amp beta x0 y0   h=1+amp*exp(-beta r^2), hu=hv=0
!   3 Euler pulse: amp beta x0 y0   rho=p=1+amp*exp(-beta r^2), u=v=0
subroutine qinit(ndim_in,meqn,mbc,nx,xlower,dx,q,maux,aux)
    implicit double precision (a-h,o-z)
    integer :: nx(2)
    double precision :: xlower(2), dx(2)
    double precision :: q(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, meqn)
    dimension aux(*)
    common /param/ gamma, gamma1
    double precision :: st(4,4)
    integer :: kind, k
    open(unit=8, file='qinit.data', status='old', form='formatted')
    read(8,*) kind
    select case (kind)
    case (1)
        read(8,*) x0, y0
        do k = 1, 4
            read(8,*) st(1,k), st(2,k), st(3,k), st(4,k)
        enddo
        do j = 1, nx(2)
            y = xlower(2) + (j-0.5d0)*dx(2)
            do i = 1, nx(1)
                x = xlower(1) + (i-0.5d0)*dx(1)
                if (x > x0 .and. y > y0) then
                    k = 1
                else if (x <= x0 .and. y > y0) then
                    k = 2
                else if (x <= x0 .and. y <= y0) then
                    k = 3
                else
                    k = 4
                endif
                q(i,j,1) = st(1,k)
                q(i,j,2) = st(1,k)*st(2,k)
                q(i,j,3) = st(1,k)*st(3,k)
                q(i,j,4) = st(4,k)/gamma1 + 0.5d0*st(1,k)*(st(2,k)**2 + st(3,k)**2)
            enddo
        enddo
    case (2)
        read(8,*) amp, beta, x0, y0
        do j = 1, nx(2)
            y = xlower(2) + (j-0.5d0)*dx(2)
            do i = 1, nx(1)
                x = xlower(1) + (i-0.5d0)*dx(1)
                q(i,j,1) = 1.d0 + amp*dexp(-beta*((x-x0)**2 + (y-y0)**2))
                q(i,j,2) = 0.d0
                q(i,j,3) = 0.d0
            enddo
        enddo
    case (3)
        read(8,*) amp, beta, x0, y0
        do j = 1, nx(2)
            y = xlower(2) + (j-0.5d0)*dx(2)
            do i = 1, nx(1)
                x = xlower(1) + (i-0.5d0)*dx(1)
                e = amp*dexp(-beta*((x-x0)**2 + (y-y0)**2))
                q(i,j,1) = 1.d0 + e
                q(i,j,2) = 0.d0
                q(i,j,3) = 0.d0
                q(i,j,4) = (1.d0 + e)/gamma1
            enddo
        enddo
    case default
        write(6,*) 'qinit: unknown kind', kind
        stop 1
    end select
    close(8)
end subroutine qinit
