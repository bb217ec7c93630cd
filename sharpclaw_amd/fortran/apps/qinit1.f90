! Sample initial conditions of the benchmark configurations (user hook).
! Project-written 1D initial conditions, argument list as called by
! src/main.f90:169 (qinit(ndim,meqn,mbc,nx,xlower,dx,q,maux,aux)).
! qinit.data (list-directed): kind, then its parameters.
!   1 advection:  beta x0 a b     q = exp(-beta(x-x0)^2) + [a<=x<=b]
!   2 Euler Shu-Osher: rhol ul pl xs   right of xs: rho=1+0.2sin(5x),u=0,p=1
!   3 shallow water dam break: hl hr x0     (u=0)
!   4 acoustics pulse: amp beta x0          p=amp*exp(-beta(x-x0)^2), u=0
!   5 Euler pulse: amp beta x0   rho=1+amp*e, u=0, p=1+amp*e
! Cell centres x = xlower + (i-1/2)dx as apps/advection/1d/qinit.f90:17.
subroutine qinit(ndim_in,meqn,mbc,nx,xlower,dx,q,maux,aux)
    implicit double precision (a-h,o-z)
    integer :: nx(1)
    double precision :: xlower(1), dx(1)
    double precision :: q(1-mbc:nx(1)+mbc, meqn)
    dimension aux(*)
    common /param/ gamma, gamma1
    integer :: kind
    open(unit=8, file='qinit.data', status='old', form='formatted')
    read(8,*) kind
    select case (kind)
    case (1)
        read(8,*) beta, x0, xa, xb
        do i = 1, nx(1)
            x = xlower(1) + (i-0.5d0)*dx(1)
            q(i,1) = dexp(-beta*(x-x0)**2)
            if (x >= xa .and. x <= xb) q(i,1) = q(i,1) + 1.d0
        enddo
    case (2)
        read(8,*) rhol, ul, pl, xs
        do i = 1, nx(1)
            x = xlower(1) + (i-0.5d0)*dx(1)
            if (x < xs) then
                q(i,1) = rhol
                q(i,2) = rhol*ul
                q(i,3) = pl/gamma1 + 0.5d0*rhol*ul**2
            else
                q(i,1) = 1.d0 + 0.2d0*dsin(5.d0*x)
                q(i,2) = 0.d0
                q(i,3) = 1.d0/gamma1
            endif
        enddo
    case (3)
        read(8,*) hl, hr, x0
        do i = 1, nx(1)
            x = xlower(1) + (i-0.5d0)*dx(1)
            if (x < x0) then
                q(i,1) = hl
            else
                q(i,1) = hr
            endif
            q(i,2) = 0.d0
        enddo
    case (4)
        read(8,*) amp, beta, x0
        do i = 1, nx(1)
            x = xlower(1) + (i-0.5d0)*dx(1)
            q(i,1) = amp*dexp(-beta*(x-x0)**2)
            q(i,2) = 0.d0
        enddo
    case (5)
        read(8,*) amp, beta, x0
        do i = 1, nx(1)
            x = xlower(1) + (i-0.5d0)*dx(1)
            e = amp*dexp(-beta*(x-x0)**2)
            q(i,1) = 1.d0 + e
            q(i,2) = 0.d0
            q(i,3) = (1.d0 + e)/gamma1
        enddo
    case default
        write(6,*) 'qinit: unknown kind', kind
        stop 1
    end select
    close(8)
end subroutine qinit
