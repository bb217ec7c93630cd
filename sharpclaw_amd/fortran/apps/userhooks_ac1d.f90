! Sample user hooks of a 1D acoustics application (project-written), the
! kinds the reference's applications link in place of the library's:
!   bc     a moving wall on the left for mthbc(1) = 0 (the pattern of
!          apps/acoustics/1d/interface/bc1.f95 and apps/elasticity/stegoton/
!          bc1.f: mirror the interior and set the normal velocity to the
!          wall's), the library's rules 1-3 on either side otherwise;
!   src    linear damping, dq = -dt*alpha*q (the step's source, src1.f90);
!   b4step a time reversal: once t reaches trtime, the velocity q(:,2) is
!          negated (apps/elasticity/stegoton/b4step1.f:19-24).
! Their parameters come from userhooks.data (list-directed: a1 omega alpha
! trtime): wall velocity a1*sin(omega*t); q(:,1) is the pressure and
! q(:,2) the velocity (apps/acoustics/1d/homogeneous/rp1.f).
module userhooks_ac1d
    implicit none
    double precision :: a1, omega, alpha, trtime
    logical :: loaded = .false., reversed = .false.
contains
    subroutine load_userhooks
        if (loaded) return
        open(unit=9, file='userhooks.data', status='old', form='formatted')
        read(9,*) a1, omega, alpha, trtime
        close(9)
        loaded = .true.
    end subroutine load_userhooks
end module userhooks_ac1d

subroutine bc(maxmx, meqn, mbc, mx, xlower, dx, q, maux, aux, t, dt, mthbc)
    use userhooks_ac1d
    implicit none
    integer :: maxmx, meqn, mbc, mx, maux
    integer :: mthbc(2)
    double precision :: xlower, dx, t, dt
    double precision :: q(1-mbc:mx+mbc, meqn), aux(*)
    integer :: ibc, m
    double precision :: vwall
    call load_userhooks
    do ibc = 1, mbc
        select case (mthbc(1))
        case (0)
            vwall = a1*dsin(omega*t)
            do m = 1, meqn
                q(1-ibc,m) = q(ibc,m)
            enddo
            q(1-ibc,2) = 2.d0*vwall - q(ibc,2)
        case (1)
            do m = 1, meqn
                q(1-ibc,m) = q(1,m)
            enddo
        case (2)
            do m = 1, meqn
                q(1-ibc,m) = q(mx+1-ibc,m)
            enddo
        case (3)
            do m = 1, meqn
                q(1-ibc,m) = q(ibc,m)
            enddo
            q(1-ibc,2) = -q(ibc,2)
        end select
        select case (mthbc(2))
        case (1)
            do m = 1, meqn
                q(mx+ibc,m) = q(mx,m)
            enddo
        case (2)
            do m = 1, meqn
                q(mx+ibc,m) = q(ibc,m)
            enddo
        case (3)
            do m = 1, meqn
                q(mx+ibc,m) = q(mx+1-ibc,m)
            enddo
            q(mx+ibc,2) = -q(mx+1-ibc,2)
        case default
            write(6,*) '*** ERROR *** mthbc(2)=', mthbc(2), ' has no rule in this bc'
            stop
        end select
    enddo
end subroutine bc

subroutine src(q, dq, aux, t, dt)
    use ClawParams
    use userhooks_ac1d
    implicit none
    double precision, intent(in) :: q(1-mbc:nx(1)+mbc, meqn)
    double precision :: aux, t, dt
    double precision, intent(out) :: dq(1-mbc:nx(1)+mbc, meqn)
    integer :: i, m
    call load_userhooks
    do m = 1, meqn
        do i = 1, nx(1)
            dq(i,m) = -dt*alpha*q(i,m)
        enddo
    enddo
end subroutine src

subroutine b4step(maxnx, mbc, nx, meqn, q, xlower, dx, t, dt, maux, aux)
    use userhooks_ac1d
    implicit none
    integer :: maxnx, mbc, meqn, maux
    integer :: nx(*)
    double precision :: xlower(*), dx(*), t, dt
    double precision :: q(1-mbc:nx(1)+mbc, meqn), aux(*)
    integer :: i
    call load_userhooks
    if (t + 1.d-10 >= trtime .and. .not. reversed) then
        do i = 1-mbc, nx(1)+mbc
            q(i,2) = -q(i,2)
        enddo
        reversed = .true.
    endif
end subroutine b4step
