! Default source of a 2D grid (src/2d/src2.f90): dq = 0 in the interior.
! Called by the driver's host-side step only; an application links its own
! src in its place.
subroutine src(q, dq, aux, t, dt)
    use ClawParams
    implicit none
    double precision, intent(in) :: q, aux, t, dt
    double precision :: dq(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, meqn)
    integer :: i, j, m
    do m = 1, meqn
        do j = 1, nx(2)
            do i = 1, nx(1)
                dq(i,j,m) = 0.d0
            enddo
        enddo
    enddo
end subroutine src
