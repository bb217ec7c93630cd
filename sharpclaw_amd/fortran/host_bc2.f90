! The bc call of a 2D step (src/2d/step2d.f90:25), in a file of its own so
! that the 1D and 2D argument lists of the user's bc never meet in one
! compilation unit.
subroutine host_bc2(q, aux, t, dt)
    use ClawParams
    implicit none
    double precision :: q(*), aux(*), t, dt
    call bc(meqn, mbc, nx, xlower, dx, q, maux, aux, t, dt, mthbc)
end subroutine host_bc2
