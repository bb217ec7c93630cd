! The bc call of a 1D step (src/1d/step1d.f90:20), in a file of its own so
! that the 1D and 2D argument lists of the user's bc never meet in one
! compilation unit.
subroutine host_bc1(q, aux, t, dt)
    use ClawParams
    implicit none
    double precision :: q(*), aux(*), t, dt
    call bc(maxnx, meqn, mbc, nx(1), xlower(1), dx(1), q, maux, aux, t, dt, mthbc)
end subroutine host_bc1
