! Default b4step (argument list as src/sharpclaw.f90:129 calls it): nothing
! to do before a step.  An application links its own in its place.
subroutine b4step(maxnx, mbc, nx, meqn, q, xlower, dx, t, dt, maux, aux)
    implicit none
    integer :: maxnx, mbc, meqn, maux
    integer :: nx(*)
    double precision :: xlower(*), dx(*), q(*), aux(*), t, dt
    return
end subroutine b4step
