! Default setaux (argument list as src/main.f90:165 calls it): no aux
! values.  An application links its own in its place.
subroutine setaux(maxnx, mbc, nx, xlower, dx, maux, aux)
    implicit none
    integer :: maxnx, mbc, maux
    integer :: nx(*)
    double precision :: xlower(*), dx(*), aux(*)
    return
end subroutine setaux
