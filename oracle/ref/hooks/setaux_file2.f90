! Project hook substituted for src/setaux.f90 in the 2D Euler and shallow-water
! oracle builds: the aux array aux(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, maux) is
! read, column-major float64, from aux.bin in the working directory (written
! by the tests; the capacity column mcapa of flux1.f90:56-61).  main.f90:165
! calls it only when maux > 0.
subroutine setaux(maxmx,mbc,nx,xlower,dx,maux,aux)
    implicit none
    integer :: maxmx, mbc, maux
    integer :: nx(2)
    double precision :: xlower(2), dx(2)
    double precision :: aux(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, maux)
    open(unit=52, file='aux.bin', access='stream', form='unformatted', status='old')
    read(52) aux
    close(52)
end subroutine setaux
