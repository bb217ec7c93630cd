! Binary frame writer substituted for src/2d/out2.f90 in the oracle build.
! Writes fort.bNNNN: int32 ndim, nx, ny, meqn, then float64 t and the
! interior q(1:nx,1:ny,1:meqn) in the reference column-major layout, after
! setting |q| < 1e-99 to zero in q itself as out2.f90:60-67 does.
subroutine output(meqn,mbc,nx,xlower,dx,q,t,iframe,aux,maux)
    implicit double precision (a-h,o-z)
    integer :: nx(2)
    double precision :: q(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, meqn)
    double precision :: aux(*)
    character(len=10) :: fname
    do m = 1, meqn
        do j = 1, nx(2)
            do i = 1, nx(1)
                if (dabs(q(i,j,m)) .lt. 1d-99) q(i,j,m) = 0.d0
            enddo
        enddo
    enddo
    write(fname,'(a6,i4.4)') 'fort.b', iframe
    open(unit=51, file=fname, access='stream', form='unformatted', status='replace')
    write(51) 2, nx(1), nx(2), meqn
    write(51) t
    write(51) (((q(i,j,m), i=1,nx(1)), j=1,nx(2)), m=1,meqn)
    close(51)
end subroutine output
