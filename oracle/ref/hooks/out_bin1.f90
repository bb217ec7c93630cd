! Binary frame writer substituted for src/1d/out1.f90 in the oracle build
! (link-time substitution, as the reference apps do for their own
! routines).  Writes fort.bNNNN: int32 ndim, mx, meqn, then float64 t and
! the interior q(1:mx,1:meqn) in the reference column-major layout.
! out1.f90 prints only 8 significant digits (out1.f90:70); parity needs all.
! Like out1.f90:63-68 it first sets |q| < 1e-99 to zero in q itself, which
! the run then continues from.
subroutine output(meqn,mbc,mx,xlower,dx,q,t,iframe,aux,maux)
    implicit double precision (a-h,o-z)
    double precision :: q(1-mbc:mx+mbc, meqn)
    double precision :: aux(*)
    character(len=10) :: fname
    do m = 1, meqn
        do i = 1, mx
            if (dabs(q(i,m)) .lt. 1d-99) q(i,m) = 0.d0
        enddo
    enddo
    write(fname,'(a6,i4.4)') 'fort.b', iframe
    open(unit=51, file=fname, access='stream', form='unformatted', status='replace')
    write(51) 1, mx, meqn
    write(51) t
    write(51) ((q(i,m), i=1,mx), m=1,meqn)
    close(51)
end subroutine output
