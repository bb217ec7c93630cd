! b4step hook with the argument list used by src/sharpclaw.f90:129.  It
! appends (told, dt) of every new step attempt to dtlog.bin so the exact
! time-step sequence of the reference run can be compared.
subroutine b4step(maxnx,mbc,nx,meqn,q,xlower,dx,told,dt,maux,aux)
    implicit double precision (a-h,o-z)
    dimension nx(*), xlower(*), dx(*), q(*), aux(*)
    logical, save :: opened = .false.
    if (.not. opened) then
        open(unit=52, file='dtlog.bin', access='stream', form='unformatted', status='replace')
        opened = .true.
    endif
    write(52) told, dt
    flush(52)
end subroutine b4step
