! Default user hooks with the argument lists the drivers use:
! b4step as called by src/sharpclaw.f90:129, setaux as src/main.f90:165.
subroutine b4step(maxnx,mbc,nx,meqn,q,xlower,dx,t,dt,maux,aux)
    implicit double precision (a-h,o-z)
    dimension nx(*), xlower(*), dx(*), q(*), aux(*)
    return
end subroutine b4step

subroutine setaux(maxnx,mbc,nx,xlower,dx,maux,aux)
    implicit double precision (a-h,o-z)
    dimension nx(*), xlower(*), dx(*), aux(*)
    return
end subroutine setaux

! skip the '#' header of a data file and the first line after it
subroutine skip_data_header(iunit)
    integer, intent(in) :: iunit
    character(len=1) :: c
    c = '#'
    do while (c == '#')
        read(iunit, '(a1)') c
    enddo
end subroutine skip_data_header
