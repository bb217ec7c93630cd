! setprob for linear advection: setprob.data holds u, beta, ic (the layout
! of apps/advection/1d/setprob.data); u goes to /comrp/.
subroutine setprob
    implicit double precision (a-h,o-z)
    common /comrp/ u
    common /comic/ beta, ic
    open(unit=7, file='setprob.data', status='old', form='formatted')
    call skip_data_header(7)
    read(7,*) u
    read(7,*) beta
    read(7,*) ic
    close(7)
end subroutine setprob
