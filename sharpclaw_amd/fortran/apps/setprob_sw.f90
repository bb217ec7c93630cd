! setprob for shallow water: setprob.data holds g.
subroutine setprob
    implicit double precision (a-h,o-z)
    common /comrp/ grav
    open(unit=7, file='setprob.data', status='old', form='formatted')
    call skip_data_header(7)
    read(7,*) grav
    close(7)
end subroutine setprob
