! Shallow-water setprob: reads g into /comrp/ (the block read by
! rp1sw.f:39 and rpn2sw.f:45) from setprob.data.
subroutine setprob
    implicit double precision (a-h,o-z)
    character*25 fname
    common /comrp/ grav
    fname = 'setprob.data'
    call opendatafile(7, fname)
    read(7,*) grav
    close(7)
end subroutine setprob
