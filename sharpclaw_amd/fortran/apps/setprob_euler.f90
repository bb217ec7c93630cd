! setprob for the Euler equations: setprob.data holds gamma.
subroutine setprob
    implicit double precision (a-h,o-z)
    common /param/ gamma, gamma1
    open(unit=7, file='setprob.data', status='old', form='formatted')
    call skip_data_header(7)
    read(7,*) gamma
    close(7)
    gamma1 = gamma - 1.d0
end subroutine setprob
