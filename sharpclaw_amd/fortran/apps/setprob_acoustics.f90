! setprob for acoustics: setprob.data holds rho, K, beta (the layout of
! apps/acoustics/1d/homogeneous/setprob.data); sound speed and impedance
! follow as cc = sqrt(K/rho), zz = cc*rho.
subroutine setprob
    implicit double precision (a-h,o-z)
    common /cparam/ rho, bulk, cc, zz
    common /cqinit/ beta
    open(unit=7, file='setprob.data', status='old', form='formatted')
    call skip_data_header(7)
    read(7,*) rho
    read(7,*) bulk
    read(7,*) beta
    close(7)
    cc = dsqrt(bulk/rho)
    zz = cc*rho
end subroutine setprob
