! Project-written eigenvector hook for the 1D shallow-water equations.
! The reference ships no shallow-water evec.  Roe averages follow
! apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/rp1sw.f:47-49
! (ubar from sqrt(h)-weighted velocities, cbar = sqrt(g*(hl+hr)/2)),
! evaluated on the cell averages q(i-1), q(i).
! Argument list as called by src/flux1.f90:105.
subroutine evec(maxm,meqn,mbc,mx,q,auxl,auxr,evl,evr)
    implicit double precision (a-h,o-z)
    dimension q(1-mbc:maxm+mbc, meqn)
    dimension auxl(*), auxr(*)
    dimension evl(1-mbc:maxm+mbc, meqn, meqn)
    dimension evr(1-mbc:maxm+mbc, meqn, meqn)
    common /comrp/ grav

    do i = 2-mbc, mx+mbc
        ubar = (q(i-1,2)/dsqrt(q(i-1,1)) + q(i,2)/dsqrt(q(i,1))) / &
               (dsqrt(q(i-1,1)) + dsqrt(q(i,1)))
        cbar = dsqrt(0.5d0*grav*(q(i-1,1) + q(i,1)))

        evr(i,1,1) = 1.d0
        evr(i,1,2) = 1.d0
        evr(i,2,1) = ubar - cbar
        evr(i,2,2) = ubar + cbar

        z = 0.5d0/cbar
        evl(i,1,1) = (ubar + cbar)*z
        evl(i,1,2) = -z
        evl(i,2,1) = -(ubar - cbar)*z
        evl(i,2,2) = z
    enddo
    return
end subroutine evec
