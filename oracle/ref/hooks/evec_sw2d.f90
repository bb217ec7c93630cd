! Project-written eigenvector hook for the 2D shallow-water equations
! (q = h, hu, hv).  Roe averages as apps/shallowWater/2dRadialDam/rpn2sw.f:80-87
! (h = (hl+hr)/2, sqrt(h)-weighted u, v, a = sqrt(g*h)), on cell averages.
! Characteristic fields are ordered (u-c, shear, u+c) like rpn2sw's waves.
! The direction comes from module evec_dir (flux1 passes no ixy).
subroutine evec(maxm,meqn,mbc,mx,q,auxl,auxr,evl,evr)
    use ClawParams, only: nx
    use evec_dir
    implicit double precision (a-h,o-z)
    dimension q(1-mbc:maxm+mbc, meqn)
    dimension auxl(*), auxr(*)
    dimension evl(1-mbc:maxm+mbc, meqn, meqn)
    dimension evr(1-mbc:maxm+mbc, meqn, meqn)
    common /comrp/ g
    integer :: ixy, mu, mv

    ixy = evec_next_ixy(nx(1), nx(2))
    if (ixy == 1) then
        mu = 2
        mv = 3
    else
        mu = 3
        mv = 2
    endif

    do i = 2-mbc, mx+mbc
        h = (q(i-1,1) + q(i,1))*0.50d0
        hsqrtl = dsqrt(q(i-1,1))
        hsqrtr = dsqrt(q(i,1))
        hsq2 = hsqrtl + hsqrtr
        u = (q(i-1,mu)/hsqrtl + q(i,mu)/hsqrtr) / hsq2
        v = (q(i-1,mv)/hsqrtl + q(i,mv)/hsqrtr) / hsq2
        a = dsqrt(g*h)

        evr(i,1,1) = 1.d0
        evr(i,mu,1) = u - a
        evr(i,mv,1) = v
        evr(i,1,2) = 0.d0
        evr(i,mu,2) = 0.d0
        evr(i,mv,2) = 1.d0
        evr(i,1,3) = 1.d0
        evr(i,mu,3) = u + a
        evr(i,mv,3) = v

        z = 0.5d0/a
        evl(i,1,1) = (u + a)*z
        evl(i,1,mu) = -z
        evl(i,1,mv) = 0.d0
        evl(i,2,1) = -v
        evl(i,2,mu) = 0.d0
        evl(i,2,mv) = 1.d0
        evl(i,3,1) = -(u - a)*z
        evl(i,3,mu) = z
        evl(i,3,mv) = 0.d0
    enddo
    return
end subroutine evec
