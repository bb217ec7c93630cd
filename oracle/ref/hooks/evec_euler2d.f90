! Project-written eigenvector hook for the 2D Euler equations
! (q = rho, rho*u, rho*v, E), one slice at a time.  mu/mv select the
! normal/tangential momentum like apps/shallowWater/2dRadialDam/rpn2sw.f:60-66.
! Roe averages as in rpn2_euler.f90 of these hooks, on cell averages.
! Characteristic fields are ordered (u-c, entropy, shear, u+c).
! The direction comes from module evec_dir (flux1 passes no ixy).
subroutine evec(maxm,meqn,mbc,mx,q,auxl,auxr,evl,evr)
    use ClawParams, only: nx
    use evec_dir
    implicit double precision (a-h,o-z)
    dimension q(1-mbc:maxm+mbc, meqn)
    dimension auxl(*), auxr(*)
    dimension evl(1-mbc:maxm+mbc, meqn, meqn)
    dimension evr(1-mbc:maxm+mbc, meqn, meqn)
    common /param/ gamma, gamma1
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
        rhsqrtl = dsqrt(q(i-1,1))
        rhsqrtr = dsqrt(q(i,1))
        pl = gamma1*(q(i-1,4) - 0.5d0*(q(i-1,mu)**2 + q(i-1,mv)**2)/q(i-1,1))
        pr = gamma1*(q(i,4) - 0.5d0*(q(i,mu)**2 + q(i,mv)**2)/q(i,1))
        rhsq2 = rhsqrtl + rhsqrtr
        u = (q(i-1,mu)/rhsqrtl + q(i,mu)/rhsqrtr) / rhsq2
        v = (q(i-1,mv)/rhsqrtl + q(i,mv)/rhsqrtr) / rhsq2
        enth = ((q(i-1,4)+pl)/rhsqrtl + (q(i,4)+pr)/rhsqrtr) / rhsq2
        u2v2 = u**2 + v**2
        a2 = gamma1*(enth - 0.5d0*u2v2)
        a = dsqrt(a2)

        ! right eigenvectors (columns)
        evr(i,1,1) = 1.d0
        evr(i,mu,1) = u - a
        evr(i,mv,1) = v
        evr(i,4,1) = enth - u*a

        evr(i,1,2) = 1.d0
        evr(i,mu,2) = u
        evr(i,mv,2) = v
        evr(i,4,2) = 0.5d0*u2v2

        evr(i,1,3) = 0.d0
        evr(i,mu,3) = 0.d0
        evr(i,mv,3) = 1.d0
        evr(i,4,3) = v

        evr(i,1,4) = 1.d0
        evr(i,mu,4) = u + a
        evr(i,mv,4) = v
        evr(i,4,4) = enth + u*a

        ! left eigenvectors (rows): L = R^{-1}, b = gamma-1, c^2 = a2
        b = gamma1
        h2 = 0.5d0/a2
        bk = 0.5d0*b*u2v2
        evl(i,1,1) = (bk + u*a)*h2
        evl(i,1,mu) = -(b*u + a)*h2
        evl(i,1,mv) = -b*v*h2
        evl(i,1,4) = b*h2

        evl(i,2,1) = 1.d0 - bk/a2
        evl(i,2,mu) = b*u/a2
        evl(i,2,mv) = b*v/a2
        evl(i,2,4) = -b/a2

        evl(i,3,1) = -v
        evl(i,3,mu) = 0.d0
        evl(i,3,mv) = 1.d0
        evl(i,3,4) = 0.d0

        evl(i,4,1) = (bk - u*a)*h2
        evl(i,4,mu) = -(b*u - a)*h2
        evl(i,4,mv) = -b*v*h2
        evl(i,4,4) = b*h2
    enddo
    return
end subroutine evec
