! Project-written eigenvector hook for the 1D Euler equations.
!
! The reference's apps/euler/shockentropy/evec1eu.f does not compile
! (duplicate label 20, undefined H, evl never filled).  This hook keeps
! its right eigenvectors (evec1eu.f:49-57) and its Roe averages (the same
! formulas as apps/euler/shockentropy/rp1eu.f:44-55, evaluated on the cell
! averages q(i-1), q(i)), and fills evl with the closed-form inverse of R.
! Argument list as called by src/flux1.f90:105.
subroutine evec(maxm,meqn,mbc,mx,q,auxl,auxr,evl,evr)
    implicit double precision (a-h,o-z)
    dimension q(1-mbc:maxm+mbc, meqn)
    dimension auxl(*), auxr(*)
    dimension evl(1-mbc:maxm+mbc, meqn, meqn)
    dimension evr(1-mbc:maxm+mbc, meqn, meqn)
    common /param/ gamma, gamma1

    do i = 2-mbc, mx+mbc
        rhsqrtl = dsqrt(q(i-1,1))
        rhsqrtr = dsqrt(q(i,1))
        pl = gamma1*(q(i-1,3) - 0.5d0*(q(i-1,2)**2)/q(i-1,1))
        pr = gamma1*(q(i,3) - 0.5d0*(q(i,2)**2)/q(i,1))
        rhsq2 = rhsqrtl + rhsqrtr
        u = (q(i-1,2)/rhsqrtl + q(i,2)/rhsqrtr) / rhsq2
        enth = ((q(i-1,3)+pl)/rhsqrtl + (q(i,3)+pr)/rhsqrtr) / rhsq2
        a2 = gamma1*(enth - 0.5d0*u**2)
        a = dsqrt(a2)

        evr(i,1,1) = 1.d0
        evr(i,1,2) = 1.d0
        evr(i,1,3) = 1.d0
        evr(i,2,1) = u - a
        evr(i,2,2) = u
        evr(i,2,3) = u + a
        evr(i,3,1) = enth - u*a
        evr(i,3,2) = 0.5d0*u**2
        evr(i,3,3) = enth + u*a

        ! L = R^{-1} with b = gamma-1, c^2 = a2
        b = gamma1
        h2 = 0.5d0/a2
        bk = 0.5d0*b*u**2
        evl(i,1,1) = (bk + u*a)*h2
        evl(i,1,2) = -(b*u + a)*h2
        evl(i,1,3) = b*h2
        evl(i,2,1) = 1.d0 - bk/a2
        evl(i,2,2) = b*u/a2
        evl(i,2,3) = -b/a2
        evl(i,3,1) = (bk - u*a)*h2
        evl(i,3,2) = -(b*u - a)*h2
        evl(i,3,3) = b*h2
    enddo
    return
end subroutine evec
