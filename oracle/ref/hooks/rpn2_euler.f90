! Project-written Roe solver (rpn2) for the 2D Euler equations; the
! reference ships none.  Structure: normal/tangential momentum selection
! as apps/shallowWater/2dRadialDam/rpn2sw.f:60-66; Roe averages, wave
! strengths and the entropy fix as apps/euler/shockentropy/rp1eu.f:44-213
! extended by the tangential momentum.  mwaves = 3: the entropy and
! shear jumps both travel at speed u and share wave 2.
! Argument list of the reference rp plugins (rp1eu.f:5-6, rpn2sw.f:5-6).
subroutine rp(ixy,maxm,meqn,mwaves,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
    implicit double precision (a-h,o-z)
    dimension   ql(1-mbc:maxm+mbc, meqn)
    dimension   qr(1-mbc:maxm+mbc, meqn)
    dimension    s(1-mbc:maxm+mbc, mwaves)
    dimension wave(1-mbc:maxm+mbc, meqn, mwaves)
    dimension amdq(1-mbc:maxm+mbc, meqn)
    dimension apdq(1-mbc:maxm+mbc, meqn)
    dimension auxl(*), auxr(*)
    dimension delta(4)
    common /param/ gamma, gamma1
    integer :: mu, mv

    if (ixy == 1) then
        mu = 2
        mv = 3
    else
        mu = 3
        mv = 2
    endif

    do i = 2-mbc, mx+mbc
        rhsqrtl = dsqrt(qr(i-1,1))
        rhsqrtr = dsqrt(ql(i,1))
        pl = gamma1*(qr(i-1,4) - 0.5d0*(qr(i-1,mu)**2 + qr(i-1,mv)**2)/qr(i-1,1))
        pr = gamma1*(ql(i,4) - 0.5d0*(ql(i,mu)**2 + ql(i,mv)**2)/ql(i,1))
        rhsq2 = rhsqrtl + rhsqrtr
        u = (qr(i-1,mu)/rhsqrtl + ql(i,mu)/rhsqrtr) / rhsq2
        v = (qr(i-1,mv)/rhsqrtl + ql(i,mv)/rhsqrtr) / rhsq2
        enth = ((qr(i-1,4)+pl)/rhsqrtl + (ql(i,4)+pr)/rhsqrtr) / rhsq2
        u2v2 = u**2 + v**2
        a2 = gamma1*(enth - 0.5d0*u2v2)
        a = dsqrt(a2)

        delta(1) = ql(i,1) - qr(i-1,1)
        delta(2) = ql(i,mu) - qr(i-1,mu)
        delta(3) = ql(i,mv) - qr(i-1,mv)
        delta(4) = ql(i,4) - qr(i-1,4)
        a3 = delta(3) - v*delta(1)
        a2s = gamma1/a2 * ((enth - u2v2)*delta(1) + u*delta(2) &
              + v*delta(3) - delta(4))
        a4 = (delta(2) + (a - u)*delta(1) - a*a2s) / (2.d0*a)
        a1 = delta(1) - a2s - a4

        wave(i,1,1) = a1
        wave(i,mu,1) = a1*(u - a)
        wave(i,mv,1) = a1*v
        wave(i,4,1) = a1*(enth - u*a)
        s(i,1) = u - a

        wave(i,1,2) = a2s
        wave(i,mu,2) = a2s*u
        wave(i,mv,2) = a2s*v + a3
        wave(i,4,2) = a2s*0.5d0*u2v2 + a3*v
        s(i,2) = u

        wave(i,1,3) = a4
        wave(i,mu,3) = a4*(u + a)
        wave(i,mv,3) = a4*v
        wave(i,4,3) = a4*(enth + u*a)
        s(i,3) = u + a
    enddo

    ! entropy fix (transonic rarefactions), as rp1eu.f:126-201
    do i = 2-mbc, mx+mbc
        rhoim1 = qr(i-1,1)
        pim1 = gamma1*(qr(i-1,4) - 0.5d0*(qr(i-1,mu)**2 + qr(i-1,mv)**2)/rhoim1)
        cim1 = dsqrt(gamma*pim1/rhoim1)
        s0 = qr(i-1,mu)/rhoim1 - cim1

        if (s0 >= 0.d0 .and. s(i,1) > 0.d0) then
            do m = 1, meqn
                amdq(i,m) = 0.d0
            enddo
            cycle
        endif

        rho1 = qr(i-1,1) + wave(i,1,1)
        rhou1 = qr(i-1,mu) + wave(i,mu,1)
        rhov1 = qr(i-1,mv) + wave(i,mv,1)
        en1 = qr(i-1,4) + wave(i,4,1)
        p1 = gamma1*(en1 - 0.5d0*(rhou1**2 + rhov1**2)/rho1)
        c1 = dsqrt(gamma*p1/rho1)
        s1 = rhou1/rho1 - c1
        if (s0 < 0.d0 .and. s1 > 0.d0) then
            sfract = s0 * (s1 - s(i,1)) / (s1 - s0)
        else if (s(i,1) < 0.d0) then
            sfract = s(i,1)
        else
            sfract = 0.d0
        endif
        do m = 1, meqn
            amdq(i,m) = sfract*wave(i,m,1)
        enddo

        if (s(i,2) >= 0.d0) cycle
        do m = 1, meqn
            amdq(i,m) = amdq(i,m) + s(i,2)*wave(i,m,2)
        enddo

        rhoi = ql(i,1)
        pi = gamma1*(ql(i,4) - 0.5d0*(ql(i,mu)**2 + ql(i,mv)**2)/rhoi)
        ci = dsqrt(gamma*pi/rhoi)
        s3 = ql(i,mu)/rhoi + ci

        rho2 = ql(i,1) - wave(i,1,3)
        rhou2 = ql(i,mu) - wave(i,mu,3)
        rhov2 = ql(i,mv) - wave(i,mv,3)
        en2 = ql(i,4) - wave(i,4,3)
        p2 = gamma1*(en2 - 0.5d0*(rhou2**2 + rhov2**2)/rho2)
        c2 = dsqrt(gamma*p2/rho2)
        s2 = rhou2/rho2 + c2
        if (s2 < 0.d0 .and. s3 > 0.d0) then
            sfract = s2 * (s3 - s(i,3)) / (s3 - s2)
        else if (s(i,3) < 0.d0) then
            sfract = s(i,3)
        else
            cycle
        endif
        do m = 1, meqn
            amdq(i,m) = amdq(i,m) + sfract*wave(i,m,3)
        enddo
    enddo

    do m = 1, meqn
        do i = 2-mbc, mx+mbc
            df = 0.d0
            do mw = 1, mwaves
                df = df + s(i,mw)*wave(i,m,mw)
            enddo
            apdq(i,m) = df - amdq(i,m)
        enddo
    enddo
    return
end subroutine rp
