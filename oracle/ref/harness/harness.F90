! Oracle harness: calls single routines of the reference library, compiled
! unmodified from its sources, on binary inputs so that the C restatement
! and the device kernels can be compared with the reference itself.
!
!   harness rp    in out   -> rp(ixy,...)                    (app rp plugin)
!   harness evec  in out   -> evec(mx,meqn,mbc,mx,q,...)      (flux1.f90:105)
!   harness weno  in out   -> weno5 / weno5_char              (reconstruct.f90:77,146)
!   harness step  in out   -> step(q,g,dq,aux,dt,cfl,t,...)   (step1d.f90 / step2d.f90)
!
! setprob is called first, so setprob.data must be in the working
! directory.  All files are raw little-endian streams (int32, float64).
program harness
    use ClawData
    use ClawParams
    use reconstruct
    use evec_dir
    implicit none
    external rp, src, tfluct
    character(len=16) :: mode
    character(len=256) :: fin, fout
    integer :: ixy, mx, mq, mw, nb, mx2, charf, i
    integer :: hdr(18)
    double precision :: rhdr(6)
    double precision, allocatable :: ql(:,:), qr(:,:), wave(:,:,:), s(:,:)
    double precision, allocatable :: amdq(:,:), apdq(:,:), evl(:,:,:), evr(:,:,:)
    double precision, allocatable :: q1(:,:), aux1(:,:)
#if NDIM == 2
    double precision, allocatable :: qs(:,:,:), dqs(:,:,:), auxs(:,:,:)
#else
    double precision, allocatable :: qs(:,:), dqs(:,:), auxs(:,:)
#endif
    double precision :: dt, t, cfl
    type(griddat) :: g

    call get_command_argument(1, mode)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    call setprob

    open(unit=21, file=fin, access='stream', form='unformatted', status='old')
    open(unit=22, file=fout, access='stream', form='unformatted', status='replace')

    select case (trim(mode))
    case ('rp')
        read(21) ixy, mx, mq, mw, nb
        allocate(ql(1-nb:mx+nb, mq), qr(1-nb:mx+nb, mq))
        allocate(wave(1-nb:mx+nb, mq, mw), s(1-nb:mx+nb, mw))
        allocate(amdq(1-nb:mx+nb, mq), apdq(1-nb:mx+nb, mq), aux1(1-nb:mx+nb, 1))
        read(21) ql, qr
        wave = 0.d0; s = 0.d0; amdq = 0.d0; apdq = 0.d0; aux1 = 0.d0
        call rp(ixy, mx, mq, mw, nb, mx, ql, qr, aux1, aux1, wave, s, amdq, apdq)
        write(22) wave, s, amdq, apdq

    case ('evec')
        read(21) ixy, mx, mq, nb
        ixy_forced = ixy
        allocate(q1(1-nb:mx+nb, mq), aux1(1-nb:mx+nb, 1))
        allocate(evl(1-nb:mx+nb, mq, mq), evr(1-nb:mx+nb, mq, mq))
        read(21) q1
        evl = 0.d0; evr = 0.d0; aux1 = 0.d0
        call evec(mx, mq, nb, mx, q1, aux1, aux1, evl, evr)
        write(22) evl, evr

    case ('weno')
        read(21) mx2, mq, charf
        allocate(q1(mx2, mq), ql(mx2, mq), qr(mx2, mq))
        allocate(evl(mx2, mq, mq), evr(mx2, mq, mq))
        read(21) q1
        if (charf == 2) read(21) evl, evr
        call alloc_recon_workspace(mx2-6, 3, mq, mq, 2, charf)
        ql = 0.d0; qr = 0.d0
        if (charf == 2) then
            call weno5_char(q1, ql, qr, evl, evr)
        else
            call weno5(q1, ql, qr)
        endif
        write(22) ql, qr

    case ('step')
        ! header: ndim nx ny meqn mwaves char_decomp mthbc(4) tfluct_solver
        !         lim_type mthlim(4) maux mcapa; q, then aux when maux > 0
        read(21) hdr
        read(21) rhdr      ! dx dy xlower ylower dt t
        ndim = hdr(1)
        if (ndim /= NDIM) stop 'harness: ndim of input does not match this build'
        allocate(nx(ndim), dx(ndim), xlower(ndim), xupper(ndim), mthbc(2*ndim))
        nx(1) = hdr(2)
        if (ndim == 2) nx(2) = hdr(3)
        meqn = hdr(4); mwaves = hdr(5); char_decomp = hdr(6)
        do i = 1, 2*ndim
            mthbc(i) = hdr(6+i)
        enddo
        dx(1) = rhdr(1); xlower(1) = rhdr(3)
        if (ndim == 2) then
            dx(2) = rhdr(2); xlower(2) = rhdr(4)
        endif
        dt = rhdr(5); t = rhdr(6)
        mbc = 3; maux = hdr(17); mcapa = hdr(18); lim_type = hdr(12); tfluct_solver = hdr(11)
        src_term = 0; multid_recon = 0
        allocate(mthlim(mwaves))
        do i = 1, mwaves
            mthlim(i) = hdr(12 + min(i, 4))
        enddo
        maxnx = maxval(nx)
        call alloc_recon_workspace(maxnx, mbc, meqn, mwaves, lim_type, char_decomp)
        allocate(g%s(1-mbc:maxnx+mbc, mwaves), g%wave(1-mbc:maxnx+mbc, meqn, mwaves))
        allocate(g%ql(1-mbc:maxnx+mbc, meqn), g%qr(1-mbc:maxnx+mbc, meqn))
        allocate(g%dtdx(1-mbc:maxnx+mbc))
        allocate(g%amdq(1-mbc:maxnx+mbc, meqn), g%apdq(1-mbc:maxnx+mbc, meqn))
        allocate(g%amdq2(1-mbc:maxnx+mbc, meqn), g%apdq2(1-mbc:maxnx+mbc, meqn))
        allocate(g%evl(1-mbc:maxnx+mbc, meqn, meqn), g%evr(1-mbc:maxnx+mbc, meqn, meqn))
        allocate(g%auxl(1-mbc:maxnx+mbc, maux), g%auxr(1-mbc:maxnx+mbc, maux))
#if NDIM == 2
        allocate(g%q1d(1-mbc:maxnx+mbc, meqn), g%dq1d(1-mbc:maxnx+mbc, meqn))
        allocate(g%aux2(1-mbc:maxnx+mbc, maux))
        allocate(qs(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, meqn))
        allocate(dqs(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, meqn))
        allocate(auxs(1-mbc:nx(1)+mbc, 1-mbc:nx(2)+mbc, max(1, maux)))
#else
        allocate(qs(1-mbc:nx(1)+mbc, meqn), dqs(1-mbc:nx(1)+mbc, meqn))
        allocate(auxs(1-mbc:nx(1)+mbc, max(1, maux)))
#endif
        read(21) qs
        dqs = 0.d0; auxs = 0.d0
        if (maux > 0) read(21) auxs
        call step(qs, g, dqs, auxs, dt, cfl, t, rp, src, tfluct)
        write(22) cfl
        write(22) qs, dqs

    case default
        stop 'harness: unknown mode'
    end select
    close(21)
    close(22)
end program harness
