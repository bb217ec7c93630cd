! Fortran host driver for the device path: the control flow of the
! reference program (src/main.f90:21-270) and of its time loop
! (src/sharpclaw.f90:35-235), with every array operation of the hot path
! -- qold = q, the SSP stages, q = qold -- done on the device through the
! C ABI (module sharpclaw_gpu).  The user's hooks setprob, qinit, setaux,
! b4step, bc and src are linked in exactly as for the reference (an
! application's object replaces the default of libschooks<ndim>.a), the
! problem parameters sit in module ClawParams under the reference's names,
! sharpclaw.data is read with the same positional layout, and frames are
! written as fort.qNNNN / fort.tNNNN / fort.info in the reference's formats.
!
! Device options live in gpu.data (so sharpclaw.data stays as the reference
! reads it), one value per line:  physics id (1 advection, 2 acoustics,
! 3 Euler, 4 shallow water), precision (0 float64, 1 float32), device
! ordinal, efix (-1 default), and optionally
!   - a device source id (0 none; 1 the radial shallow-water source of
!     1dRadialSymm-qwave/src1.f90, computed inside the device stage);
!   - the period of the swirl app's b4step velocity factor (2D advection:
!     the device applies that factor itself; 0 = aux as uploaded);
!   - hooks: 0 (default) the device holds q for the whole run and b4step
!     sees the host q of the last frame; 1 live b4step: q is fetched from
!     the device before b4step and q and aux are sent back after it, so
!     b4step reads and changes the current solution as in the reference
!     (sharpclaw.f90:129); 2 host step: q stays in host memory, every
!     stage calls the application's bc and src on it (step1d.f90:20-29,
!     step2d.f90:25-36) and the device computes flux1/flux2 through
!     sc_flux, with the SSP combinations of ts_ssp* on the host -- the mode
!     for user boundary conditions (mthbc 0) and user source terms, which
!     the fused device stages cannot call.  Any mthbc of 0 selects mode 2.
! The physics constants are taken from the common blocks the user's setprob
! fills (/param/ gamma,gamma1; /comrp/ u or g; /cparam/ rho,bulk,cc,zz).
module xsclaw_data
    use iso_c_binding
    use ClawParams
    implicit none
    integer :: nout, outstyle, nstepout, nstop, max_steps, dt_variable
    integer :: time_integrator, verbosity
    double precision :: tfinal, dt_initial, dt_max, cfl_max, cfl_desired, t0
    double precision, allocatable :: tout(:)
    double precision, allocatable :: q(:), aux(:)
    ! host step (hooks mode 2): qold and the SSP registers of
    ! sharpclaw.f90:68-111 in host memory
    double precision, allocatable :: qold(:), r1(:), r2(:)
    integer :: hooks
contains
    ! src/opendatafile.f:45-53: skip the '#' header and the first line after it
    subroutine open_data_file(iunit, fname)
        integer, intent(in) :: iunit
        character(len=*), intent(in) :: fname
        character(len=1) :: c
        open(unit=iunit, file=fname, status='old', form='formatted')
        c = '#'
        do while (c == '#')
            read(iunit, '(a1)') c
        enddo
    end subroutine open_data_file
end module xsclaw_data

program xsclaw_gpu
    use iso_c_binding
    use sharpclaw_gpu
    use xsclaw_data
    implicit none
    type(sc_params) :: prm
    type(sc_run_ctl) :: ctl
    type(c_ptr) :: solver
    integer :: md, i, n, iframe, info, physics, precision, device, efix, source, ios
    double precision :: tperiod
    integer :: steps_taken
    double precision :: tstart, tend, dtout
    double precision :: cfl_maxused, cfl_last, dt_minused, dt_maxused, dt_last
    double precision :: gamma_c, gamma1_c, comrp_v, rho_c, bulk_c, cc_c, zz_c
    common /param/ gamma_c, gamma1_c
    common /comrp/ comrp_v
    common /cparam/ rho_c, bulk_c, cc_c, zz_c

    call open_data_file(55, 'sharpclaw.data')
    read(55,*) ndim
    allocate(nx(ndim))
    do md = 1, ndim
        read(55,*) nx(md)
    enddo
    maxnx = maxval(nx)
    read(55,*) nout
    read(55,*) outstyle
    select case (outstyle)
    case (1)
        read(55,*) tfinal
        nstepout = 1
    case (2)
        allocate(tout(nout))
        read(55,*) (tout(i), i=1,nout)
        nstepout = 1
    case (3)
        read(55,*) nstepout, nstop
        nout = nstop
    end select
    read(55,*) dt_initial
    read(55,*) dt_max
    read(55,*) cfl_max
    read(55,*) cfl_desired
    read(55,*) max_steps
    read(55,*) dt_variable
    read(55,*) time_integrator
    read(55,*) verbosity
    read(55,*) src_term
    read(55,*) mcapa
    read(55,*) maux
    read(55,*) tfluct_solver
    read(55,*) char_decomp
    read(55,*) meqn
    read(55,*) mwaves
    read(55,*) lim_type
    allocate(mthlim(mwaves))
    read(55,*) (mthlim(i), i=1,mwaves)
    read(55,*) t0
    allocate(xlower(ndim), xupper(ndim), dx(ndim))
    do md = 1, ndim
        read(55,*) xlower(md)
        read(55,*) xupper(md)
    enddo
    read(55,*) mbc
    allocate(mthbc(2*ndim))
    do md = 1, 2*ndim
        read(55,*) mthbc(md)
    enddo
    close(55)
    multid_recon = 0
    rord = 0

    if ((mthbc(1) == 2 .and. mthbc(2) /= 2) .or. (mthbc(2) == 2 .and. mthbc(1) /= 2)) then
        write(6,*) '*** ERROR ***  periodic boundary conditions'
        write(6,*) ' require mthbc(1) and mthbc(2) BOTH be set to 2'
        stop 1
    endif
    do i = 1, ndim
        dx(i) = (xupper(i) - xlower(i)) / float(nx(i))
    enddo
    if (outstyle == 1) dtout = (tfinal - t0) / float(nout)

    open(unit=56, file='gpu.data', status='old', form='formatted')
    read(56,*) physics
    read(56,*) precision
    read(56,*) device
    read(56,*) efix
    source = 0
    tperiod = 0.d0
    hooks = 0
    read(56,*,iostat=ios) source
    if (ios /= 0) source = 0
    if (ios == 0) read(56,*,iostat=ios) tperiod
    if (ios /= 0) tperiod = 0.d0
    if (ios == 0) read(56,*,iostat=ios) hooks
    if (ios /= 0) hooks = 0
    close(56)
    if (any(mthbc == 0)) hooks = 2
    if (hooks < 0 .or. hooks > 2) then
        write(6,*) '*** ERROR *** gpu.data: hooks must be 0, 1 or 2'
        stop 1
    endif

    call setprob

    if (ndim == 1) then
        allocate(q((nx(1)+2*mbc)*meqn))
        allocate(aux(max(1,(nx(1)+2*mbc)*maux)))
    else
        allocate(q((nx(1)+2*mbc)*(nx(2)+2*mbc)*meqn))
        allocate(aux(max(1,(nx(1)+2*mbc)*(nx(2)+2*mbc)*maux)))
    endif
    q = 0.d0
    aux = 0.d0
    if (maux > 0) call setaux(maxnx, mbc, nx, xlower, dx, maux, aux)
    call qinit(ndim, meqn, mbc, nx, xlower, dx, q, maux, aux)

    prm%ndim = ndim
    prm%nx = 1
    prm%nx(1:ndim) = nx
    prm%meqn = meqn
    prm%mwaves = mwaves
    prm%mbc = mbc
    prm%maux = maux
    prm%mcapa = mcapa
    prm%lim_type = lim_type
    prm%char_decomp = char_decomp
    prm%tfluct_solver = tfluct_solver
    prm%time_integrator = time_integrator
    prm%src_term = source
    prm%mthbc = 1
    prm%mthbc(1:2*ndim) = mthbc
    prm%xlower = 0.d0
    prm%xlower(1:ndim) = xlower
    prm%dx = 1.d0
    prm%dx(1:ndim) = dx
    prm%physics = physics
    prm%efix = efix
    prm%gamma = gamma_c
    prm%grav = comrp_v
    prm%u = comrp_v
    prm%rho = rho_c
    prm%bulk = bulk_c
    prm%precision = precision
    prm%device = device
    prm%tperiod = tperiod
    prm%mthlim = 0
    prm%mthlim(1:min(4, mwaves)) = mthlim(1:min(4, mwaves))
    call sc_check(sc_create(prm, solver), 'sc_create')
    if (maux > 0) call sc_check(sc_set_aux(solver, aux), 'sc_set_aux')
    call sc_check(sc_set_q(solver, q), 'sc_set_q')

    ctl%dt_max = dt_max
    ctl%cfl_max = cfl_max
    ctl%cfl_desired = cfl_desired
    ctl%max_steps = max_steps
    ctl%dt_variable = dt_variable

    ! frames go through sc_output_begin/wait: the |q| < 1e-99 flush of the
    ! reference's writers (out1.f90:63-68) then happens in the device q the
    ! run continues from, as it does in the reference's q
    if (hooks == 2) then
        allocate(qold(size(q)), r1(size(q)))
        qold = q
        r1 = 0.d0
        if (time_integrator == 4) then
            allocate(r2(size(q)))
            r2 = 0.d0
        endif
    endif
    call get_frame
    call write_frame(t0, 0)
    write(6,601) 0, t0
    open(10, file='fort.info', status='unknown', form='formatted')
    tend = t0
    do n = 1, nout
        tstart = tend
        select case (outstyle)
        case (1)
            tend = tstart + dtout
        case (2)
            tend = tout(n)
        case (3)
            tend = tstart - 1.d0
        end select
        call gpu_sharpclaw(solver, ctl, tstart, tend, cfl_maxused, cfl_last, dt_minused, &
                           dt_maxused, dt_last, steps_taken, info)
        if (info /= 0) then
            write(6,*) '*** ERROR in sharpclaw ***  info =', info
            if (info == 3) write(6,*) '***   dt_variable=1 and cfl_desired > cfl_max'
            if (info == 11) write(6,*) '***   Too many times steps, n > max_steps'
            if (info == 12) write(6,*) '***   The Courant number is greater than cfl_max'
        endif
        dt_initial = dt_last
        iframe = n / nstepout
        if (iframe*nstepout == n) then
            call get_frame
            call write_frame(tend, iframe)
            write(6,601) iframe, tend
            write(10,1010) tend, info, dt_minused, dt_maxused, dt_last, cfl_maxused, &
                           cfl_last, steps_taken
        endif
    enddo
601 format('main: Frame ',i4,' matlab plot files done at time t =',d12.4,/)
1010 format('tend =',d15.4,/,'info =',i5,/,'smallest dt =',d15.4,/,'largest dt =', &
            d15.4,/,'last dt =',d15.4,/,'largest cfl =',d15.4,/,'last cfl =',d15.4,/, &
            'steps taken =',i4,/)
    close(10)
    call sc_check(sc_destroy(solver), 'sc_destroy')

contains

    ! q of a frame into the host array.  On the device (hooks 0, 1) it goes
    ! through sc_output_begin/wait: the |q| < 1e-99 flush of the reference's
    ! writers (out1.f90:63-68) then happens in the device q the run
    ! continues from, as it does in the reference's q; in host-step mode q
    ! is already here and write_frame flushes it.
    subroutine get_frame
        if (hooks == 2) return
        call sc_check(sc_output_begin(solver), 'sc_output_begin')
        call sc_check(sc_output_wait(solver, q), 'sc_output_wait')
    end subroutine get_frame

    ! The user's b4step before a new step (sharpclaw.f90:129).  With live
    ! hooks it works on the current device solution: q (ghost cells filled
    ! by the boundary rules) is fetched first, and q and aux are sent back.
    subroutine user_b4step(told, dt)
        double precision, intent(in) :: told, dt
        if (hooks == 1) call sc_check(sc_get_q(solver, q), 'sc_get_q')
        call b4step(maxnx, mbc, nx, meqn, q, xlower, dx, told, dt, maux, aux)
        if (hooks == 1) call sc_check(sc_set_q(solver, q), 'sc_set_q')
        if (hooks >= 1 .and. maux > 0) call sc_check(sc_set_aux(solver, aux), 'sc_set_aux')
    end subroutine user_b4step

    ! step(q,g,dq,aux,dt,cfl,t,rp,src,tfluct) of src/1d/step1d.f90:2-31 and
    ! src/2d/step2d.f90:2-41 on host arrays: the application's bc fills the
    ! ghost cells, its src sets dq, and the device adds dt*L(q) (sc_flux).
    subroutine host_step(qs, dqs, t, dt, cfl)
        double precision, intent(inout) :: qs(:), dqs(:)
        double precision, intent(in) :: t, dt
        double precision, intent(out) :: cfl
        if (ndim == 1) then
            call host_bc1(qs, aux, t, dt)
        else
            call host_bc2(qs, aux, t, dt)
        endif
        call src(qs, dqs, aux, t, dt)
        call sc_check(sc_flux(solver, qs, dqs, dt, t, cfl), 'sc_flux')
    end subroutine host_step

    ! One SSP step on the host arrays, with the stage times and combinations
    ! of ts_ssp22 / ts_ssp42 / ts_ssp33 / ts_ssp104 (sharpclaw.f90:247-464);
    ! cfl is the last stage's.
    subroutine host_ts(dt, told, cfl)
        double precision, intent(in) :: dt, told
        double precision, intent(out) :: cfl
        double precision :: third, sixth
        select case (time_integrator)
        case (1)
            call host_step(q, r1, told, dt, cfl)
            q = q + r1
            call host_step(q, r1, told + dt, dt, cfl)
            q = 0.5d0*(qold + q + r1)
        case (2)
            third = 1.d0/3.d0
            call host_step(q, r1, told, dt, cfl)
            q = q + third*r1
            call host_step(q, r1, told + third*dt, dt, cfl)
            q = q + third*r1
            call host_step(q, r1, told + 2.d0*third*dt, dt, cfl)
            q = q + third*r1
            call host_step(q, r1, told + dt, dt, cfl)
            q = 0.25d0*qold + 0.75d0*q + 0.25d0*r1
        case (3)
            r1 = 0.d0
            call host_step(q, r1, told, dt, cfl)
            q = q + r1
            call host_step(q, r1, told + dt, dt, cfl)
            q = 0.75d0*qold + 0.25d0*(q + r1)
            call host_step(q, r1, told + 0.5d0*dt, dt, cfl)
            q = (qold + 2.d0*(q + r1))/3.d0
        case default
            sixth = 1.d0/6.d0
            call host_step(q, r1, told, dt, cfl)
            q = q + sixth*r1
            call host_step(q, r1, told + dt/6.d0, dt, cfl)
            q = q + sixth*r1
            call host_step(q, r1, told + dt/3.d0, dt, cfl)
            q = q + sixth*r1
            call host_step(q, r1, told + 0.5d0*dt, dt, cfl)
            q = q + sixth*r1
            call host_step(q, r1, told + 2.d0/3.d0*dt, dt, cfl)
            r2 = q + sixth*r1
            r1 = 0.04d0*qold + 0.36d0*r2
            q = 0.6d0*qold + 0.4d0*r2
            call host_step(q, r2, told + 1.d0/3.d0*dt, dt, cfl)
            q = q + sixth*r2
            call host_step(q, r2, told + 0.5d0*dt, dt, cfl)
            q = q + sixth*r2
            call host_step(q, r2, told + 2.d0/3.d0*dt, dt, cfl)
            q = q + sixth*r2
            call host_step(q, r2, told + 5.d0/6.d0*dt, dt, cfl)
            q = q + sixth*r2
            call host_step(q, r2, told + dt, dt, cfl)
            q = r1 + 0.6d0*q + 0.1d0*r2
        end select
    end subroutine host_ts

    ! The time loop of one output interval (sharpclaw.f90:35-235).  qold is
    ! saved at every new step (a buffer renaming on the device); the SSP step
    ! runs as fused device stages (sc_ts), or stage by stage through the
    ! application's bc / src and sc_flux in host-step mode; the CFL of its
    ! last stage drives dt.
    subroutine gpu_sharpclaw(solver, ctl, tstart, tend, cfl_maxused, cfl_last, dt_minused, &
                             dt_maxused, dt_last, steps_taken, info)
        type(c_ptr), intent(in) :: solver
        type(sc_run_ctl), intent(in) :: ctl
        double precision, intent(in) :: tstart
        double precision, intent(inout) :: tend
        double precision, intent(out) :: cfl_maxused, cfl_last, dt_minused, dt_maxused, dt_last
        integer, intent(out) :: steps_taken, info
        double precision :: t, told, dt, cfl, cflmax, dtmin, dtmax
        integer :: maxn, it
        logical :: retake
        info = 0
        t = tstart
        told = tstart
        maxn = ctl%max_steps
        dt = dt_initial
        cflmax = 0.d0
        cfl = 0.d0
        dtmin = dt
        dtmax = dt
        steps_taken = 0
        if (ctl%dt_variable == 0) then
            if (tend < tstart) then
                maxn = 1
            else
                maxn = int((tend - tstart + 1d-10) / dt)
                if (dabs(maxn*dt - (tend-tstart)) > 1d-5*(tend-tstart)) then
                    info = 2
                    goto 900
                endif
            endif
        endif
        if (ctl%dt_variable == 1 .and. ctl%cfl_desired > ctl%cfl_max) then
            info = 3
            goto 900
        endif
        retake = .false.
        do it = 1, maxn
            if (.not. retake) then
                told = t
                if (told + dt > tend .and. tstart < tend) dt = tend - told
                call user_b4step(told, dt)
                if (hooks == 2) then
                    qold = q
                else
                    call sc_check(sc_save_qold(solver), 'sc_save_qold')
                endif
            endif
            retake = .false.
            t = told + dt
            if (hooks == 2) then
                call host_ts(dt, told, cfl)
            else
                call sc_check(sc_ts(solver, dt, told, cfl), 'sc_ts')
            endif
            if (verbosity == 1) write(6,602) it, cfl, dt, t
            if (ctl%dt_variable == 1) then
                if (cfl > 0.d0) then
                    dt = dmin1(ctl%dt_max, dt * ctl%cfl_desired / cfl)
                    dtmin = dmin1(dt, dtmin)
                    dtmax = dmax1(dt, dtmax)
                else
                    dt = ctl%dt_max
                endif
            endif
            if (cfl <= ctl%cfl_max) then
                cflmax = dmax1(cfl, cflmax)
            else
                t = told
                if (hooks == 2) then
                    q = qold
                else
                    call sc_check(sc_restore_qold(solver), 'sc_restore_qold')
                endif
                if (verbosity == 1) write(6,*) 'SClaw rejecting step... Courant number too large'
                if (ctl%dt_variable == 1) then
                    retake = .true.
                else
                    cflmax = dmax1(cfl, cflmax)
                    exit
                endif
            endif
            steps_taken = steps_taken + 1
            if (t >= tend) exit
        enddo
        if (ctl%dt_variable == 1 .and. t < tend .and. steps_taken == maxn) info = 11
        if (ctl%dt_variable == 0 .and. cflmax > ctl%cfl_max) info = 12
900     continue
        tend = t
        cfl_maxused = cflmax
        cfl_last = cfl
        dt_minused = dtmin
        dt_maxused = dtmax
        dt_last = dt
602     format('SClaw... Step',i4,'   Courant number =',f6.3,'  dt =',d12.4,'  t =',d12.4)
    end subroutine gpu_sharpclaw

    ! fort.qNNNN / fort.tNNNN (and fort.aNNNN in 2D) in the layouts of
    ! src/1d/out1.f90:32-104 and src/2d/out2.f90:26-111, including the flush
    ! of |q| < 1e-99 to zero.
    subroutine write_frame(t, iframe)
        double precision, intent(in) :: t
        integer, intent(in) :: iframe
        character(len=10) :: f1, f2, f3
        integer :: ii, jj, m, lx, ly, k
        write(f1,'(a6,i4.4)') 'fort.q', iframe
        write(f2,'(a6,i4.4)') 'fort.t', iframe
        write(f3,'(a6,i4.4)') 'fort.a', iframe
        open(unit=50, file=f1, status='unknown', form='formatted')
        open(unit=60, file=f2, status='unknown', form='formatted')
        lx = nx(1) + 2*mbc
        if (ndim == 1) then
            write(50,1001) 1, 1, nx(1)
            write(50,1002) xlower(1), dx(1)
            do ii = 1, nx(1)
                do m = 1, meqn
                    k = (ii + mbc) + (m-1)*lx
                    if (dabs(q(k)) < 1d-99) q(k) = 0.d0
                enddo
                write(50,1005) (q((ii + mbc) + (m-1)*lx), m=1,meqn)
            enddo
            write(50,*) ' '
            write(50,*) ' '
            write(60,1000) t, meqn, 1, maux, 1
        else
            ly = nx(2) + 2*mbc
            write(50,2001) 1, 1, nx(1), nx(2)
            write(50,2002) xlower(1), xlower(2), dx(1), dx(2)
            do jj = 1, nx(2)
                do ii = 1, nx(1)
                    do m = 1, meqn
                        k = (ii + mbc) + lx*((jj + mbc - 1) + ly*(m-1))
                        if (dabs(q(k)) < 1d-99) q(k) = 0.d0
                    enddo
                    write(50,2005) (q((ii + mbc) + lx*((jj + mbc - 1) + ly*(m-1))), m=1,meqn)
                enddo
                write(50,*) ' '
            enddo
            write(50,*) ' '
            ! out2.f90:17 writes the aux file too (outaux = .true.)
            open(unit=70, file=f3, status='unknown', form='formatted')
            write(70,2001) 1, 1, nx(1), nx(2)
            write(70,2002) xlower(1), xlower(2), dx(1), dx(2)
            do jj = 1, nx(2)
                do ii = 1, nx(1)
                    write(70,2005) (aux((ii + mbc) + lx*((jj + mbc - 1) + ly*(m-1))), m=1,maux)
                enddo
                write(70,*) ' '
            enddo
            write(70,*) ' '
            close(70)
            write(60,1000) t, meqn, 1, maux, 2
        endif
        close(50)
        close(60)
1001    format(i5,'                 grid_number',/,i5,'                 AMR_level',/, &
               i5,'                 mx')
1002    format(e18.8,'    xlow', /, e18.8,'    dx', /)
1005    format(4e16.8)
1000    format(e26.16,'    time', /, i5,'                 meqn'/, i5,'                 ngrids'/, &
               i5,'                 maux'/, i5,'                 ndim'/,/)
2001    format(i5,'                 grid_number',/,i5,'                 AMR_level',/, &
               i5,'                 mx',/,i5,'                 my')
2002    format(e26.16,'    xlow', /, e26.16,'    ylow', /, e26.16,'    dx', /, &
               e26.16,'    dy',/)
2005    format(4e26.16)
    end subroutine write_frame

end program xsclaw_gpu
