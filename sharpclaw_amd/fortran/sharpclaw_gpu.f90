! ISO_C_BINDING interface to libsharpclaw_amd.so (include/sharpclaw_amd.h).
!
! This is the thin layer a Fortran SharpClaw driver uses: the parameter
! block mirrors module ClawParams (src/ClawParams.f90:4-14) plus the physics
! common blocks, the solver handle owns the device copies of q / qold / the
! RK registers of module Global (src/1d/Global1.f90), and every array is the
! caller's own Fortran array q(1-mbc:nx+mbc[,1-mbc:ny+mbc], meqn).
module sharpclaw_gpu
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: SC_OK = 0, SC_ERR_ARG = 1, SC_ERR_UNSUPPORTED = 2
    integer(c_int), parameter :: SC_ERR_HIP = 3, SC_ERR_NOMEM = 4, SC_ERR_STATE = 5
    integer(c_int), parameter :: SC_PHYS_ADVECTION = 1, SC_PHYS_ACOUSTICS = 2
    integer(c_int), parameter :: SC_PHYS_EULER = 3, SC_PHYS_SHALLOW_WATER = 4
    integer(c_int), parameter :: SC_FP64 = 0, SC_FP32 = 1

    type, bind(C) :: sc_params
        integer(c_int) :: ndim
        integer(c_int) :: nx(2)
        integer(c_int) :: meqn, mwaves, mbc, maux, mcapa
        integer(c_int) :: lim_type, char_decomp, tfluct_solver, time_integrator, src_term
        integer(c_int) :: mthbc(4)
        real(c_double) :: xlower(2), dx(2)
        integer(c_int) :: physics, efix
        real(c_double) :: gamma, grav, u, rho, bulk
        integer(c_int) :: precision, device
        integer(c_int) :: mthlim(4)
        real(c_double) :: tperiod
    end type sc_params

    type, bind(C) :: sc_run_ctl
        real(c_double) :: dt_max, cfl_max, cfl_desired
        integer(c_int) :: max_steps, dt_variable
    end type sc_run_ctl

    type, bind(C) :: sc_run_stats
        real(c_double) :: t, cfl_maxused, cfl_last, dt_minused, dt_maxused, dt_last
        integer(c_int) :: steps_taken, info, rejected
        integer(c_long_long) :: stage_evals
        real(c_double) :: gpu_ms
    end type sc_run_stats

    interface
        function sc_last_error_c() bind(C, name="sc_last_error")
            import :: c_ptr
            type(c_ptr) :: sc_last_error_c
        end function
        integer(c_int) function sc_create(params, solver) bind(C, name="sc_create")
            import :: c_int, c_ptr, sc_params
            type(sc_params), intent(in) :: params
            type(c_ptr), intent(out) :: solver
        end function
        integer(c_int) function sc_destroy(solver) bind(C, name="sc_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: solver
        end function
        integer(c_int) function sc_set_q(solver, q) bind(C, name="sc_set_q")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), intent(in) :: q(*)
        end function
        integer(c_int) function sc_get_q(solver, q) bind(C, name="sc_get_q")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), intent(inout) :: q(*)
        end function
        integer(c_int) function sc_set_aux(solver, aux) bind(C, name="sc_set_aux")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), intent(in) :: aux(*)
        end function
        integer(c_int) function sc_output_begin(solver) bind(C, name="sc_output_begin")
            import :: c_int, c_ptr
            type(c_ptr), value :: solver
        end function
        integer(c_int) function sc_output_wait(solver, q) bind(C, name="sc_output_wait")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), intent(inout) :: q(*)
        end function
        integer(c_int) function sc_save_qold(solver) bind(C, name="sc_save_qold")
            import :: c_int, c_ptr
            type(c_ptr), value :: solver
        end function
        integer(c_int) function sc_restore_qold(solver) bind(C, name="sc_restore_qold")
            import :: c_int, c_ptr
            type(c_ptr), value :: solver
        end function
        integer(c_int) function sc_step(solver, dt, t, dq, cfl) bind(C, name="sc_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), value :: dt, t
            real(c_double), intent(inout) :: dq(*)
            real(c_double), intent(out) :: cfl
        end function
        integer(c_int) function sc_flux(solver, q, dq, dt, t, cfl) bind(C, name="sc_flux")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), intent(in) :: q(*)
            real(c_double), intent(inout) :: dq(*)
            real(c_double), value :: dt, t
            real(c_double), intent(out) :: cfl
        end function
        integer(c_int) function sc_ts(solver, dt, told, cfl) bind(C, name="sc_ts")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), value :: dt, told
            real(c_double), intent(out) :: cfl
        end function
        integer(c_int) function sc_advance(solver, ctl, tstart, tend, dt, stats) &
                bind(C, name="sc_advance")
            import :: c_int, c_ptr, c_double, sc_run_ctl, sc_run_stats
            type(c_ptr), value :: solver
            type(sc_run_ctl), intent(in) :: ctl
            real(c_double), value :: tstart, tend
            real(c_double), intent(inout) :: dt
            type(sc_run_stats), intent(out) :: stats
        end function
        integer(c_int) function sc_sums(solver, sums) bind(C, name="sc_sums")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: solver
            real(c_double), intent(out) :: sums(*)
        end function
        integer(c_int) function sc_rp(params, ixy, mx, ql, qr, wave, s, amdq, apdq) &
                bind(C, name="sc_rp")
            import :: c_int, c_double, sc_params
            type(sc_params), intent(in) :: params
            integer(c_int), value :: ixy, mx
            real(c_double), intent(in) :: ql(*), qr(*)
            real(c_double), intent(inout) :: wave(*), s(*), amdq(*), apdq(*)
        end function
        integer(c_int) function sc_evec(params, ixy, mx, q, evl, evr) bind(C, name="sc_evec")
            import :: c_int, c_double, sc_params
            type(sc_params), intent(in) :: params
            integer(c_int), value :: ixy, mx
            real(c_double), intent(in) :: q(*)
            real(c_double), intent(inout) :: evl(*), evr(*)
        end function
        integer(c_int) function sc_weno5(params, mx2, q, ql, qr) bind(C, name="sc_weno5")
            import :: c_int, c_double, sc_params
            type(sc_params), intent(in) :: params
            integer(c_int), value :: mx2
            real(c_double), intent(in) :: q(*)
            real(c_double), intent(inout) :: ql(*), qr(*)
        end function
        integer(c_int) function sc_weno5_char(params, mx2, q, evl, evr, ql, qr) &
                bind(C, name="sc_weno5_char")
            import :: c_int, c_double, sc_params
            type(sc_params), intent(in) :: params
            integer(c_int), value :: mx2
            real(c_double), intent(in) :: q(*), evl(*), evr(*)
            real(c_double), intent(inout) :: ql(*), qr(*)
        end function
        integer(c_int) function sc_flux1(params, ixy, mx, q1d, dq1d, dt, cfl) &
                bind(C, name="sc_flux1")
            import :: c_int, c_double, sc_params
            type(sc_params), intent(in) :: params
            integer(c_int), value :: ixy, mx
            real(c_double), intent(in) :: q1d(*)
            real(c_double), intent(inout) :: dq1d(*)
            real(c_double), value :: dt
            real(c_double), intent(out) :: cfl
        end function
    end interface

contains

    ! The reference reports a bad option with write(6,*) + stop
    ! (flux1.f90:111-115); a nonzero status from the library does the same.
    subroutine sc_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        type(c_ptr) :: p
        integer :: n
        if (rc == SC_OK) return
        p = sc_last_error_c()
        n = 0
        text = ' '
        if (c_associated(p)) then
            call c_f_pointer(p, msg, [512])
            do while (n < 512)
                if (msg(n+1) == c_null_char) exit
                n = n + 1
                text(n:n) = msg(n)
            enddo
            write(6,*) '*** ERROR *** ', where, ': ', text(1:n)
        else
            write(6,*) '*** ERROR *** ', where, ': status ', rc
        endif
        stop 1
    end subroutine sc_check

end module sharpclaw_gpu
