! Direction bookkeeping for the 2D characteristic-decomposition hooks.
!
! The reference flux1 calls evec without the sweep direction ixy
! (src/flux1.f90:105, "call evec(mx,meqn,mbc,mx,q1d,aux,aux,g%evl,g%evr)"),
! while src/2d/flux2.f90 calls flux1 for rows j=0..ny+1 with ixy=1
! (flux2.f90:36-46) and then for columns i=0..nx+1 with ixy=2
! (flux2.f90:68-79).  Each flux1 call invokes evec exactly once in the
! characteristic mode, so the call count modulo (nx+ny+4) recovers ixy.
! A harness may force the direction instead (ixy_forced > 0).
module evec_dir
    implicit none
    integer :: ixy_forced = 0
    integer :: ncalls = 0
contains
    integer function evec_next_ixy(nxv, nyv)
        integer, intent(in) :: nxv, nyv
        integer :: k
        if (ixy_forced > 0) then
            evec_next_ixy = ixy_forced
            return
        endif
        k = mod(ncalls, nxv + nyv + 4)
        if (k < nyv + 2) then
            evec_next_ixy = 1
        else
            evec_next_ixy = 2
        endif
        ncalls = ncalls + 1
    end function evec_next_ixy
end module evec_dir
