! The problem parameters under the names and types of the reference's
! module ClawParams (src/ClawParams.f90:1-15), filled by the driver from
! sharpclaw.data as src/main.f90:61-148 fills them, so that user routines
! written against the reference (src, bc, b4step: `use ClawParams`) compile
! unchanged against this driver.
module ClawParams
    implicit none
    integer :: ndim, mcapa, maux, meqn, mwaves, mbc, maxnx
    integer, allocatable :: nx(:)
    double precision, allocatable :: xlower(:), xupper(:), dx(:)
    integer, allocatable :: mthbc(:)
    integer :: src_term, tfluct_solver, char_decomp, lim_type, multid_recon
    integer, allocatable :: mthlim(:)
    integer :: rord
end module ClawParams
