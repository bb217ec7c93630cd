! Skip the '#' header of a data file and the first line after it
! (src/opendatafile.f:45-53).
subroutine skip_data_header(iunit)
    implicit none
    integer, intent(in) :: iunit
    character(len=1) :: c
    c = '#'
    do while (c == '#')
        read(iunit, '(a1)') c
    enddo
end subroutine skip_data_header
