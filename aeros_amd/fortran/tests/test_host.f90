program test_host
    ! Host-only checks of the Fortran binding (no GPU needed): the grid-size rule
    ! of the reference (tests/test_grid.f90:25-37) through the C ABI, and the
    ! library's refusal of an unsupported configuration.
    use, intrinsic :: iso_c_binding
    use aer_sht_c
    use aeros_spectral, only : aeros_sht_grid_size
    implicit none
    integer :: nlon, nlat, nfail
    type(c_ptr) :: h
    nfail = 0
    call aeros_sht_grid_size(21, nlon, nlat);  if (nlon /=  64 .or. nlat /=  32) nfail = nfail + 1
    call aeros_sht_grid_size(31, nlon, nlat);  if (nlon /=  96 .or. nlat /=  48) nfail = nfail + 1
    call aeros_sht_grid_size(42, nlon, nlat);  if (nlon /= 128 .or. nlat /=  64) nfail = nfail + 1
    call aeros_sht_grid_size(85, nlon, nlat);  if (nlon /= 256 .or. nlat /= 128) nfail = nfail + 1
    call aeros_sht_grid_size(106, nlon, nlat); if (nlon /= 320 .or. nlat /= 160) nfail = nfail + 1
    ! Condon-Shortley phase requested: refused
    if (aer_sht_create(21_c_int, 21_c_int, 1_c_int, AER_SHT_ORTHONORMAL, h) == AER_OK) nfail = nfail + 1
    if (aer_sht_create(21_c_int, 21_c_int, 1_c_int, AER_SHT_ORTHONORMAL + AER_SHT_NO_CS_PHASE, h) /= AER_OK) &
        nfail = nfail + 1
    if (aer_sht_destroy(h) /= AER_OK) nfail = nfail + 1
    if (nfail == 0) then
        write(*,*) "test_host:: PASS"
    else
        write(*,*) "test_host:: FAIL, ", nfail
        stop 1
    end if
end program test_host
