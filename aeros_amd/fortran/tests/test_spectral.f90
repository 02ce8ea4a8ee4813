program test_spectral
    ! Acceptance test of the GPU-backed aeros_spectral, restating the reference's
    ! tests/test_spectral.f90 with its tolerances: grid sizing, coefficient
    ! indexing, the round trips (< 20 eps), quadrature, the exact Laplacian, and
    ! the pool bit-identity -- plus the level-batched extension against the
    ! per-level calls, bit for bit. Exits non-zero on failure.

    use aeros_amd_kinds, only : dp, wp, wp_sh, pi, r_earth
    use aeros_spectral

    implicit none

    type(aeros_sht_class) :: sht
    integer, parameter :: trunc = 21
    real(wp),       allocatable :: field(:,:), field2(:,:)
    complex(wp_sh), allocatable :: coeffs(:), coeffs2(:)
    real(dp) :: tol, err, total, expect
    integer  :: lm, l, m, nfail, nlon, nlat

    nfail = 0
    tol = 20.0_dp*real(epsilon(1.0_wp), dp)

    ! grid sizes of the standard truncations (tests/test_grid.f90:25-37)
    call aeros_sht_grid_size(21, nlon, nlat);  call check(nlon ==  64 .and. nlat ==  32, "T21 -> 64x32", nfail)
    call aeros_sht_grid_size(31, nlon, nlat);  call check(nlon ==  96 .and. nlat ==  48, "T31 -> 96x48", nfail)
    call aeros_sht_grid_size(42, nlon, nlat);  call check(nlon == 128 .and. nlat ==  64, "T42 -> 128x64", nfail)
    call aeros_sht_grid_size(85, nlon, nlat);  call check(nlon == 256 .and. nlat == 128, "T85 -> 256x128", nfail)

    call aeros_sht_init(sht, trunc, quick=.TRUE.)
    write(*,*) "test_spectral:: T", sht%trunc, " grid ", sht%nlon, "x", sht%nlat, " nlm ", sht%nlm

    allocate(field(sht%nlon,sht%nlat), field2(sht%nlon,sht%nlat))
    allocate(coeffs(sht%nlm), coeffs2(sht%nlm))

    call check(sht%nlon >= 3*trunc + 1, "grid is quadratically unaliased", nfail)
    call check(sht%nlat == sht%nlon/2,  "nlat = nlon/2", nfail)
    call check(sht%nlm == (trunc+1)*(trunc+2)/2, "nlm = (T+1)(T+2)/2", nfail)

    do m = 0, sht%mmax
        do l = m, sht%lmax
            lm = aeros_sht_lm(sht, l, m)
            if (lm < 1 .or. lm > sht%nlm) then
                nfail = nfail + 1
                cycle
            end if
            if (sht%l_of_lm(lm) /= l .or. sht%m_of_lm(lm) /= m) nfail = nfail + 1
        end do
    end do
    call check(aeros_sht_lm(sht, trunc+1, 0) == -1, "out-of-truncation index returns -1", nfail)
    call check(sht%sinlat(1) > sht%sinlat(sht%nlat), "latitude runs north to south", nfail)
    call check(all(sht%gauss_w == sht%gauss_w(sht%nlat:1:-1)), "Gauss weights exactly symmetric", nfail)

    call fill_test_spectrum(coeffs, sht)
    call aeros_sht_synthesis(sht, coeffs, field)
    field2 = field
    call aeros_sht_analysis(sht, field2, coeffs2)
    err = maxval(abs(coeffs2 - coeffs))/maxval(abs(coeffs))
    call check(err < tol, "spectral -> grid -> spectral round trip", nfail)
    write(*,"(a40,es12.3,a,es12.3)") "   round-trip relative error ", err, "  tol ", tol
    call check(all(field2 == field), "analysis leaves its input intact", nfail)

    field  = 1.0_wp
    total  = aeros_sht_surface_integral(sht, field)
    expect = 4.0_dp*pi*r_earth*r_earth
    err    = abs(total - expect)/expect
    call check(err < tol, "surface integral of 1 = 4*pi*a^2", nfail)

    coeffs = (0.0_wp_sh, 0.0_wp_sh)
    coeffs(aeros_sht_lm(sht,1,0)) = (1.0_wp_sh, 0.0_wp_sh)
    call aeros_sht_synthesis(sht, coeffs, field)
    total = aeros_sht_surface_integral(sht, field)
    call check(abs(total)/expect < tol, "surface integral of Y_10 = 0", nfail)

    call fill_test_spectrum(coeffs, sht)
    coeffs2 = coeffs
    call aeros_sht_laplacian(sht, coeffs2)
    err = 0.0_dp
    do lm = 1, sht%nlm
        l = sht%l_of_lm(lm)
        err = max(err, abs(coeffs2(lm) - coeffs(lm)*(-real(l*(l+1),dp)/(r_earth*r_earth))))
    end do
    call check(err == 0.0_dp, "Laplacian eigenvalue -l(l+1)/a^2 exact", nfail)
    call aeros_sht_laplacian(sht, coeffs2, inverse=.TRUE.)
    coeffs(aeros_sht_lm(sht,0,0)) = (0.0_wp_sh, 0.0_wp_sh)
    err = maxval(abs(coeffs2 - coeffs))/maxval(abs(coeffs))
    call check(err < 1.0e-12_dp, "inverse Laplacian recovers l > 0", nfail)

    call test_vector_roundtrip(sht, tol, nfail)
    call test_levels_and_pool(trunc, nfail)

    call aeros_sht_end(sht)

    write(*,*) ""
    if (nfail == 0) then
        write(*,*) "test_spectral:: PASS"
    else
        write(*,*) "test_spectral:: FAIL, ", nfail, " check(s) failed"
        stop 1
    end if

contains

    subroutine fill_test_spectrum(coeffs, sht)
        ! amplitude 1/(l+1)^2, cos(3l+m) + i sin(l+2m), m = 0 real
        ! (the reference's fixture, tests/test_spectral.f90:147-172)
        complex(wp_sh), intent(out) :: coeffs(:)
        type(aeros_sht_class), intent(in) :: sht
        integer  :: lm, l, m
        real(dp) :: amp
        do lm = 1, sht%nlm
            l = sht%l_of_lm(lm)
            m = sht%m_of_lm(lm)
            amp = 1.0_dp/real((l+1)*(l+1), dp)
            coeffs(lm) = cmplx(amp*cos(real(3*l + m, dp)), amp*sin(real(l + 2*m, dp)), wp_sh)
            if (m == 0) coeffs(lm) = cmplx(real(coeffs(lm)), 0.0_wp_sh, wp_sh)
        end do
    end subroutine fill_test_spectrum

    subroutine test_vector_roundtrip(sht, tol, nfail)
        type(aeros_sht_class), intent(in) :: sht
        real(dp), intent(in) :: tol
        integer, intent(inout) :: nfail
        real(wp),       allocatable :: vth(:,:), vph(:,:)
        complex(wp_sh), allocatable :: slm(:), tlm(:), slm2(:), tlm2(:)
        real(dp) :: err

        allocate(vth(sht%nlon,sht%nlat), vph(sht%nlon,sht%nlat))
        allocate(slm(sht%nlm), tlm(sht%nlm), slm2(sht%nlm), tlm2(sht%nlm))
        call fill_test_spectrum(slm, sht)
        call fill_test_spectrum(tlm, sht)
        slm(aeros_sht_lm(sht,0,0)) = (0.0_wp_sh, 0.0_wp_sh)
        tlm(aeros_sht_lm(sht,0,0)) = (0.0_wp_sh, 0.0_wp_sh)
        call aeros_sht_synthesis_vec(sht, slm, tlm, vth, vph)
        call aeros_sht_analysis_vec(sht, vth, vph, slm2, tlm2)
        err = max(maxval(abs(slm2 - slm)), maxval(abs(tlm2 - tlm))) &
                / max(maxval(abs(slm)), maxval(abs(tlm)))
        call check(err < tol, "spheroidal/toroidal round trip", nfail)
        write(*,"(a40,es12.3)") "   vector round-trip relative error ", err
    end subroutine test_vector_roundtrip

    subroutine test_levels_and_pool(trunc, nfail)
        ! The level-batched launch and an OpenMP level loop over the pool must
        ! both reproduce the serial per-level calls exactly
        ! (tests/test_spectral.f90:209-283).
        !$ use omp_lib
        integer, intent(in) :: trunc
        integer, intent(inout) :: nfail
        integer, parameter :: nlev = 16
        type(aeros_sht_class), pointer :: s
        type(aeros_sht_pool_class) :: pool
        type(aeros_sht_class) :: ref
        real(wp),       allocatable :: fld_s(:,:,:), fld_p(:,:,:), fld_b(:,:,:)
        complex(wp_sh), allocatable :: coef(:,:), out_s(:,:), out_p(:,:), out_b(:,:)
        integer :: k, nt

        nt = 1
        !$ nt = min(omp_get_max_threads(), 4)
        call aeros_sht_init(ref, trunc, quick=.TRUE.)
        call aeros_sht_pool_init(pool, trunc, nthreads=nt, quick=.TRUE.)
        call check(pool%nthreads == nt, "pool built one handle per thread", nfail)
        allocate(fld_s(ref%nlon,ref%nlat,nlev), fld_p(ref%nlon,ref%nlat,nlev), fld_b(ref%nlon,ref%nlat,nlev))
        allocate(coef(ref%nlm,nlev), out_s(ref%nlm,nlev), out_p(ref%nlm,nlev), out_b(ref%nlm,nlev))
        do k = 1, nlev
            call fill_test_spectrum(coef(:,k), ref)
            coef(:,k) = coef(:,k)*real(k, wp_sh)
        end do
        do k = 1, nlev
            call aeros_sht_synthesis(ref, coef(:,k), fld_s(:,:,k))
            call aeros_sht_analysis(ref, fld_s(:,:,k), out_s(:,k))
        end do
        !$omp parallel do private(k,s) schedule(static) num_threads(nt)
        do k = 1, nlev
            s => aeros_sht_pool_get(pool)
            call aeros_sht_synthesis(s, coef(:,k), fld_p(:,:,k))
            call aeros_sht_analysis(s, fld_p(:,:,k), out_p(:,k))
        end do
        !$omp end parallel do
        call check(all(fld_p == fld_s), "pooled grid fields are bit-identical to serial", nfail)
        call check(all(out_p == out_s), "pooled coefficients are bit-identical to serial", nfail)
        call aeros_sht_synthesis_levels(ref, coef, fld_b)
        call aeros_sht_analysis_levels(ref, fld_b, out_b)
        call check(all(fld_b == fld_s), "level-batched synthesis is bit-identical to per-level", nfail)
        call check(all(out_b == out_s), "level-batched analysis is bit-identical to per-level", nfail)
        write(*,"(a44,es12.3,es12.3)") "   batched vs per-level max |diff| ", &
            maxval(abs(fld_b - fld_s)), maxval(abs(out_b - out_s))
        call aeros_sht_pool_end(pool)
        call aeros_sht_end(ref)
    end subroutine test_levels_and_pool

    subroutine check(ok, label, nfail)
        logical, intent(in) :: ok
        character(len=*), intent(in) :: label
        integer, intent(inout) :: nfail
        if (ok) then
            write(*,*) "  ok   : ", trim(label)
        else
            write(*,*) "  FAIL : ", trim(label)
            nfail = nfail + 1
        end if
    end subroutine check

end program test_spectral
