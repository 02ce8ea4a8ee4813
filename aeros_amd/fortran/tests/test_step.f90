program test_step
    ! Held-Suarez dynamics steps driven from Fortran through the C ABI, in the order of the
    ! reference's aeros_timestep_step (src/dynamics/aeros_timestep.f90:776-1137): the
    ! spectral->grid transform half, the column terms with Held-Suarez forcing, the
    ! grid->spectral half, the semi-implicit leapfrog with diffusion and RAW filter, and the
    ! mass fixer, on device arrays, at T21L8 from the Held-Suarez start (isothermal 264 K
    ! plus the aeros_hs_perturb seed, src/physics/aeros_held_suarez.f90:295-336).
    !
    ! Checks: the state stays finite and the fixer holds the mass at its target. It also
    ! writes the vertical grid, the initial state and the final state to test_step.bin
    ! (stream access) so that a host in another language can replay the same steps and
    ! compare.
    use, intrinsic :: iso_c_binding
    use aer_sht_c
    use aer_dyn_c
    implicit none

    integer, parameter :: T = 21, NLEV = 8, NSTEP = 6
    real(c_double), parameter :: DT = 1800.0d0, P0 = 1.0d5
    type(c_ptr) :: cfg, vg, si
    type(aer_sht_info) :: info
    integer(c_int) :: nlat, nlon
    integer :: nlm, k, l, m, lm, n, q, j, io, ic, inew, itmp, u
    real(c_double) :: a_half(NLEV + 1), b_half(NLEV + 1), target, h, hcur, eps, y00
    real(c_double), target :: mass_end(1)
    complex(c_double_complex), allocatable, target :: x0(:, :, :), x1(:, :, :)
    type(c_ptr) :: X(4, 3), R(4), G(10), C(5), dmass, dlnr
    integer(c_size_t) :: s3, s2, g3, g2
    integer(c_int) :: flags
    logical :: ok

    call check(aer_sht_grid_size(int(T, c_int), nlon, nlat))
    call check(aer_sht_create(int(T, c_int), int(T, c_int), 1_c_int, &
               ior(AER_SHT_ORTHONORMAL, AER_SHT_NO_CS_PHASE), cfg))
    call check(aer_sht_set_grid(cfg, ior(AER_SHT_PHI_CONTIGUOUS, AER_SHT_QUICK_INIT), 1.0d-10, nlat, nlon))
    call check(aer_sht_get_info(cfg, info))
    nlm = info%nlm

    ! uniform sigma levels under a 1000 Pa top (the Held-Suarez grid)
    a_half = 0.0d0
    b_half = 0.0d0
    a_half(1) = 1000.0d0
    do j = 1, NLEV - 1
        b_half(j + 1) = 0.01d0 + 0.99d0 * real(j, c_double) / real(NLEV, c_double)
    end do
    b_half(NLEV + 1) = 1.0d0
    call check(aer_vgrid_create(int(NLEV, c_int), a_half, b_half, vg))
    call check(aer_semiimp_create(cfg, vg, 300.0d0, P0, DT, 0.5d0, si))
    call check(aer_semiimp_set_diffusion(si, 6.0d0 * 3600.0d0, 0.0d0, 6_c_int))

    ! initial state: (vor, div, temp) as (nlm, nlev), ln p_s in column 1 of the 4th field
    allocate(x0(nlm, NLEV, 4), x1(nlm, NLEV, 4))
    x0 = (0.0d0, 0.0d0)
    y00 = sqrt(4.0d0 * acos(-1.0d0))
    x0(1, :, 3) = cmplx(264.0d0 * y00, 0.0d0, c_double_complex)
    x0(1, 1, 4) = cmplx(log(P0) * y00, 0.0d0, c_double_complex)
    do k = 1, NLEV
        do m = 1, min(8, T)
            do l = m, min(m + 4, T)
                lm = (m * (2 * T + 2 - (m + 1))) / 2 + l + 1
                x0(lm, k, 3) = x0(lm, k, 3) + cmplx(cos(real(3 * l + 5 * m + k, c_double)), &
                                                    sin(real(7 * l + 2 * m + 3 * k, c_double)), c_double_complex)
            end do
        end do
    end do

    s3 = int(16, c_size_t) * nlm * NLEV
    s2 = int(16, c_size_t) * nlm
    g3 = int(8, c_size_t) * nlon * nlat * NLEV
    g2 = int(8, c_size_t) * nlon * nlat
    do q = 1, 3
        do n = 1, 3
            X(n, q) = dalloc(s3)
        end do
        X(4, q) = dalloc(s2)
    end do
    do n = 1, 3
        R(n) = dalloc(s3)
    end do
    R(4) = dalloc(s2)
    do n = 1, 7
        G(n) = dalloc(g3)
    end do
    do n = 8, 10
        G(n) = dalloc(g2)
    end do
    do n = 1, 4
        C(n) = dalloc(g3)
    end do
    C(5) = dalloc(g2)
    dmass = dalloc(8_c_size_t)
    dlnr = dalloc(8_c_size_t)

    ! X^(-1) := X^0 (the start-up step is forward over dt)
    ic = 1
    io = 3
    inew = 2
    do q = 1, 3, 2
        do n = 1, 3
            call check(aer_memcpy_h2d(X(n, q), c_loc(x0(1, 1, n)), s3))
        end do
        call check(aer_memcpy_h2d(X(4, q), c_loc(x0(1, 1, 4)), s2))
    end do
    call check(aer_mass_d(cfg, X(4, ic), 0_c_int, 0.0d0, dmass, dlnr))
    call check(aer_sht_sync(cfg))
    call check(aer_memcpy_d2h(c_loc(mass_end), dmass, 8_c_size_t))
    target = mass_end(1)

    hcur = DT
    do n = 0, NSTEP - 1
        h = DT
        if (n > 0) h = 2.0d0 * DT
        if (h /= hcur) then
            call check(aer_semiimp_set_step(si, h))
            hcur = h
        end if
        call check(aer_tendency_grid_transforms_d(cfg, int(NLEV, c_int), X(1, ic), X(2, ic), X(3, ic), X(4, ic), &
                   G(1), G(2), G(3), G(4), G(5), G(6), G(7), G(8), G(9), G(10)))
        call check(aer_column_terms_d(cfg, vg, 1_c_int, G(1), G(2), G(3), G(4), G(5), G(6), G(7), G(8), G(9), &
                   G(10), c_null_ptr, C(1), C(2), C(3), C(4), C(5)))
        call check(aer_tendency_spectral_d(cfg, int(NLEV, c_int), C(1), C(2), C(3), C(4), C(5), &
                   R(1), R(2), R(3), R(4)))
        eps = 0.0d0
        if (n > 0) eps = 0.06d0
        flags = ior(AER_STEP_SEMI_IMPLICIT, AER_STEP_DIFFUSE)
        if (eps > 0.0d0) flags = ior(flags, AER_STEP_FILTER)
        call check(aer_leapfrog_d(si, flags, eps, 0.53d0, X(1, io), X(2, io), X(3, io), X(4, io), &
                   X(1, ic), X(2, ic), X(3, ic), X(4, ic), R(1), R(2), R(3), R(4), &
                   X(1, inew), X(2, inew), X(3, inew), X(4, inew)))
        call check(aer_mass_d(cfg, X(4, inew), 1_c_int, target, dmass, dlnr))
        ! the filtered X^n becomes X^(n-1), X^(n+1) becomes X^n
        itmp = io
        io = ic
        ic = inew
        inew = itmp
    end do

    call check(aer_mass_d(cfg, X(4, ic), 0_c_int, 0.0d0, dmass, dlnr))
    call check(aer_sht_sync(cfg))
    call check(aer_memcpy_d2h(c_loc(mass_end), dmass, 8_c_size_t))
    x1 = (0.0d0, 0.0d0)
    do n = 1, 3
        call check(aer_memcpy_d2h(c_loc(x1(1, 1, n)), X(n, ic), s3))
    end do
    call check(aer_memcpy_d2h(c_loc(x1(1, 1, 4)), X(4, ic), s2))

    open(newunit=u, file="test_step.bin", access="stream", form="unformatted", status="replace")
    write(u) int(NLEV, c_int), int(nlm, c_int), int(NSTEP, c_int), DT, target
    write(u) a_half, b_half
    do n = 1, 3
        write(u) x0(:, :, n)
    end do
    write(u) x0(:, 1, 4)
    do n = 1, 3
        write(u) x1(:, :, n)
    end do
    write(u) x1(:, 1, 4)
    close(u)

    ok = all(abs(real(x1(:, :, 1:3))) < huge(1.0d0)) .and. all(abs(aimag(x1(:, :, 1:3))) < huge(1.0d0))
    ok = ok .and. abs(mass_end(1) - target) <= 1.0d-13 * target
    print '(a, es23.15, a, es23.15)', "mass target ", target, "  after the steps ", mass_end(1)
    do q = 1, 3
        do n = 1, 4
            call check(aer_dev_free(X(n, q)))
        end do
    end do
    do n = 1, 4
        call check(aer_dev_free(R(n)))
    end do
    do n = 1, 10
        call check(aer_dev_free(G(n)))
    end do
    do n = 1, 5
        call check(aer_dev_free(C(n)))
    end do
    call check(aer_dev_free(dmass))
    call check(aer_dev_free(dlnr))
    call check(aer_semiimp_destroy(si))
    call check(aer_vgrid_destroy(vg))
    call check(aer_sht_destroy(cfg))
    if (ok) then
        print '(a)', "test_step:: PASS"
    else
        print '(a)', "test_step:: FAIL"
        error stop 1
    end if

contains

    subroutine check(rc)
        integer(c_int), intent(in) :: rc
        if (rc /= AER_OK) then
            print '(a, i0, 2a)', "test_step: C ABI error ", rc, ": ", aer_c_string(aer_last_error())
            error stop 1
        end if
    end subroutine check

    function dalloc(nbytes) result(p)
        integer(c_size_t), intent(in) :: nbytes
        type(c_ptr) :: p
        call check(aer_dev_malloc(nbytes, p))
    end function dalloc

end program test_step
