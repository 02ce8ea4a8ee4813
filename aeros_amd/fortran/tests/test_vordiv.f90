program test_vordiv
    ! The reference's tests/test_vordiv.f90 restated against the GPU-backed modules:
    ! solid-body rotation (closed form), pure divergence, the round trip and the
    ! l = 0 rule, with its tolerance 1e3 eps; plus the all-levels fused routines
    ! against the per-level ones. Exits non-zero on failure.

    use aeros_amd_kinds, only : dp, wp, wp_sh, pi, r_earth
    use aeros_spectral
    use aeros_vordiv

    implicit none

    type(aeros_sht_class) :: sht
    integer, parameter :: trunc = 21
    real(wp) :: tol
    integer  :: nfail

    nfail = 0
    tol   = 1.0e3_wp*epsilon(1.0_wp)
    call aeros_sht_init(sht, trunc, quick=.TRUE.)
    call test_solid_body(sht, tol, nfail)
    call test_pure_divergence(sht, tol, nfail)
    call test_roundtrip_and_levels(sht, tol, nfail)
    call test_l0_discarded(sht, nfail)
    call aeros_sht_end(sht)

    write(*,*) ""
    if (nfail == 0) then
        write(*,*) "test_vordiv:: PASS"
    else
        write(*,*) "test_vordiv:: FAIL, ", nfail, " check(s) failed"
        stop 1
    end if

contains

    subroutine test_solid_body(sht, tol, nfail)
        type(aeros_sht_class), intent(in) :: sht
        real(wp), intent(in) :: tol
        integer, intent(inout) :: nfail
        real(wp), parameter :: u0 = 40.0_wp
        real(wp)       :: u(sht%nlon,sht%nlat), v(sht%nlon,sht%nlat), zeta_g(sht%nlon,sht%nlat)
        complex(wp_sh) :: vor(sht%nlm), div(sht%nlm)
        real(dp)       :: expect, err, scale, nrm
        integer        :: i, j, lm10

        do j = 1, sht%nlat
            do i = 1, sht%nlon
                u(i,j) = u0*real(sin(sht%colat(j)), wp)
                v(i,j) = 0.0_wp
            end do
        end do
        call aeros_vordiv_from_uv(sht, u, v, vor, div)
        err = maxval(abs(div))/maxval(abs(vor))
        call check(err < tol, "solid-body rotation is non-divergent", nfail)
        lm10   = aeros_sht_lm(sht, 1, 0)
        nrm    = sqrt(3.0_dp/(4.0_dp*pi))
        expect = (2.0_dp*real(u0, dp)/r_earth)/nrm
        err = abs(real(vor(lm10), dp) - expect)/abs(expect)
        call check(err < tol, "zeta_(1,0) has the analytic value", nfail)
        write(*,"(a44,es12.3)") "   zeta_10 relative error ", err
        vor(lm10) = (0.0_wp_sh, 0.0_wp_sh)
        call check(maxval(abs(vor))/abs(expect) < tol, "no other vorticity coefficient is excited", nfail)
        call aeros_vordiv_from_uv(sht, u, v, vor, div)
        call aeros_sht_synthesis(sht, vor, zeta_g)
        err   = 0.0_dp
        scale = 2.0_dp*real(u0, dp)/r_earth
        do j = 1, sht%nlat
            do i = 1, sht%nlon
                err = max(err, abs(real(zeta_g(i,j), dp) - scale*sht%sinlat(j)))
            end do
        end do
        call check(err/scale < tol, "grid-space zeta = (2 U0/a) sin(lat)", nfail)
        call check(real(zeta_g(1,1), dp) > 0.0_dp, "westerly flow gives positive vorticity in the north", nfail)
        call check(real(zeta_g(1,sht%nlat), dp) < 0.0_dp, "westerly flow gives negative vorticity in the south", nfail)
    end subroutine test_solid_body

    subroutine test_pure_divergence(sht, tol, nfail)
        type(aeros_sht_class), intent(in) :: sht
        real(wp), intent(in) :: tol
        integer, intent(inout) :: nfail
        real(wp)       :: u(sht%nlon,sht%nlat), v(sht%nlon,sht%nlat)
        complex(wp_sh) :: vor(sht%nlm), div(sht%nlm), vor2(sht%nlm), div2(sht%nlm)
        vor = (0.0_wp_sh, 0.0_wp_sh)
        div = (0.0_wp_sh, 0.0_wp_sh)
        div(aeros_sht_lm(sht, 2, 1)) = cmplx(1.0e-6_dp, -4.0e-7_dp, wp_sh)
        call aeros_uv_from_vordiv(sht, vor, div, u, v)
        call aeros_vordiv_from_uv(sht, u, v, vor2, div2)
        call check(maxval(abs(vor2))/maxval(abs(div)) < tol, "a divergent field carries no vorticity", nfail)
        call check(maxval(abs(div2 - div))/maxval(abs(div)) < tol, "divergence survives the round trip", nfail)
    end subroutine test_pure_divergence

    subroutine test_roundtrip_and_levels(sht, tol, nfail)
        type(aeros_sht_class), intent(in) :: sht
        real(wp), intent(in) :: tol
        integer, intent(inout) :: nfail
        integer, parameter :: nlev = 6
        real(wp)       :: u(sht%nlon,sht%nlat), v(sht%nlon,sht%nlat)
        real(wp)       :: ub(sht%nlon,sht%nlat,nlev), vb(sht%nlon,sht%nlat,nlev)
        complex(wp_sh) :: vor(sht%nlm), div(sht%nlm), vor2(sht%nlm), div2(sht%nlm)
        complex(wp_sh) :: vl(sht%nlm,nlev), dl(sht%nlm,nlev), vl2(sht%nlm,nlev), dl2(sht%nlm,nlev)
        real(dp)       :: amp, err, scale, e2
        integer        :: lm, l, m, k

        do lm = 1, sht%nlm
            l = sht%l_of_lm(lm)
            m = sht%m_of_lm(lm)
            if (l == 0) then
                vor(lm) = (0.0_wp_sh, 0.0_wp_sh)
                div(lm) = (0.0_wp_sh, 0.0_wp_sh)
                cycle
            end if
            amp = 1.0e-5_dp/real(l*l, dp)
            vor(lm) = cmplx(amp*cos(real(3*l+m,dp)), amp*sin(real(l+2*m,dp)), wp_sh)
            div(lm) = cmplx(0.3_dp*amp*sin(real(l+m,dp)), 0.3_dp*amp*cos(real(2*l+m,dp)), wp_sh)
            if (m == 0) then
                vor(lm) = cmplx(real(vor(lm)), 0.0_wp_sh, wp_sh)
                div(lm) = cmplx(real(div(lm)), 0.0_wp_sh, wp_sh)
            end if
        end do
        call aeros_uv_from_vordiv(sht, vor, div, u, v)
        call aeros_vordiv_from_uv(sht, u, v, vor2, div2)
        scale = max(maxval(abs(vor)), maxval(abs(div)))
        err   = max(maxval(abs(vor2 - vor)), maxval(abs(div2 - div)))/scale
        call check(err < tol, "(zeta,D) -> (u,v) -> (zeta,D) is the identity", nfail)
        write(*,"(a44,es12.3)") "   round-trip relative error ", err
        call check(maxval(abs(u)) > 1.0_wp .and. maxval(abs(u)) < 200.0_wp, "wind speeds are physically plausible", nfail)
        ! all levels at once against the per-level routine
        do k = 1, nlev
            vl(:,k) = vor*real(k, wp_sh)
            dl(:,k) = div*real(nlev + 1 - k, wp_sh)
        end do
        call aeros_uv_from_vordiv_levels(sht, vl, dl, ub, vb)
        call aeros_vordiv_from_uv_levels(sht, ub, vb, vl2, dl2)
        err = 0.0_dp
        e2 = 0.0_dp
        do k = 1, nlev
            call aeros_uv_from_vordiv(sht, vl(:,k), dl(:,k), u, v)
            err = max(err, maxval(abs(ub(:,:,k) - u))/maxval(abs(u)), maxval(abs(vb(:,:,k) - v))/maxval(abs(v)))
            e2 = max(e2, maxval(abs(vl2(:,k) - vl(:,k)))/maxval(abs(vl(:,k))), &
                         maxval(abs(dl2(:,k) - dl(:,k)))/maxval(abs(dl(:,k))))
        end do
        call check(err < 1.0e-13_dp, "fused all-level (u,v) matches the per-level routine", nfail)
        call check(e2 < tol, "fused all-level round trip is the identity", nfail)
    end subroutine test_roundtrip_and_levels

    subroutine test_l0_discarded(sht, nfail)
        type(aeros_sht_class), intent(in) :: sht
        integer, intent(inout) :: nfail
        real(wp)       :: u(sht%nlon,sht%nlat), v(sht%nlon,sht%nlat), u2(sht%nlon,sht%nlat), v2(sht%nlon,sht%nlat)
        complex(wp_sh) :: vor(sht%nlm), div(sht%nlm), vor_rt(sht%nlm), div_rt(sht%nlm)
        integer        :: lm00

        lm00 = aeros_sht_lm(sht, 0, 0)
        vor = (0.0_wp_sh, 0.0_wp_sh)
        div = (0.0_wp_sh, 0.0_wp_sh)
        vor(aeros_sht_lm(sht,3,2)) = cmplx(1.0e-5_dp, 2.0e-6_dp, wp_sh)
        call aeros_uv_from_vordiv(sht, vor, div, u, v)
        call aeros_vordiv_from_uv(sht, u, v, vor_rt, div_rt)
        call check(vor_rt(lm00) == (0.0_wp_sh, 0.0_wp_sh), "analysis returns zeta_00 = 0", nfail)
        call check(div_rt(lm00) == (0.0_wp_sh, 0.0_wp_sh), "analysis returns D_00 = 0", nfail)
        vor(lm00) = cmplx(1.0e-3_dp, 0.0_wp_sh, wp_sh)
        div(lm00) = cmplx(1.0e-3_dp, 0.0_wp_sh, wp_sh)
        call aeros_uv_from_vordiv(sht, vor, div, u2, v2)
        call check(all(u2 == u) .and. all(v2 == v), "a nonzero degree-0 input is discarded, bit-for-bit", nfail)
    end subroutine test_l0_discarded

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

end program test_vordiv
