module aeros_vordiv
    ! Vorticity/divergence <-> (u,v) and the physical gradient, with the API of the
    ! reference's aeros_vordiv (src/dynamics/aeros_vordiv.f90), over the GPU-backed
    ! aeros_spectral of this package.
    !
    ! Mapping (reference src/dynamics/aeros_vordiv.f90:11-51):
    !   u = V_phi,  v = -V_theta,
    !   S_lm = -a/(l(l+1)) D_lm,  T_lm = +a/(l(l+1)) zeta_lm,  l = 0 discarded;
    !   dfdx = (1/a) g_phi (east),  dfdy = -(1/a) g_theta (north).
    !
    ! The per-level routines do the factors on the host around one vector
    ! transform, as the reference does. The *_levels routines take all levels at
    ! once and run the factors inside the fused device kernels
    ! (aer_uv_from_vordiv_d and friends): one launch pair per call.

    use, intrinsic :: iso_c_binding
    use aeros_amd_kinds, only : dp, wp, wp_sh, r_earth, io_unit_err
    use aeros_spectral,  only : aeros_sht_class, aeros_sht_analysis_vec, &
                                aeros_sht_synthesis_vec, aeros_sht_gradient
    use aer_sht_c

    implicit none

    private

    public :: aeros_uv_from_vordiv
    public :: aeros_vordiv_from_uv
    public :: aeros_gradient
    public :: aeros_uv_from_vordiv_levels
    public :: aeros_vordiv_from_uv_levels

contains

    subroutine aeros_uv_from_vordiv(sht, vor, div, u, v)
        ! Spectral vorticity and divergence -> grid winds, one level.
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous :: vor(:), div(:)
        real(wp), intent(out), contiguous :: u(:,:), v(:,:)

        complex(wp_sh) :: slm(sht%nlm), tlm(sht%nlm)
        real(wp)       :: vth(sht%nlon, sht%nlat)
        real(dp)       :: fac
        integer        :: lm, l

        do lm = 1, sht%nlm
            l = sht%l_of_lm(lm)
            if (l == 0) then
                slm(lm) = (0.0_wp_sh, 0.0_wp_sh)
                tlm(lm) = (0.0_wp_sh, 0.0_wp_sh)
            else
                fac = r_earth/real(l*(l+1), dp)
                slm(lm) = -fac*div(lm)
                tlm(lm) =  fac*vor(lm)
            end if
        end do
        call aeros_sht_synthesis_vec(sht, slm, tlm, vth, u)
        v = -vth
    end subroutine aeros_uv_from_vordiv

    subroutine aeros_vordiv_from_uv(sht, u, v, vor, div)
        ! Grid winds -> spectral vorticity and divergence, one level; inputs preserved.
        type(aeros_sht_class), intent(in) :: sht
        real(wp), intent(in), contiguous :: u(:,:), v(:,:)
        complex(wp_sh), intent(out), contiguous :: vor(:), div(:)

        complex(wp_sh) :: slm(sht%nlm), tlm(sht%nlm)
        real(wp)       :: vth(sht%nlon, sht%nlat), vph(sht%nlon, sht%nlat)
        real(dp)       :: fac
        integer        :: lm, l

        vth = -v
        vph =  u
        call aeros_sht_analysis_vec(sht, vth, vph, slm, tlm)
        do lm = 1, sht%nlm
            l = sht%l_of_lm(lm)
            if (l == 0) then
                vor(lm) = (0.0_wp_sh, 0.0_wp_sh)
                div(lm) = (0.0_wp_sh, 0.0_wp_sh)
            else
                fac = real(l*(l+1), dp)/r_earth
                div(lm) = -fac*slm(lm)
                vor(lm) =  fac*tlm(lm)
            end if
        end do
    end subroutine aeros_vordiv_from_uv

    subroutine aeros_gradient(sht, f, dfdx, dfdy)
        ! Physical horizontal gradient, east/north components.
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous :: f(:)
        real(wp), intent(out), contiguous :: dfdx(:,:), dfdy(:,:)

        real(wp) :: gth(sht%nlon, sht%nlat)
        real(wp) :: rinv

        call aeros_sht_gradient(sht, f, gth, dfdx)
        rinv = 1.0_wp/real(r_earth, wp)
        dfdx = dfdx*rinv
        dfdy = -gth*rinv
    end subroutine aeros_gradient

    ! === all levels in one fused launch pair =================================

    subroutine check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc == AER_OK) return
        write(io_unit_err,*) trim(where)//":: error: ", aer_c_string(aer_last_error())
        error stop 1
    end subroutine check

    subroutine aeros_uv_from_vordiv_levels(sht, vor, div, u, v)
        ! vor, div (nlm, nlev) -> u, v (nlon, nlat, nlev).
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous, target :: vor(:,:), div(:,:)
        real(wp), intent(out), contiguous, target :: u(:,:,:), v(:,:,:)
        type(c_ptr) :: d(4)
        integer(c_size_t) :: bq, bg
        integer :: nlev, k

        nlev = size(vor, 2)
        bq = int(16, c_size_t)*sht%nlm*nlev
        bg = int(8, c_size_t)*sht%nlon*sht%nlat*nlev
        call check(aer_dev_malloc(bq, d(1)), "aeros_uv_from_vordiv_levels")
        call check(aer_dev_malloc(bq, d(2)), "aeros_uv_from_vordiv_levels")
        call check(aer_dev_malloc(bg, d(3)), "aeros_uv_from_vordiv_levels")
        call check(aer_dev_malloc(bg, d(4)), "aeros_uv_from_vordiv_levels")
        call check(aer_memcpy_h2d(d(1), c_loc(vor), bq), "aeros_uv_from_vordiv_levels")
        call check(aer_memcpy_h2d(d(2), c_loc(div), bq), "aeros_uv_from_vordiv_levels")
        call check(aer_uv_from_vordiv_d(sht%cfg, int(nlev, c_int), d(1), d(2), d(3), d(4)), &
                   "aeros_uv_from_vordiv_levels")
        call check(aer_sht_sync(sht%cfg), "aeros_uv_from_vordiv_levels")
        call check(aer_memcpy_d2h(c_loc(u), d(3), bg), "aeros_uv_from_vordiv_levels")
        call check(aer_memcpy_d2h(c_loc(v), d(4), bg), "aeros_uv_from_vordiv_levels")
        do k = 1, 4
            call check(aer_dev_free(d(k)), "aeros_uv_from_vordiv_levels")
        end do
    end subroutine aeros_uv_from_vordiv_levels

    subroutine aeros_vordiv_from_uv_levels(sht, u, v, vor, div)
        ! u, v (nlon, nlat, nlev) -> vor, div (nlm, nlev); inputs preserved.
        type(aeros_sht_class), intent(in) :: sht
        real(wp), intent(in), contiguous, target :: u(:,:,:), v(:,:,:)
        complex(wp_sh), intent(out), contiguous, target :: vor(:,:), div(:,:)
        type(c_ptr) :: d(4)
        integer(c_size_t) :: bq, bg
        integer :: nlev, k

        nlev = size(u, 3)
        bq = int(16, c_size_t)*sht%nlm*nlev
        bg = int(8, c_size_t)*sht%nlon*sht%nlat*nlev
        call check(aer_dev_malloc(bg, d(1)), "aeros_vordiv_from_uv_levels")
        call check(aer_dev_malloc(bg, d(2)), "aeros_vordiv_from_uv_levels")
        call check(aer_dev_malloc(bq, d(3)), "aeros_vordiv_from_uv_levels")
        call check(aer_dev_malloc(bq, d(4)), "aeros_vordiv_from_uv_levels")
        call check(aer_memcpy_h2d(d(1), c_loc(u), bg), "aeros_vordiv_from_uv_levels")
        call check(aer_memcpy_h2d(d(2), c_loc(v), bg), "aeros_vordiv_from_uv_levels")
        call check(aer_vordiv_from_uv_d(sht%cfg, int(nlev, c_int), d(1), d(2), d(3), d(4)), &
                   "aeros_vordiv_from_uv_levels")
        call check(aer_sht_sync(sht%cfg), "aeros_vordiv_from_uv_levels")
        call check(aer_memcpy_d2h(c_loc(vor), d(3), bq), "aeros_vordiv_from_uv_levels")
        call check(aer_memcpy_d2h(c_loc(div), d(4), bq), "aeros_vordiv_from_uv_levels")
        do k = 1, 4
            call check(aer_dev_free(d(k)), "aeros_vordiv_from_uv_levels")
        end do
    end subroutine aeros_vordiv_from_uv_levels

end module aeros_vordiv
