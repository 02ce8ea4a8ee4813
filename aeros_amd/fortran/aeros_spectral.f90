module aeros_spectral
    ! Drop-in replacement of the reference's aeros_spectral module
    ! (src/aeros_spectral.f90): the same public names, type fields and argument
    ! conventions, with the transforms running on an MI355X through the C ABI of
    ! libaer_sht.so instead of SHTns.
    !
    ! CONVENTIONS (as in the reference, src/aeros_spectral.f90:17-26):
    !   orthonormal harmonics, no Condon-Shortley phase;
    !   spatial fields (nlon, nlat), longitude contiguous, latitude north to south;
    !   triangular truncation lmax = mmax = trunc, mres = 1;
    !   double precision on both sides, arrays passed by address with no copy.
    !
    ! Differences a caller can see:
    !   * analyses leave their input fields intact (the reference allows SHTns to
    !     destroy them, src/aeros_spectral.f90:125-126);
    !   * `nthreads`, `quick` and `cache` are accepted and change nothing: there is
    !     no auto-tuning step and no SHTns thread pool;
    !   * every library failure stops the program with a message, where the
    !     reference ignores SHTns return values (src/aeros_spectral.f90:190);
    !   * extensions: aeros_sht_set_precision, and the level-batched
    !     aeros_sht_synthesis_levels / aeros_sht_analysis_levels, which run all
    !     levels of a 3-D field in one launch.

    use, intrinsic :: iso_c_binding
    !$ use omp_lib
    use aeros_amd_kinds, only : sp, dp, wp, wp_sh, pi, r_earth, io_unit_err
    use aer_sht_c

    implicit none

    private

    type aeros_sht_class
        type(c_ptr) :: cfg = c_null_ptr    ! opaque handle of the gfx950 library

        integer :: trunc = 0
        integer :: lmax  = 0
        integer :: mmax  = 0
        integer :: mres  = 1
        integer :: nlon  = 0
        integer :: nlat  = 0
        integer :: nlm   = 0

        integer, allocatable :: l_of_lm(:)
        integer, allocatable :: m_of_lm(:)

        real(dp), allocatable :: sinlat(:)  ! sin(latitude) = cos(colatitude), (nlat)
        real(dp), allocatable :: colat(:)   ! colatitude [rad], (nlat), ascending
        real(dp), allocatable :: lon(:)     ! longitude [rad], (nlon)
        real(dp), allocatable :: gauss_w(:) ! quadrature weights, (nlat), sum = 2
    end type aeros_sht_class

    type aeros_sht_pool_class
        ! One handle per OpenMP thread, as the reference's pool
        ! (src/aeros_spectral.f90:87-113): a handle is not safe for concurrent calls.
        integer :: nthreads = 0
        type(aeros_sht_class), allocatable :: sht(:)
    end type aeros_sht_pool_class

    public :: aeros_sht_class
    public :: aeros_sht_init
    public :: aeros_sht_end
    public :: aeros_sht_grid_size
    public :: aeros_sht_pool_class
    public :: aeros_sht_pool_init
    public :: aeros_sht_pool_end
    public :: aeros_sht_pool_get
    public :: aeros_sht_analysis
    public :: aeros_sht_synthesis
    public :: aeros_sht_analysis_vec
    public :: aeros_sht_synthesis_vec
    public :: aeros_sht_gradient
    public :: aeros_sht_laplacian
    public :: aeros_sht_surface_integral
    public :: aeros_sht_lm
    ! extensions
    public :: aeros_sht_set_precision
    public :: aeros_sht_synthesis_levels
    public :: aeros_sht_analysis_levels

contains

    subroutine check_rc(rc, where)
        ! Fatal on any library failure, with the library's own message.
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where

        if (rc == AER_OK) return
        write(io_unit_err,*) trim(where)//":: error: ", aer_c_string(aer_strerror(rc)), &
                                " -- ", aer_c_string(aer_last_error())
        error stop 1
    end subroutine check_rc

    subroutine aeros_sht_init(sht, trunc, nlon, nlat, nthreads, quick, cache)
        ! Configure the truncation and its Gaussian grid on the current GPU
        ! (reference src/aeros_spectral.f90:147-279).

        type(aeros_sht_class), intent(inout) :: sht
        integer, intent(in) :: trunc
        integer, intent(in), optional :: nlon, nlat
        integer, intent(in), optional :: nthreads
        logical, intent(in), optional :: quick
        logical, intent(in), optional :: cache

        type(aer_sht_info) :: info
        integer(c_int), allocatable :: li(:), mi(:)
        real(c_double), allocatable :: wts_half(:)
        integer :: nlon_, nlat_, layout, nh, i
        logical :: quick_, cache_

        quick_ = .FALSE.
        if (present(quick)) quick_ = quick
        cache_ = .FALSE.
        if (present(cache)) cache_ = cache
        if (present(nthreads)) then
            continue   ! no library-internal threading to configure
        end if

        call aeros_sht_grid_size(trunc, nlon_, nlat_)
        if (present(nlon)) then
            if (nlon > 0) nlon_ = nlon
        end if
        if (present(nlat)) then
            if (nlat > 0) nlat_ = nlat
        end if

        call aeros_sht_end(sht)

        sht%trunc = trunc
        sht%lmax  = trunc
        sht%mmax  = trunc
        sht%mres  = 1

        layout = AER_SHT_PHI_CONTIGUOUS
        if (quick_) then
            layout = layout + AER_SHT_QUICK_INIT
        else
            layout = layout + AER_SHT_GAUSS
        end if
        if (cache_) layout = layout + AER_SHT_LOAD_SAVE_CFG

        call check_rc(aer_sht_create(int(sht%lmax, c_int), int(sht%mmax, c_int), 1_c_int, &
                        AER_SHT_ORTHONORMAL + AER_SHT_NO_CS_PHASE, sht%cfg), "aeros_sht_init")
        call check_rc(aer_sht_set_grid(sht%cfg, int(layout, c_int), 1.0e-10_c_double, &
                        int(nlat_, c_int), int(nlon_, c_int)), "aeros_sht_init")

        call check_rc(aer_sht_get_info(sht%cfg, info), "aeros_sht_init")
        sht%nlat = info%nlat
        sht%nlon = info%nphi
        sht%nlm  = info%nlm

        if (info%nspat /= sht%nlon*sht%nlat) then
            write(io_unit_err,*) "aeros_sht_init:: error: padded spatial layout, nspat = ", &
                                    info%nspat, " /= nlon*nlat = ", sht%nlon*sht%nlat
            stop 1
        end if

        allocate(li(sht%nlm), mi(sht%nlm))
        call check_rc(aer_sht_get_lm(sht%cfg, li, mi), "aeros_sht_init")
        allocate(sht%l_of_lm(sht%nlm), sht%m_of_lm(sht%nlm))
        sht%l_of_lm = int(li)
        sht%m_of_lm = int(mi)

        allocate(sht%sinlat(sht%nlat), sht%colat(sht%nlat))
        call check_rc(aer_sht_get_cos_theta(sht%cfg, sht%sinlat), "aeros_sht_init")
        sht%colat = acos(sht%sinlat)

        allocate(sht%lon(sht%nlon))
        do i = 1, sht%nlon
            sht%lon(i) = 2.0_dp*pi*real(i-1, dp)/real(sht%nlon, dp)
        end do

        nh = info%nlat_2
        allocate(wts_half(nh))
        call check_rc(aer_sht_gauss_wts(sht%cfg, wts_half), "aeros_sht_init")
        allocate(sht%gauss_w(sht%nlat))
        sht%gauss_w(1:nh) = wts_half
        sht%gauss_w(sht%nlat:sht%nlat-nh+1:-1) = wts_half

    end subroutine aeros_sht_init

    subroutine aeros_sht_end(sht)
        type(aeros_sht_class), intent(inout) :: sht
        integer(c_int) :: rc

        if (c_associated(sht%cfg)) rc = aer_sht_destroy(sht%cfg)
        sht%cfg = c_null_ptr
        sht%trunc = 0; sht%lmax = 0; sht%mmax = 0
        sht%nlon  = 0; sht%nlat = 0; sht%nlm  = 0
        if (allocated(sht%l_of_lm)) deallocate(sht%l_of_lm)
        if (allocated(sht%m_of_lm)) deallocate(sht%m_of_lm)
        if (allocated(sht%sinlat))  deallocate(sht%sinlat)
        if (allocated(sht%colat))   deallocate(sht%colat)
        if (allocated(sht%lon))     deallocate(sht%lon)
        if (allocated(sht%gauss_w)) deallocate(sht%gauss_w)
    end subroutine aeros_sht_end

    subroutine aeros_sht_set_precision(sht, prec)
        ! AER_PREC_FP64 (default), AER_PREC_FP32 or AER_PREC_SPLIT for later transforms.
        type(aeros_sht_class), intent(in) :: sht
        integer, intent(in) :: prec
        call check_rc(aer_sht_set_precision(sht%cfg, int(prec, c_int)), "aeros_sht_set_precision")
    end subroutine aeros_sht_set_precision

    subroutine aeros_sht_pool_init(pool, trunc, nlon, nlat, nthreads, quick, cache)
        ! One handle per thread, created serially (reference src/aeros_spectral.f90:305-342).
        type(aeros_sht_pool_class), intent(inout) :: pool
        integer, intent(in) :: trunc
        integer, intent(in), optional :: nlon, nlat
        integer, intent(in), optional :: nthreads
        logical, intent(in), optional :: quick
        logical, intent(in), optional :: cache
        integer :: nt, i

        nt = 1
        !$ nt = omp_get_max_threads()
        if (present(nthreads)) then
            if (nthreads > 0) nt = nthreads
        end if
        call aeros_sht_pool_end(pool)
        pool%nthreads = nt
        allocate(pool%sht(nt))
        do i = 1, nt
            call aeros_sht_init(pool%sht(i), trunc, nlon=nlon, nlat=nlat, nthreads=1, &
                                quick=quick, cache=cache)
        end do
    end subroutine aeros_sht_pool_init

    subroutine aeros_sht_pool_end(pool)
        type(aeros_sht_pool_class), intent(inout) :: pool
        integer :: i
        if (allocated(pool%sht)) then
            do i = 1, size(pool%sht)
                call aeros_sht_end(pool%sht(i))
            end do
            deallocate(pool%sht)
        end if
        pool%nthreads = 0
    end subroutine aeros_sht_pool_end

    function aeros_sht_pool_get(pool) result(sht)
        ! The calling thread's handle (reference src/aeros_spectral.f90:365-397).
        type(aeros_sht_pool_class), intent(in), target :: pool
        type(aeros_sht_class), pointer :: sht
        integer :: tid

        tid = 1
        !$ tid = omp_get_thread_num() + 1
        if (tid > pool%nthreads) then
            write(io_unit_err,*) "aeros_sht_pool_get:: error: thread ", tid, &
                                    " exceeds pool size ", pool%nthreads
            error stop 1
        end if
        sht => pool%sht(tid)
    end function aeros_sht_pool_get

    subroutine aeros_sht_grid_size(trunc, nlon, nlat)
        ! Smallest quadratically unaliased Gaussian grid (src/aeros_spectral.f90:399-438).
        integer, intent(in)  :: trunc
        integer, intent(out) :: nlon, nlat
        integer(c_int) :: a, b
        call check_rc(aer_sht_grid_size(int(trunc, c_int), a, b), "aeros_sht_grid_size")
        nlon = a
        nlat = b
    end subroutine aeros_sht_grid_size

    integer function aeros_sht_lm(sht, l, m) result(lm)
        ! 1-based index of (l, m); -1 outside the truncation (src/aeros_spectral.f90:460-482).
        type(aeros_sht_class), intent(in) :: sht
        integer, intent(in) :: l, m

        if (m < 0 .or. m > sht%mmax*sht%mres .or. l < m .or. l > sht%lmax &
                .or. mod(m, sht%mres) /= 0) then
            lm = -1
            return
        end if
        lm = ((m/sht%mres)*(2*sht%lmax + 2 - (m + sht%mres)))/2 + l + 1
    end function aeros_sht_lm

    ! === Scalar transforms ===================================================

    subroutine aeros_sht_analysis(sht, field, coeffs)
        ! Spatial field (nlon,nlat) -> spectral coefficients (nlm); field preserved.
        type(aeros_sht_class), intent(in) :: sht
        real(dp), intent(inout), contiguous, target :: field(:,:)
        complex(wp_sh), intent(out), contiguous, target :: coeffs(:)
        call check_rc(aer_spat_to_SH(sht%cfg, c_loc(field), c_loc(coeffs)), "aeros_sht_analysis")
    end subroutine aeros_sht_analysis

    subroutine aeros_sht_synthesis(sht, coeffs, field)
        ! Spectral coefficients (nlm) -> spatial field (nlon,nlat).
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous, target :: coeffs(:)
        real(dp), intent(out), contiguous, target :: field(:,:)
        call check_rc(aer_SH_to_spat(sht%cfg, c_loc(coeffs), c_loc(field)), "aeros_sht_synthesis")
    end subroutine aeros_sht_synthesis

    ! === Vector transforms ===================================================
    ! (v_theta, v_phi) = grad_h(S) + curl_h(T) on the unit sphere, v_theta positive
    ! southward (reference src/aeros_spectral.f90:517-528).

    subroutine aeros_sht_analysis_vec(sht, vth, vph, slm, tlm)
        type(aeros_sht_class), intent(in) :: sht
        real(dp), intent(inout), contiguous, target :: vth(:,:), vph(:,:)
        complex(wp_sh), intent(out), contiguous, target :: slm(:), tlm(:)
        call check_rc(aer_spat_to_SHsphtor(sht%cfg, c_loc(vth), c_loc(vph), c_loc(slm), c_loc(tlm)), &
                        "aeros_sht_analysis_vec")
    end subroutine aeros_sht_analysis_vec

    subroutine aeros_sht_synthesis_vec(sht, slm, tlm, vth, vph)
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous, target :: slm(:), tlm(:)
        real(dp), intent(out), contiguous, target :: vth(:,:), vph(:,:)
        call check_rc(aer_SHsphtor_to_spat(sht%cfg, c_loc(slm), c_loc(tlm), c_loc(vth), c_loc(vph)), &
                        "aeros_sht_synthesis_vec")
    end subroutine aeros_sht_synthesis_vec

    subroutine aeros_sht_gradient(sht, coeffs, gth, gph)
        ! gth = dF/dtheta, gph = (1/sin theta) dF/dphi on the unit sphere.
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous, target :: coeffs(:)
        real(wp), intent(out), contiguous, target :: gth(:,:)
        real(wp), intent(out), contiguous, target :: gph(:,:)
        call check_rc(aer_SHsph_to_spat(sht%cfg, c_loc(coeffs), c_loc(gth), c_loc(gph)), &
                        "aeros_sht_gradient")
    end subroutine aeros_sht_gradient

    ! === Level-batched transforms (extension) ================================

    subroutine aeros_sht_synthesis_levels(sht, coeffs, field)
        ! coeffs(nlm, nlev) -> field(nlon, nlat, nlev): every level in one launch.
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(in), contiguous, target :: coeffs(:,:)
        real(dp), intent(out), contiguous, target :: field(:,:,:)
        type(c_ptr) :: dq, dg
        integer(c_size_t) :: bq, bg
        integer :: nlev

        nlev = size(coeffs, 2)
        bq = int(16, c_size_t)*sht%nlm*nlev
        bg = int(8, c_size_t)*sht%nlon*sht%nlat*nlev
        call check_rc(aer_dev_malloc(bq, dq), "aeros_sht_synthesis_levels")
        call check_rc(aer_dev_malloc(bg, dg), "aeros_sht_synthesis_levels")
        call check_rc(aer_memcpy_h2d(dq, c_loc(coeffs), bq), "aeros_sht_synthesis_levels")
        call check_rc(aer_synth_d(sht%cfg, int(nlev, c_int), dq, int(2*sht%nlm, c_long), dg, &
                        int(sht%nlon*sht%nlat, c_long)), "aeros_sht_synthesis_levels")
        call check_rc(aer_sht_sync(sht%cfg), "aeros_sht_synthesis_levels")
        call check_rc(aer_memcpy_d2h(c_loc(field), dg, bg), "aeros_sht_synthesis_levels")
        call check_rc(aer_dev_free(dq), "aeros_sht_synthesis_levels")
        call check_rc(aer_dev_free(dg), "aeros_sht_synthesis_levels")
    end subroutine aeros_sht_synthesis_levels

    subroutine aeros_sht_analysis_levels(sht, field, coeffs)
        ! field(nlon, nlat, nlev) -> coeffs(nlm, nlev): every level in one launch.
        type(aeros_sht_class), intent(in) :: sht
        real(dp), intent(in), contiguous, target :: field(:,:,:)
        complex(wp_sh), intent(out), contiguous, target :: coeffs(:,:)
        type(c_ptr) :: dq, dg
        integer(c_size_t) :: bq, bg
        integer :: nlev

        nlev = size(field, 3)
        bq = int(16, c_size_t)*sht%nlm*nlev
        bg = int(8, c_size_t)*sht%nlon*sht%nlat*nlev
        call check_rc(aer_dev_malloc(bq, dq), "aeros_sht_analysis_levels")
        call check_rc(aer_dev_malloc(bg, dg), "aeros_sht_analysis_levels")
        call check_rc(aer_memcpy_h2d(dg, c_loc(field), bg), "aeros_sht_analysis_levels")
        call check_rc(aer_anal_d(sht%cfg, int(nlev, c_int), dg, int(sht%nlon*sht%nlat, c_long), dq, &
                        int(2*sht%nlm, c_long)), "aeros_sht_analysis_levels")
        call check_rc(aer_sht_sync(sht%cfg), "aeros_sht_analysis_levels")
        call check_rc(aer_memcpy_d2h(c_loc(coeffs), dq, bq), "aeros_sht_analysis_levels")
        call check_rc(aer_dev_free(dq), "aeros_sht_analysis_levels")
        call check_rc(aer_dev_free(dg), "aeros_sht_analysis_levels")
    end subroutine aeros_sht_analysis_levels

    ! === Spectral operators (exact, on the host) =============================

    subroutine aeros_sht_laplacian(sht, coeffs, inverse)
        ! lap(Y_lm) = -l(l+1)/a^2 Y_lm; with inverse=.TRUE. l = 0 is set to zero
        ! (src/aeros_spectral.f90:588-630, same arithmetic so the result is exact).
        type(aeros_sht_class), intent(in) :: sht
        complex(wp_sh), intent(inout) :: coeffs(:)
        logical, intent(in), optional :: inverse
        real(dp) :: fac, a2
        integer  :: lm, l
        logical  :: inv

        inv = .FALSE.
        if (present(inverse)) inv = inverse
        a2 = r_earth*r_earth
        do lm = 1, sht%nlm
            l = sht%l_of_lm(lm)
            if (inv) then
                if (l == 0) then
                    coeffs(lm) = (0.0_wp_sh, 0.0_wp_sh)
                    cycle
                end if
                fac = -a2/real(l*(l+1), dp)
            else
                fac = -real(l*(l+1), dp)/a2
            end if
            coeffs(lm) = coeffs(lm)*fac
        end do
    end subroutine aeros_sht_laplacian

    real(dp) function aeros_sht_surface_integral(sht, field) result(total)
        ! a^2 sum_j w_j (2 pi/nlon) sum_i f_ij, row sums first, in dp
        ! (src/aeros_spectral.f90:632-666).
        type(aeros_sht_class), intent(in) :: sht
        real(dp), intent(in) :: field(:,:)
        real(dp) :: dlon, row
        integer  :: i, j

        dlon  = 2.0_dp*pi/real(sht%nlon, dp)
        total = 0.0_dp
        do j = 1, sht%nlat
            row = 0.0_dp
            do i = 1, sht%nlon
                row = row + field(i,j)
            end do
            total = total + sht%gauss_w(j)*row*dlon
        end do
        total = total*r_earth*r_earth
    end function aeros_sht_surface_integral

end module aeros_spectral
