module aer_sht_c
    ! iso_c_binding interfaces to the C ABI of the gfx950 transform library
    ! (include/aer_sht.h). This takes the place of the SHTns Fortran header the
    ! reference includes (`include 'shtns.f03'`, reference src/aeros_spectral.f90:41-45).
    !
    ! Spectral arrays cross as complex(c_double_complex), grid arrays as
    ! real(c_double), both by address with no copy on the Fortran side.

    use, intrinsic :: iso_c_binding

    implicit none

    private

    integer(c_int), parameter, public :: AER_OK = 0
    integer(c_int), parameter, public :: AER_SHT_ORTHONORMAL    = int(z'1', c_int)
    integer(c_int), parameter, public :: AER_SHT_NO_CS_PHASE    = int(z'2', c_int)
    integer(c_int), parameter, public :: AER_SHT_PHI_CONTIGUOUS = int(z'10', c_int)
    integer(c_int), parameter, public :: AER_SHT_GAUSS          = int(z'20', c_int)
    integer(c_int), parameter, public :: AER_SHT_QUICK_INIT     = int(z'40', c_int)
    integer(c_int), parameter, public :: AER_SHT_LOAD_SAVE_CFG  = int(z'80', c_int)
    integer(c_int), parameter, public :: AER_PREC_FP64 = 0, AER_PREC_FP32 = 1, AER_PREC_SPLIT = 2

    type, bind(C), public :: aer_sht_info
        integer(c_int) :: lmax, mmax, mres, nlm, nlat, nphi, nspat, nlat_2
    end type aer_sht_info

    public :: aer_strerror, aer_last_error
    public :: aer_sht_grid_size, aer_sht_create, aer_sht_set_grid, aer_sht_get_info
    public :: aer_sht_get_lm, aer_sht_get_cos_theta, aer_sht_gauss_wts, aer_sht_destroy
    public :: aer_sht_set_precision, aer_sht_set_stream, aer_sht_sync
    public :: aer_spat_to_SH, aer_SH_to_spat, aer_spat_to_SHsphtor, aer_SHsphtor_to_spat
    public :: aer_SHsph_to_spat
    public :: aer_synth_d, aer_anal_d, aer_synth_vec_d, aer_anal_vec_d, aer_synth_grad_d
    public :: aer_uv_from_vordiv_d, aer_vordiv_from_uv_d, aer_gradient_d
    public :: aer_tendency_grid_transforms_d, aer_tendency_spectral_d
    public :: aer_dev_malloc, aer_dev_free, aer_memcpy_h2d, aer_memcpy_d2h
    public :: aer_c_string

    interface
        function aer_strerror(status) bind(C, name="aer_strerror") result(msg)
            import :: c_int, c_ptr
            integer(c_int), value :: status
            type(c_ptr) :: msg
        end function
        function aer_last_error() bind(C, name="aer_last_error") result(msg)
            import :: c_ptr
            type(c_ptr) :: msg
        end function
        function aer_sht_grid_size(trunc, nlon, nlat) bind(C, name="aer_sht_grid_size") result(rc)
            import :: c_int
            integer(c_int), value :: trunc
            integer(c_int), intent(out) :: nlon, nlat
            integer(c_int) :: rc
        end function
        function aer_sht_create(lmax, mmax, mres, norm, cfg) bind(C, name="aer_sht_create") result(rc)
            import :: c_int, c_ptr
            integer(c_int), value :: lmax, mmax, mres, norm
            type(c_ptr), intent(out) :: cfg
            integer(c_int) :: rc
        end function
        function aer_sht_set_grid(cfg, layout, eps, nlat, nphi) bind(C, name="aer_sht_set_grid") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cfg
            integer(c_int), value :: layout, nlat, nphi
            real(c_double), value :: eps
            integer(c_int) :: rc
        end function
        function aer_sht_get_info(cfg, info) bind(C, name="aer_sht_get_info") result(rc)
            import :: c_int, c_ptr, aer_sht_info
            type(c_ptr), value :: cfg
            type(aer_sht_info), intent(out) :: info
            integer(c_int) :: rc
        end function
        function aer_sht_get_lm(cfg, li, mi) bind(C, name="aer_sht_get_lm") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg
            integer(c_int), intent(out) :: li(*), mi(*)
            integer(c_int) :: rc
        end function
        function aer_sht_get_cos_theta(cfg, ct) bind(C, name="aer_sht_get_cos_theta") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cfg
            real(c_double), intent(out) :: ct(*)
            integer(c_int) :: rc
        end function
        function aer_sht_gauss_wts(cfg, w) bind(C, name="aer_sht_gauss_wts") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cfg
            real(c_double), intent(out) :: w(*)
            integer(c_int) :: rc
        end function
        function aer_sht_destroy(cfg) bind(C, name="aer_sht_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg
            integer(c_int) :: rc
        end function
        function aer_sht_set_precision(cfg, prec) bind(C, name="aer_sht_set_precision") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg
            integer(c_int), value :: prec
            integer(c_int) :: rc
        end function
        function aer_sht_set_stream(cfg, stream) bind(C, name="aer_sht_set_stream") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, stream
            integer(c_int) :: rc
        end function
        function aer_sht_sync(cfg) bind(C, name="aer_sht_sync") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg
            integer(c_int) :: rc
        end function

        ! ---- per-call transforms on host arrays
        function aer_spat_to_SH(cfg, vr, qlm) bind(C, name="aer_spat_to_SH") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vr, qlm
            integer(c_int) :: rc
        end function
        function aer_SH_to_spat(cfg, qlm, vr) bind(C, name="aer_SH_to_spat") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, qlm, vr
            integer(c_int) :: rc
        end function
        function aer_spat_to_SHsphtor(cfg, vt, vp, slm, tlm) bind(C, name="aer_spat_to_SHsphtor") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vt, vp, slm, tlm
            integer(c_int) :: rc
        end function
        function aer_SHsphtor_to_spat(cfg, slm, tlm, vt, vp) bind(C, name="aer_SHsphtor_to_spat") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, slm, tlm, vt, vp
            integer(c_int) :: rc
        end function
        function aer_SHsph_to_spat(cfg, slm, vt, vp) bind(C, name="aer_SHsph_to_spat") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, slm, vt, vp
            integer(c_int) :: rc
        end function

        ! ---- batched transforms on device arrays
        function aer_synth_d(cfg, nfield, qlm, qstride, vr, gstride) bind(C, name="aer_synth_d") result(rc)
            import :: c_int, c_ptr, c_long
            type(c_ptr), value :: cfg, qlm, vr
            integer(c_int), value :: nfield
            integer(c_long), value :: qstride, gstride
            integer(c_int) :: rc
        end function
        function aer_anal_d(cfg, nfield, vr, gstride, qlm, qstride) bind(C, name="aer_anal_d") result(rc)
            import :: c_int, c_ptr, c_long
            type(c_ptr), value :: cfg, vr, qlm
            integer(c_int), value :: nfield
            integer(c_long), value :: gstride, qstride
            integer(c_int) :: rc
        end function
        function aer_synth_vec_d(cfg, nfield, slm, tlm, qstride, vt, vp, gstride) &
                bind(C, name="aer_synth_vec_d") result(rc)
            import :: c_int, c_ptr, c_long
            type(c_ptr), value :: cfg, slm, tlm, vt, vp
            integer(c_int), value :: nfield
            integer(c_long), value :: qstride, gstride
            integer(c_int) :: rc
        end function
        function aer_anal_vec_d(cfg, nfield, vt, vp, gstride, slm, tlm, qstride) &
                bind(C, name="aer_anal_vec_d") result(rc)
            import :: c_int, c_ptr, c_long
            type(c_ptr), value :: cfg, vt, vp, slm, tlm
            integer(c_int), value :: nfield
            integer(c_long), value :: gstride, qstride
            integer(c_int) :: rc
        end function
        function aer_synth_grad_d(cfg, nfield, slm, qstride, gt, gp, gstride) &
                bind(C, name="aer_synth_grad_d") result(rc)
            import :: c_int, c_ptr, c_long
            type(c_ptr), value :: cfg, slm, gt, gp
            integer(c_int), value :: nfield
            integer(c_long), value :: qstride, gstride
            integer(c_int) :: rc
        end function
        function aer_uv_from_vordiv_d(cfg, nlev, vor, div, u, v) bind(C, name="aer_uv_from_vordiv_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vor, div, u, v
            integer(c_int), value :: nlev
            integer(c_int) :: rc
        end function
        function aer_vordiv_from_uv_d(cfg, nlev, u, v, vor, div) bind(C, name="aer_vordiv_from_uv_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, u, v, vor, div
            integer(c_int), value :: nlev
            integer(c_int) :: rc
        end function
        function aer_gradient_d(cfg, nlev, f, dfdx, dfdy) bind(C, name="aer_gradient_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, f, dfdx, dfdy
            integer(c_int), value :: nlev
            integer(c_int) :: rc
        end function
        function aer_tendency_grid_transforms_d(cfg, nlev, vor, div, temp, lnps, u, v, vor_g, div_g, &
                t_g, dtdx, dtdy, lnps_g, dlnpsdx, dlnpsdy) &
                bind(C, name="aer_tendency_grid_transforms_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vor, div, temp, lnps, u, v, vor_g, div_g, t_g, dtdx, dtdy
            type(c_ptr), value :: lnps_g, dlnpsdx, dlnpsdy
            integer(c_int), value :: nlev
            integer(c_int) :: rc
        end function
        function aer_tendency_spectral_d(cfg, nlev, ae, an, ek, dtdt, dlnpsdt, tvor, tdiv, ttemp, tlnps) &
                bind(C, name="aer_tendency_spectral_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, ae, an, ek, dtdt, dlnpsdt, tvor, tdiv, ttemp, tlnps
            integer(c_int), value :: nlev
            integer(c_int) :: rc
        end function

        ! ---- device memory for Fortran hosts
        function aer_dev_malloc(bytes, ptr) bind(C, name="aer_dev_malloc") result(rc)
            import :: c_int, c_ptr, c_size_t
            integer(c_size_t), value :: bytes
            type(c_ptr), intent(out) :: ptr
            integer(c_int) :: rc
        end function
        function aer_dev_free(ptr) bind(C, name="aer_dev_free") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ptr
            integer(c_int) :: rc
        end function
        function aer_memcpy_h2d(dst, src, bytes) bind(C, name="aer_memcpy_h2d") result(rc)
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
            integer(c_int) :: rc
        end function
        function aer_memcpy_d2h(dst, src, bytes) bind(C, name="aer_memcpy_d2h") result(rc)
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
            integer(c_int) :: rc
        end function
    end interface

contains

    function aer_c_string(p) result(s)
        ! Copy a NUL-terminated C string into a Fortran string.
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: cs(:)
        integer :: n

        s = ""
        if (.not. c_associated(p)) return
        call c_f_pointer(p, cs, [4096])
        n = 0
        do while (n < 4096)
            if (cs(n+1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: s)
        s = transfer(cs(1:n), s)
    end function aer_c_string

end module aer_sht_c
