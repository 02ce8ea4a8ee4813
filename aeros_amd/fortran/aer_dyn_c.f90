module aer_dyn_c
    ! iso_c_binding interfaces to the dynamics-step part of the C ABI (include/aer_sht.h):
    ! the vertical grid, the column terms with Held-Suarez forcing, the semi-implicit
    ! leapfrog completion, the mass fixer, the zonal-mean statistics and the step
    ! adjoints. A Fortran time stepper in the shape of the reference's
    ! aeros_timestep_step (src/dynamics/aeros_timestep.f90:776-1137) calls these on
    ! device arrays between aer_tendency_grid_transforms_d and aer_tendency_spectral_d.
    !
    ! Every array argument is a device address (type(c_ptr), by value), as for the
    ! batched transform entry points; host arrays go through aer_memcpy_h2d/d2h.

    use, intrinsic :: iso_c_binding

    implicit none

    private

    integer(c_int), parameter, public :: AER_STEP_SEMI_IMPLICIT = 1_c_int
    integer(c_int), parameter, public :: AER_STEP_DIFFUSE = 2_c_int
    integer(c_int), parameter, public :: AER_STEP_FILTER = 4_c_int

    public :: aer_vgrid_create, aer_vgrid_destroy
    public :: aer_column_terms_d, aer_column_terms_dteq_d, aer_column_terms_adj_d
    public :: aer_semiimp_create, aer_semiimp_set_step, aer_semiimp_set_diffusion, aer_semiimp_destroy
    public :: aer_leapfrog_d, aer_leapfrog_adj_d, aer_mass_d
    public :: aer_diag_create, aer_diag_accumulate_d, aer_diag_read, aer_diag_reset, aer_diag_destroy

    interface
        ! ---- vertical grid (host A/B tables, nlev+1 half levels, top first)
        function aer_vgrid_create(nlev, a_half, b_half, vg) bind(C, name="aer_vgrid_create") result(rc)
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: nlev
            real(c_double), intent(in) :: a_half(*), b_half(*)
            type(c_ptr), intent(out) :: vg
            integer(c_int) :: rc
        end function
        function aer_vgrid_destroy(vg) bind(C, name="aer_vgrid_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: vg
            integer(c_int) :: rc
        end function

        ! ---- column terms (+ Held-Suarez forcing); phis may be c_null_ptr
        function aer_column_terms_d(cfg, vg, held_suarez, u, v, vor_g, div_g, t_g, dtdx, dtdy, &
                lnps_g, dlnpsdx, dlnpsdy, phis, ae, an, ek, dtdt, dlnpsdt) &
                bind(C, name="aer_column_terms_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vg, u, v, vor_g, div_g, t_g, dtdx, dtdy, lnps_g, dlnpsdx, dlnpsdy
            type(c_ptr), value :: phis, ae, an, ek, dtdt, dlnpsdt
            integer(c_int), value :: held_suarez
            integer(c_int) :: rc
        end function
        function aer_column_terms_dteq_d(cfg, vg, dteq, u, v, vor_g, div_g, t_g, dtdx, dtdy, &
                lnps_g, dlnpsdx, dlnpsdy, phis, ae, an, ek, dtdt, dlnpsdt) &
                bind(C, name="aer_column_terms_dteq_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vg, dteq, u, v, vor_g, div_g, t_g, dtdx, dtdy, lnps_g, dlnpsdx
            type(c_ptr), value :: dlnpsdy, phis, ae, an, ek, dtdt, dlnpsdt
            integer(c_int) :: rc
        end function
        function aer_column_terms_adj_d(cfg, vg, held_suarez, dteq, u, v, vor_g, div_g, t_g, dtdx, dtdy, &
                lnps_g, dlnpsdx, dlnpsdy, phis, b_ae, b_an, b_ek, b_dtdt, b_dlnpsdt, &
                g_u, g_v, g_vor, g_div, g_t, g_dtdx, g_dtdy, g_lnps, g_dlnpsdx, g_dlnpsdy, g_teq) &
                bind(C, name="aer_column_terms_adj_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg, vg, dteq, u, v, vor_g, div_g, t_g, dtdx, dtdy, lnps_g, dlnpsdx
            type(c_ptr), value :: dlnpsdy, phis, b_ae, b_an, b_ek, b_dtdt, b_dlnpsdt
            type(c_ptr), value :: g_u, g_v, g_vor, g_div, g_t, g_dtdx, g_dtdy, g_lnps, g_dlnpsdx, g_dlnpsdy
            type(c_ptr), value :: g_teq
            integer(c_int), value :: held_suarez
            integer(c_int) :: rc
        end function

        ! ---- semi-implicit leapfrog completion (aeros_semiimp_class)
        function aer_semiimp_create(cfg, vg, t_ref, ps_ref, dt_step, alpha, si) &
                bind(C, name="aer_semiimp_create") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: cfg, vg
            real(c_double), value :: t_ref, ps_ref, dt_step, alpha
            type(c_ptr), intent(out) :: si
            integer(c_int) :: rc
        end function
        function aer_semiimp_set_step(si, dt_step) bind(C, name="aer_semiimp_set_step") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: si
            real(c_double), value :: dt_step
            integer(c_int) :: rc
        end function
        function aer_semiimp_set_diffusion(si, tau_diff, tau_diff_div, ndiff) &
                bind(C, name="aer_semiimp_set_diffusion") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: si
            real(c_double), value :: tau_diff, tau_diff_div
            integer(c_int), value :: ndiff
            integer(c_int) :: rc
        end function
        function aer_semiimp_destroy(si) bind(C, name="aer_semiimp_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: si
            integer(c_int) :: rc
        end function
        function aer_leapfrog_d(si, flags, eps_filter, raw_alpha, old_vor, old_div, old_temp, old_lnps, &
                now_vor, now_div, now_temp, now_lnps, tnd_vor, tnd_div, tnd_temp, tnd_lnps, &
                new_vor, new_div, new_temp, new_lnps) bind(C, name="aer_leapfrog_d") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: si, old_vor, old_div, old_temp, old_lnps, now_vor, now_div, now_temp
            type(c_ptr), value :: now_lnps, tnd_vor, tnd_div, tnd_temp, tnd_lnps
            type(c_ptr), value :: new_vor, new_div, new_temp, new_lnps
            integer(c_int), value :: flags
            real(c_double), value :: eps_filter, raw_alpha
            integer(c_int) :: rc
        end function
        function aer_leapfrog_adj_d(si, flags, eps_filter, raw_alpha, bn_vor, bn_div, bn_temp, bn_lnps, &
                bc_vor, bc_div, bc_temp, bc_lnps, go_vor, go_div, go_temp, go_lnps, &
                gc_vor, gc_div, gc_temp, gc_lnps, gt_vor, gt_div, gt_temp, gt_lnps) &
                bind(C, name="aer_leapfrog_adj_d") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: si, bn_vor, bn_div, bn_temp, bn_lnps, bc_vor, bc_div, bc_temp, bc_lnps
            type(c_ptr), value :: go_vor, go_div, go_temp, go_lnps, gc_vor, gc_div, gc_temp, gc_lnps
            type(c_ptr), value :: gt_vor, gt_div, gt_temp, gt_lnps
            integer(c_int), value :: flags
            real(c_double), value :: eps_filter, raw_alpha
            integer(c_int) :: rc
        end function
        function aer_mass_d(cfg, lnps, fix, mass_target, mass_out, lnr_out) bind(C, name="aer_mass_d") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: cfg, lnps, mass_out, lnr_out
            integer(c_int), value :: fix
            real(c_double), value :: mass_target
            integer(c_int) :: rc
        end function

        ! ---- zonal-mean statistics (aeros_diag_class)
        function aer_diag_create(cfg, nlev, diag) bind(C, name="aer_diag_create") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg
            integer(c_int), value :: nlev
            type(c_ptr), intent(out) :: diag
            integer(c_int) :: rc
        end function
        function aer_diag_accumulate_d(diag, u, v, temp, lnps_g) bind(C, name="aer_diag_accumulate_d") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: diag, u, v, temp, lnps_g
            integer(c_int) :: rc
        end function
        function aer_diag_read(diag, nsample, mean) bind(C, name="aer_diag_read") result(rc)
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: diag
            integer(c_int), intent(out) :: nsample
            real(c_double), intent(out) :: mean(*)
            integer(c_int) :: rc
        end function
        function aer_diag_reset(diag) bind(C, name="aer_diag_reset") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: diag
            integer(c_int) :: rc
        end function
        function aer_diag_destroy(diag) bind(C, name="aer_diag_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: diag
            integer(c_int) :: rc
        end function
    end interface

end module aer_dyn_c
