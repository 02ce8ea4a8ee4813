module aeros_amd_kinds
    ! The kinds and constants the transform modules need, with the reference's
    ! values (src/aeros_defs.f90:34-62, 83-92). Inside the aeros tree these names
    ! come from aeros_defs instead; this module is the only line to change.

    use, intrinsic :: iso_c_binding, only : c_double, c_float
    use, intrinsic :: iso_fortran_env, only : error_unit

    implicit none

    integer, parameter, public :: sp    = c_float
    integer, parameter, public :: dp    = c_double
    integer, parameter, public :: wp    = dp     ! working precision of grid fields
    integer, parameter, public :: wp_sh = dp     ! spectral coefficients (== c_double, no copy at the C boundary)

    real(dp), parameter, public :: pi      = 3.14159265358979323846264338327950288_dp
    real(dp), parameter, public :: r_earth = 6.371e6_dp      ! Earth mean radius [m]
    real(dp), parameter, public :: omega   = 7.292e-5_dp     ! Earth angular velocity [rad s-1]

    integer, parameter, public :: io_unit_err = error_unit

end module aeros_amd_kinds
