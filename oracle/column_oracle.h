/*
 * column_oracle.h -- FP64 CPU restatement of the grid-space column work that
 * sits between the two transform halves of an aeros dynamics step: the hybrid
 * vertical coordinate, the Simmons-Burridge column terms and the Held-Suarez
 * forcing. TEST INFRASTRUCTURE ONLY (see sht_oracle.h): the checker for the
 * HIP column kernel, never linked by the product.
 *
 * Layout: 3-D grid fields are [nlev][nlat][nlon] (a level contiguous, k = 0 at
 * the model top), 2-D fields [nlat][nlon], latitude north -> south; half-level
 * arrays have nlev+1 entries, index 0 at the top
 * (reference src/dynamics/aeros_vertical.f90:24-31).
 */
#ifndef AEROS_COLUMN_ORACLE_H
#define AEROS_COLUMN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants (reference src/aeros_defs.f90:67-103). */
#define ORC_R_D     287.05
#define ORC_CP_D    1004.64
#define ORC_KAPPA   (ORC_R_D / ORC_CP_D)
#define ORC_OMEGA   7.292e-5
#define ORC_P0      1.0e5
#define ORC_MV      (-9999.0)

/* Hybrid coefficients A (Pa), B (-) at the nlev+1 half levels, generated from
 * the stretching parameters (src/dynamics/aeros_vertical.f90:126-298, the
 * generator branch). Returns 0, or -1 on an invalid parameter. */
int    orc_vgrid_init(int nlev, double stretch_a, double stretch_r, double sigma_t,
                      double p_top, double ps_ref, double *A, double *B);

/* One column (src/dynamics/aeros_vertical.f90:344-474). */
void   orc_vgrid_pressure(int nlev, const double *A, const double *B, double ps,
                          double *p_half, double *p_full, double *dp);
void   orc_vgrid_alpha(int nlev, const double *p_half, double *alpha, double *dlnp);
void   orc_hydrostatic(int nlev, double phis, const double *temp, const double *p_half,
                       double *phi_full);

/* Held-Suarez eqs. (1)-(3) (src/physics/aeros_held_suarez.f90:225-293). */
double orc_hs_teq(double p, double sinlat2, double coslat2);
double orc_hs_kt(double sigma, double coslat4);
double orc_hs_kv(double sigma);

/* column_terms + vert_adv (src/dynamics/aeros_tendency.f90:475-630), then,
 * when hs != 0, aeros_hs_apply (src/physics/aeros_held_suarez.f90:177-223).
 * sinlat[nlat] is sin(latitude) = cos(colatitude) of the Gauss rows. */
void   orc_column_terms(int nlon, int nlat, int nlev, const double *A, const double *B,
                        const double *sinlat,
                        const double *u, const double *v, const double *vor_g,
                        const double *div_g, const double *t_g, const double *dtdx,
                        const double *dtdy, const double *lnps_g, const double *dlnpsdx,
                        const double *dlnpsdy, const double *phis,
                        double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt,
                        int hs);

#ifdef __cplusplus
}
#endif
#endif
