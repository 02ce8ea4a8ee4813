/*
 * step_oracle.h -- FP64 CPU restatement of the spectral step completion of an
 * aeros leapfrog step: the semi-implicit gravity-wave solve, the vorticity
 * leapfrog, implicit horizontal diffusion, the Robert-Asselin-Williams filter
 * and the mass fixer. TEST INFRASTRUCTURE ONLY (see sht_oracle.h).
 *
 * Spectral 3-D fields are complex [nlev][nlm] stored as interleaved doubles;
 * every operator here is real, so the real and imaginary parts are independent
 * columns. l_of_lm[nlm] maps the m-major index to the degree.
 */
#ifndef AEROS_STEP_ORACLE_H
#define AEROS_STEP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* The linearized operators and LU factors (src/dynamics/aeros_semiimp.f90:130-160). */
typedef struct orc_si {
    int nlev, lmax;
    double alpha, dt_step;
    double *dp_ref, *alpha_ref, *dlnp_ref;   /* nlev */
    double *gmat, *tmat, *wmat;              /* nlev x nlev, row-major [k][j] */
    double *nu, *pcoef;                      /* nlev */
    double *lu;                              /* (lmax+1) x nlev x nlev, row-major */
    int *piv;                                /* (lmax+1) x nlev */
} orc_si;

/* aeros_semiimp_init (src/dynamics/aeros_semiimp.f90:171-276): isothermal t_ref. */
orc_si *orc_si_create(int nlev, const double *A, const double *B, double t_ref, double ps_ref,
                      int lmax, double dt_step, double alpha);
void    orc_si_destroy(orc_si *si);
/* aeros_semiimp_set_step (:301-343). */
int     orc_si_set_step(orc_si *si, double dt_step);
/* aeros_semiimp_linear (:345-460). */
void    orc_si_linear(const orc_si *si, int nlm, const int *l_of_lm,
                      const double *div, const double *temp, const double *lnps,
                      double *ldiv, double *ltemp, double *lps);
/* aeros_semiimp_step (:462-571). */
void    orc_si_step(const orc_si *si, int nlm, const int *l_of_lm,
                    const double *old_div, const double *old_temp, const double *old_lnps,
                    const double *now_div, const double *now_temp, const double *now_lnps,
                    const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                    double *new_div, double *new_temp, double *new_lnps);
/* aeros_semiimp_gwspeed (:573-620). */
double  orc_si_gwspeed(const orc_si *si);

/* Implicit del^n diffusion factors without taper or adaptation
 * (src/dynamics/aeros_timestep.f90:560-563, 1673-1726): fac[l] =
 * 1/(1 + (h/tau) (l(l+1)/(lmax(lmax+1)))^(ndiff/2)). */
void    orc_diff_factors(int lmax, double h, double tau, int ndiff, double *fac);
/* new <- fac[l] new, per level (fac_vt for vor and temp, fac_div for div; each
 * [nlev][lmax+1]). */
void    orc_diffuse(int nlm, int nlev, int lmax, const int *l_of_lm, const double *fac_vt,
                    const double *fac_div, double *vor, double *div, double *temp);

/* raw_filter (src/dynamics/aeros_timestep.f90:1762-1815), in place on cur and
 * new, for one 3-D field of n real values (call once per field; lnps with n = 2 nlm). */
void    orc_raw_filter(long n, double eps, double raw_alpha, const double *old, double *cur,
                       double *nw);

#ifdef __cplusplus
}
#endif
#endif
