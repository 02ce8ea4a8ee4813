/*
 * sht_oracle.h -- FP64 CPU restatement of the aeros spherical-harmonic
 * transform path, used ONLY as test infrastructure: the checker in tests/,
 * the smoke check and the CPU-baseline leg of bench.py. Nothing in the
 * product (aeros_amd/, include/) links or calls it.
 *
 * Conventions (reference src/aeros_spectral.f90:17-26):
 *   orthonormal harmonics, no Condon-Shortley phase;
 *   grid (nlon, nlat), longitude contiguous, latitude north -> south;
 *   spectral m-major, l fastest, triangular lmax = mmax = T, mres = 1;
 *   complex coefficients stored as interleaved (re, im) doubles.
 *
 * Every function cites the reference line it restates. Which test pins which
 * function is listed in the header comment of sht_oracle.c.
 */
#ifndef AEROS_SHT_ORACLE_H
#define AEROS_SHT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg {
    int trunc, lmax, mmax, nlon, nlat, nlm;
    double *mu;      /* cos(colatitude) of the Gauss nodes, north -> south (nlat) */
    double *w;       /* Gauss weights, sum 2 (nlat) */
    double *sint;    /* sin(colatitude) (nlat) */
    double *P;       /* normalized P_lm(mu_j): [nlm][nlat] */
    double *D;       /* dP_lm/dtheta:          [nlm][nlat] */
    int *l_of_lm, *m_of_lm;   /* 0-based lm -> l, m */
    int fft_direct;  /* 1: exact O(N^2) DFT instead of the mixed-radix FFT */
    double *tw_fwd, *tw_inv;  /* e^{-+2 pi i k/nlon}, k < nlon (complex interleaved) */
} orc_cfg;

/* Physical constants used by the dynamics mapping (src/aeros_defs.f90:83-92). */
#define ORC_R_EARTH 6.371e6

int    orc_grid_size(int trunc, int *nlon, int *nlat);
int    orc_lm(int lmax, int l, int m);
void   orc_gauss(int n, double *mu, double *w);
void   orc_legendre(int lmax, int nlat, const double *mu, double *P, double *D);

orc_cfg *orc_create(int trunc, int nlon, int nlat);
void   orc_destroy(orc_cfg *c);
void   orc_set_fft_direct(orc_cfg *c, int direct);

/* Real DFT pieces (exposed so tests can compare them with numpy.fft). */
void   orc_rfft_forward(int n, int mmax, const double *x, double *X /* 2*(mmax+1) */);
void   orc_rfft_inverse(int n, int mmax, const double *X, double *x);

/* Scalar, vector and gradient transforms (SHTns calls at src/aeros_spectral.f90:496, 511, 539, 554, 580). */
void   orc_synth(const orc_cfg *c, const double *qlm, double *grid);
void   orc_anal(const orc_cfg *c, const double *grid, double *qlm);
void   orc_synth_vec(const orc_cfg *c, const double *slm, const double *tlm, double *vt, double *vp);
void   orc_anal_vec(const orc_cfg *c, const double *vt, const double *vp, double *slm, double *tlm);
void   orc_synth_grad(const orc_cfg *c, const double *slm, double *gt, double *gp);

/* Spectral operators (src/aeros_spectral.f90:588-666). */
void   orc_laplacian(const orc_cfg *c, double *qlm, int inverse);
double orc_surface_integral(const orc_cfg *c, const double *grid);

/* Dynamics mapping (src/dynamics/aeros_vordiv.f90:82-205). */
void   orc_uv_from_vordiv(const orc_cfg *c, const double *vor, const double *div, double *u, double *v);
void   orc_vordiv_from_uv(const orc_cfg *c, const double *u, const double *v, double *vor, double *div);
void   orc_gradient(const orc_cfg *c, const double *f, double *dfdx, double *dfdy);

/* Transform halves of the tendency (src/dynamics/aeros_tendency.f90:403-419, 449-470). */
void   orc_tendency_grid_transforms(const orc_cfg *c, int nlev,
            const double *vor, const double *div, const double *temp, const double *lnps,
            double *u, double *v, double *vor_g, double *div_g, double *t_g,
            double *dtdx, double *dtdy, double *lnps_g, double *dlnpsdx, double *dlnpsdy);
void   orc_tendency_spectral(const orc_cfg *c, int nlev,
            const double *ae, const double *an, const double *ek, const double *dtdt,
            const double *dlnpsdt,
            double *tvor, double *tdiv, double *ttemp, double *tlnps);

/* Mode-B timing harness: an OpenMP loop over independent (field, level) units. */
double orc_bench_set(const orc_cfg *c, int nlev, int nrep, int nthreads,
                     const double *spec_in, double *grid_work, double *spec_out);

/* The same scalar transforms arranged for CPU speed (cpu_port.c): hemispheric folding,
 * parity-split tables, a half-length Stockham real FFT. Timed as the CPU baseline and
 * checked against orc_synth / orc_anal. Needs an even nlat and nlon/2 of factors 2, 3, 5
 * (NULL otherwise). work: orc_port_work_doubles() doubles per concurrent call. */
typedef struct orc_port orc_port;
orc_port *orc_port_create(const orc_cfg *c);
void   orc_port_destroy(orc_port *p);
size_t orc_port_work_doubles(const orc_port *p);
void   orc_port_synth(const orc_port *p, const double *qlm, double *grid, double *work);
void   orc_port_anal(const orc_port *p, const double *grid, double *qlm, double *work);
int    orc_port_cfft(int n, int sign, const double *x, double *y);
double orc_port_bench_set(const orc_port *p, int nunit, int nrep, int nthreads,
                          const double *spec_in, double *grid_work, double *spec_out);

#ifdef __cplusplus
}
#endif
#endif
