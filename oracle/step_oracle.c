/*
 * step_oracle.c -- FP64 CPU restatement of the spectral step completion.
 * TEST INFRASTRUCTURE ONLY (see step_oracle.h).
 *
 * What pins each function (tests/test_oracle_step.py):
 *   orc_si_create/_gwspeed  the external mode within 5% of sqrt(gamma R T) and
 *                           coordinate-independent to 1% (tests/test_semiimp.f90:72-111)
 *   orc_si_linear           L against a finite-differenced nonlinear tendency of
 *                           the oracle's own tendency chain (tests/test_semiimp.f90:167-289)
 *   orc_si_step             the scheme's defining equation, re-evaluated with
 *                           orc_si_linear, at alpha = 1/2 and 1 (:291-409)
 *   orc_diff_factors        closed form at l = 0 (1) and l = lmax (1/(1+h/tau))
 *   orc_diffuse             per-degree factors applied to every m of the degree
 *   orc_raw_filter          the filter conserves old + cur + new for raw_alpha = 1/2
 *                           shifted mean; d = nu/2 (X- - 2X + X+) on a closed case
 */
#include "step_oracle.h"
#include "column_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_EARTH 6.371e6   /* src/aeros_defs.f90:84 */

/* Doolittle LU with partial pivoting, row-major, in place
 * (src/dynamics/aeros_semiimp.f90:658-713). Returns -1 if singular. */
static int lu_factor(double *a, int *piv, int n)
{
    for (int k = 0; k < n; ++k) {
        int ip = k;
        double amax = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[i * n + k]) > amax) {
                amax = fabs(a[i * n + k]);
                ip = i;
            }
        if (amax <= 0.0) return -1;
        piv[k] = ip;
        if (ip != k)
            for (int j = 0; j < n; ++j) {
                const double t = a[k * n + j];
                a[k * n + j] = a[ip * n + j];
                a[ip * n + j] = t;
            }
        for (int i = k + 1; i < n; ++i) {
            a[i * n + k] = a[i * n + k] / a[k * n + k];
            for (int j = k + 1; j < n; ++j) a[i * n + j] = a[i * n + j] - a[i * n + k] * a[k * n + j];
        }
    }
    return 0;
}

/* L U x = P b for one real column b, in place (:715-750). */
static void lu_solve(const double *a, const int *piv, double *b, int n)
{
    for (int k = 0; k < n; ++k) {
        if (piv[k] != k) {
            const double t = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = t;
        }
        for (int j = k + 1; j < n; ++j) b[j] = b[j] - a[j * n + k] * b[k];
    }
    for (int k = n - 1; k >= 0; --k) {
        for (int j = k + 1; j < n; ++j) b[k] = b[k] - a[k * n + j] * b[j];
        b[k] = b[k] / a[k * n + k];
    }
}

orc_si *orc_si_create(int nlev, const double *A, const double *B, double t_ref, double ps_ref,
                      int lmax, double dt_step, double alpha)
{
    if (nlev < 1 || lmax < 0 || alpha < 0.5 || alpha > 1.0) return NULL;
    orc_si *s = calloc(1, sizeof(orc_si));
    const int n = nlev;
    s->nlev = n;
    s->lmax = lmax;
    s->alpha = alpha;
    s->dp_ref = malloc(sizeof(double) * n);
    s->alpha_ref = malloc(sizeof(double) * n);
    s->dlnp_ref = malloc(sizeof(double) * n);
    s->gmat = calloc((size_t)n * n, sizeof(double));
    s->tmat = calloc((size_t)n * n, sizeof(double));
    s->wmat = calloc((size_t)n * n, sizeof(double));
    s->nu = malloc(sizeof(double) * n);
    s->pcoef = malloc(sizeof(double) * n);
    s->lu = malloc(sizeof(double) * (size_t)(lmax + 1) * n * n);
    s->piv = malloc(sizeof(int) * (size_t)(lmax + 1) * n);
    double *ph = malloc(sizeof(double) * (n + 1));
    orc_vgrid_pressure(n, A, B, ps_ref, ph, NULL, s->dp_ref);
    orc_vgrid_alpha(n, ph, s->alpha_ref, s->dlnp_ref);
    free(ph);
    if (A[0] <= 0.0 && B[0] <= 0.0) s->dlnp_ref[0] = 0.0;
    /* G: Phi_k = Phi_s + alpha_k R T_k + sum_{j>k} R T_j dlnp_j */
    for (int k = 0; k < n; ++k) {
        s->gmat[k * n + k] = ORC_R_D * s->alpha_ref[k];
        for (int j = k + 1; j < n; ++j) s->gmat[k * n + j] = ORC_R_D * s->dlnp_ref[j];
    }
    /* tau: kappa T_ref (omega/p) with the winds and grad ln p_s set to zero */
    for (int k = 0; k < n; ++k) {
        s->tmat[k * n + k] = ORC_KAPPA * t_ref * s->alpha_ref[k];
        for (int j = 0; j < k; ++j)
            s->tmat[k * n + j] = ORC_KAPPA * t_ref * s->dlnp_ref[k] * s->dp_ref[j] / s->dp_ref[k];
    }
    for (int k = 0; k < n; ++k) {
        s->nu[k] = s->dp_ref[k] / ps_ref;
        s->pcoef[k] = ORC_R_D * t_ref;
    }
    /* W = G tau + R T_ref (x) nu, summed in the reference's order */
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            double w = s->pcoef[k] * s->nu[j];
            for (int m = 0; m < n; ++m) w = w + s->gmat[k * n + m] * s->tmat[m * n + j];
            s->wmat[k * n + j] = w;
        }
    if (orc_si_set_step(s, dt_step) != 0) {
        orc_si_destroy(s);
        return NULL;
    }
    return s;
}

void orc_si_destroy(orc_si *s)
{
    if (!s) return;
    free(s->dp_ref); free(s->alpha_ref); free(s->dlnp_ref);
    free(s->gmat); free(s->tmat); free(s->wmat); free(s->nu); free(s->pcoef);
    free(s->lu); free(s->piv);
    free(s);
}

int orc_si_set_step(orc_si *s, double dt_step)
{
    if (dt_step <= 0.0) return -1;
    const int n = s->nlev;
    s->dt_step = dt_step;
    const double hh = s->alpha * s->alpha * dt_step * dt_step;
    const double a2 = A_EARTH * A_EARTH;
    for (int l = 0; l <= s->lmax; ++l) {
        const double lam = (double)(l * (l + 1)) / a2;
        double *a = s->lu + (size_t)l * n * n;
        for (int j = 0; j < n; ++j) {
            for (int k = 0; k < n; ++k) a[k * n + j] = hh * lam * s->wmat[k * n + j];
            a[j * n + j] = a[j * n + j] + 1.0;
        }
        if (lu_factor(a, s->piv + (size_t)l * n, n) != 0) return -1;
    }
    return 0;
}

/* L_D alone for one real column: lam (pcoef p + G t) (:436-460). */
static void lin_div(const orc_si *s, double lam, const double *t, double p, double *ld)
{
    const int n = s->nlev;
    for (int k = 0; k < n; ++k) {
        double acc = s->pcoef[k] * p;
        for (int j = k; j < n; ++j) acc = acc + s->gmat[k * n + j] * t[j];
        ld[k] = lam * acc;
    }
}

/* L for one real column (:401-434). */
static void lin_column(const orc_si *s, double lam, const double *d, const double *t, double p,
                       double *ld, double *lt, double *lp)
{
    const int n = s->nlev;
    lin_div(s, lam, t, p, ld);
    for (int k = 0; k < n; ++k) {
        double acc = 0.0;
        for (int j = 0; j <= k; ++j) acc = acc + s->tmat[k * n + j] * d[j];
        lt[k] = -acc;
    }
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc = acc + s->nu[k] * d[k];
    *lp = -acc;
}

void orc_si_linear(const orc_si *s, int nlm, const int *l_of_lm,
                   const double *div, const double *temp, const double *lnps,
                   double *ldiv, double *ltemp, double *lps)
{
    const int n = s->nlev;
    const long nc = 2L * nlm;
    double *d = malloc(sizeof(double) * n * 5), *t = d + n, *ld = t + n, *lt = ld + n;
    for (long q = 0; q < nc; ++q) {
        const int l = l_of_lm[q / 2];
        const double lam = (double)(l * (l + 1)) / (A_EARTH * A_EARTH);
        for (int k = 0; k < n; ++k) {
            d[k] = div[k * nc + q];
            t[k] = temp[k * nc + q];
        }
        double lp;
        lin_column(s, lam, d, t, lnps[q], ld, lt, &lp);
        for (int k = 0; k < n; ++k) {
            ldiv[k * nc + q] = ld[k];
            ltemp[k * nc + q] = lt[k];
        }
        lps[q] = lp;
    }
    free(d);
}

void orc_si_step(const orc_si *s, int nlm, const int *l_of_lm,
                 const double *old_div, const double *old_temp, const double *old_lnps,
                 const double *now_div, const double *now_temp, const double *now_lnps,
                 const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                 double *new_div, double *new_temp, double *new_lnps)
{
    const int n = s->nlev;
    const long nc = 2L * nlm;
    const double h = s->dt_step, a1 = s->alpha;
    const double astep = a1 * h, hh_old = a1 * (1.0 - a1) * h * h, a2 = A_EARTH * A_EARTH;
    double *buf = malloc(sizeof(double) * n * 12);
    double *dold = buf, *dnow = buf + n, *told = buf + 2 * n, *tnow = buf + 3 * n;
    double *rd = buf + 4 * n, *rt = buf + 5 * n, *ltnow = buf + 6 * n, *ldnow = buf + 7 * n;
    double *ldstar = buf + 8 * n, *tstar = buf + 9 * n, *rhs = buf + 10 * n, *dbar = buf + 11 * n;
    for (long q = 0; q < nc; ++q) {
        const int l = l_of_lm[q / 2];
        const double lam = (double)(l * (l + 1)) / a2;
        for (int k = 0; k < n; ++k) {
            dold[k] = old_div[k * nc + q];
            dnow[k] = now_div[k * nc + q];
            told[k] = old_temp[k * nc + q];
            tnow[k] = now_temp[k * nc + q];
            rd[k] = tnd_div[k * nc + q];
            rt[k] = tnd_temp[k * nc + q];
        }
        const double pold = old_lnps[q], pnow = now_lnps[q], rp = tnd_lnps[q];
        double lpnow;
        lin_column(s, lam, dnow, tnow, pnow, ldnow, ltnow, &lpnow);
        for (int k = 0; k < n; ++k) tstar[k] = told[k] + astep * (rt[k] - ltnow[k]);
        const double pstar = pold + astep * (rp - lpnow);
        lin_div(s, lam, tstar, pstar, ldstar);
        for (int k = 0; k < n; ++k) {
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc = acc + s->wmat[k * n + j] * dold[j];
            rhs[k] = dold[k] - hh_old * lam * acc + h * (rd[k] + ldstar[k] - ldnow[k]);
        }
        lu_solve(s->lu + (size_t)l * n * n, s->piv + (size_t)l * n, rhs, n);
        for (int k = 0; k < n; ++k) {
            new_div[k * nc + q] = rhs[k];
            dbar[k] = a1 * rhs[k] + (1.0 - a1) * dold[k];
        }
        for (int k = 0; k < n; ++k) {
            double acc = 0.0;
            for (int j = 0; j <= k; ++j) acc = acc + s->tmat[k * n + j] * dbar[j];
            new_temp[k * nc + q] = told[k] + h * (rt[k] - acc - ltnow[k]);
        }
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc = acc + s->nu[k] * dbar[k];
        new_lnps[q] = pold + h * (rp + (-acc) - lpnow);
    }
    free(buf);
}

double orc_si_gwspeed(const orc_si *s)
{
    const int n = s->nlev;
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) x[k] = 1.0 / sqrt((double)n);
    double lam = 0.0;
    for (int it = 0; it < 200; ++it) {
        for (int k = 0; k < n; ++k) {
            y[k] = 0.0;
            for (int j = 0; j < n; ++j) y[k] = y[k] + s->wmat[k * n + j] * x[j];
        }
        double nrm = 0.0;
        for (int k = 0; k < n; ++k) nrm += y[k] * y[k];
        nrm = sqrt(nrm);
        if (nrm <= 0.0) break;
        for (int k = 0; k < n; ++k) x[k] = y[k] / nrm;
        lam = 0.0;
        for (int k = 0; k < n; ++k) lam += x[k] * y[k];
    }
    free(x);
    free(y);
    return sqrt(lam > 0.0 ? lam : 0.0);
}

void orc_diff_factors(int lmax, double h, double tau, int ndiff, double *fac)
{
    const double lref = (double)(lmax * (lmax + 1)), rate = h / tau;
    for (int l = 0; l <= lmax; ++l) {
        /* integer power as repeated multiplication (the reference's x**(ndiff/2)) */
        const double x = (double)(l * (l + 1)) / lref;
        double r = ndiff / 2 > 0 ? x : 1.0;
        for (int i = 1; i < ndiff / 2; ++i) r = r * x;
        fac[l] = 1.0 / (1.0 + 1.0 * rate * r);
    }
}

void orc_diffuse(int nlm, int nlev, int lmax, const int *l_of_lm, const double *fac_vt,
                 const double *fac_div, double *vor, double *div, double *temp)
{
    const long nc = 2L * nlm;
    for (int k = 0; k < nlev; ++k)
        for (long q = 0; q < nc; ++q) {
            const int l = l_of_lm[q / 2];
            const double fv = fac_vt[k * (lmax + 1) + l], fd = fac_div[k * (lmax + 1) + l];
            vor[k * nc + q] = fv * vor[k * nc + q];
            div[k * nc + q] = fd * div[k * nc + q];
            temp[k * nc + q] = fv * temp[k * nc + q];
        }
}

void orc_raw_filter(long n, double eps, double raw_alpha, const double *old, double *cur, double *nw)
{
    if (eps <= 0.0) return;
    const double half_nu = 0.5 * eps, alpha = raw_alpha, beta = 1.0 - raw_alpha;
    for (long i = 0; i < n; ++i) {
        const double d = half_nu * (old[i] - 2.0 * cur[i] + nw[i]);
        cur[i] = cur[i] + alpha * d;
        nw[i] = nw[i] - beta * d;
    }
}
