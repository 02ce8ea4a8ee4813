/*
 * column_oracle.c -- FP64 CPU restatement of the grid-space column work of an
 * aeros step. TEST INFRASTRUCTURE ONLY (see column_oracle.h).
 *
 * What pins each function (tests/test_oracle_column.py):
 *   orc_vgrid_init       endpoints A(N) = 0, B(N) = 1, B(0) = 0, A(0) = p_top; pure
 *                        sigma exactly when sigma_t = 0; uniform sigma for
 *                        stretch_a = 1 (src/dynamics/aeros_vertical.f90:126-298)
 *   orc_vgrid_pressure   sum(dp) = p_s - p_top; p_full the mean of its half levels
 *   orc_vgrid_alpha      closed form for a sigma layer; alpha_1 = ln 2 at a
 *                        zero-pressure top (aeros_vertical.f90:382-426)
 *   orc_hydrostatic      isothermal closed form Phi_k = Phi_s + R T ln(p_s/p_k*)
 *                        with the Simmons-Burridge layer point
 *   orc_hs_*             the reference's point values (tests/test_held_suarez.f90:88-172)
 *   orc_column_terms     rest-state balance of the full tendency and the global
 *                        mass identity (tests/test_tendency.f90:77-107, 376-396),
 *                        pure-sigma c_k = 1, and the written-out terms of one
 *                        column computed independently in the test
 */
#include "column_oracle.h"

#include <math.h>
#include <stdlib.h>

/* Held & Suarez (1994) constants (src/physics/aeros_held_suarez.f90:92-101). */
#define HS_T_EQUATOR 315.0
#define HS_T_MIN 200.0
#define HS_DT_Y 60.0
#define HS_DTHETA_Z 10.0
#define HS_SIGMA_B 0.7
#define HS_P0 1.0e5
#define HS_K_A (1.0 / (40.0 * 86400.0))
#define HS_K_S (1.0 / (4.0 * 86400.0))
#define HS_K_F (1.0 / 86400.0)

/* ------------------------------------------------------------ vertical grid */

/* src/dynamics/aeros_vertical.f90:230-290: level placement
 * sigma = 1 - (a (1-x) + (1-a)(1-x)^r), x = j/N, squeezed onto [p_top/ps_ref, 1];
 * B ramps linearly from 0 at sigma_t to 1 at the surface; A takes the rest so
 * that p = sigma * ps_ref at p_s = ps_ref; endpoints pinned. */
int orc_vgrid_init(int nlev, double a, double r, double st, double p_top, double ps_ref,
                   double *A, double *B)
{
    if (nlev < 2 || a <= 0.0 || a > 1.0 || st < 0.0 || st >= 1.0) return -1;
    const double sig_top = p_top / ps_ref;
    for (int j = 0; j <= nlev; ++j) {
        const double x = (double)j / (double)nlev;
        const double s = 1.0 - (a * (1.0 - x) + (1.0 - a) * pow(1.0 - x, r));
        const double sh = sig_top + (1.0 - sig_top) * s;
        double b;
        if (st <= 0.0) b = sh;
        else if (sh <= st) b = 0.0;
        else b = (sh - st) / (1.0 - st);
        B[j] = b;
        A[j] = (sh - b) * ps_ref;
    }
    B[nlev] = 1.0;
    A[nlev] = 0.0;
    B[0] = 0.0;
    A[0] = p_top;
    return 0;
}

void orc_vgrid_pressure(int nlev, const double *A, const double *B, double ps,
                        double *p_half, double *p_full, double *dp)
{
    for (int j = 0; j <= nlev; ++j) p_half[j] = A[j] + B[j] * ps;
    for (int k = 1; k <= nlev; ++k) {
        if (p_full) p_full[k - 1] = 0.5 * (p_half[k] + p_half[k - 1]);
        if (dp) dp[k - 1] = p_half[k] - p_half[k - 1];
    }
}

/* alpha_k = 1 - (p_{k-1/2}/dp_k) ln(p_{k+1/2}/p_{k-1/2}); ln 2 and MV for a layer
 * whose top is at zero pressure. */
void orc_vgrid_alpha(int nlev, const double *p_half, double *alpha, double *dlnp)
{
    for (int k = 1; k <= nlev; ++k) {
        const double dpk = p_half[k] - p_half[k - 1];
        if (p_half[k - 1] <= 0.0) {
            alpha[k - 1] = log(2.0);
            if (dlnp) dlnp[k - 1] = ORC_MV;
        } else {
            const double lr = log(p_half[k] / p_half[k - 1]);
            alpha[k - 1] = 1.0 - (p_half[k - 1] / dpk) * lr;
            if (dlnp) dlnp[k - 1] = lr;
        }
    }
}

/* Upward from the surface: Phi_k = Phi_{k+1/2} + alpha_k R T_k,
 * Phi_{k-1/2} = Phi_{k+1/2} + R T_k ln(p_{k+1/2}/p_{k-1/2}). */
void orc_hydrostatic(int nlev, double phis, const double *temp, const double *p_half,
                     double *phi_full)
{
    double *alpha = malloc(sizeof(double) * nlev);
    orc_vgrid_alpha(nlev, p_half, alpha, NULL);
    double ph = phis;
    for (int k = nlev; k >= 1; --k) {
        phi_full[k - 1] = ph + alpha[k - 1] * ORC_R_D * temp[k - 1];
        if (p_half[k - 1] > 0.0) ph = ph + ORC_R_D * temp[k - 1] * log(p_half[k] / p_half[k - 1]);
        else ph = ORC_MV;
    }
    free(alpha);
}

/* --------------------------------------------------------------- Held-Suarez */

double orc_hs_teq(double p, double sinlat2, double coslat2)
{
    const double pf = p / HS_P0;
    const double t = (HS_T_EQUATOR - HS_DT_Y * sinlat2 - HS_DTHETA_Z * log(pf) * coslat2) * pow(pf, ORC_KAPPA);
    return t > HS_T_MIN ? t : HS_T_MIN;
}

static double hs_hfac(double sigma)
{
    const double f = (sigma - HS_SIGMA_B) / (1.0 - HS_SIGMA_B);
    return f > 0.0 ? f : 0.0;
}

double orc_hs_kt(double sigma, double coslat4) { return HS_K_A + (HS_K_S - HS_K_A) * hs_hfac(sigma) * coslat4; }
double orc_hs_kv(double sigma) { return HS_K_F * hs_hfac(sigma); }

/* ------------------------------------------------------------- column terms */

/* Simmons-Burridge vertical advection of level k (0-based) with mass fluxes
 * mf[0..nlev] at half levels (src/dynamics/aeros_tendency.f90:599-628). */
static double vert_adv(const double *x, const double *mf, const double *dp, int k, int nlev)
{
    double lower = 0.0, upper = 0.0;
    if (k < nlev - 1) lower = mf[k + 1] * (x[k + 1] - x[k]);
    if (k > 0) upper = mf[k] * (x[k] - x[k - 1]);
    return 0.5 * (lower + upper) / dp[k];
}

void orc_column_terms(int nlon, int nlat, int nlev, const double *A, const double *B,
                      const double *sinlat,
                      const double *u, const double *v, const double *vor_g,
                      const double *div_g, const double *t_g, const double *dtdx,
                      const double *dtdy, const double *lnps_g, const double *dlnpsdx,
                      const double *dlnpsdy, const double *phis,
                      double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt,
                      int hs)
{
    const long ng = (long)nlon * nlat;
    const int ztop = A[0] <= 0.0 && B[0] <= 0.0;
    double *ph = malloc(sizeof(double) * (nlev + 1)), *pf = malloc(sizeof(double) * nlev);
    double *dpl = malloc(sizeof(double) * nlev), *al = malloc(sizeof(double) * nlev);
    double *dl = malloc(sizeof(double) * nlev), *cf = malloc(sizeof(double) * nlev);
    double *uc = malloc(sizeof(double) * nlev), *vc = malloc(sizeof(double) * nlev);
    double *tc = malloc(sizeof(double) * nlev), *gc = malloc(sizeof(double) * nlev);
    double *cm = malloc(sizeof(double) * nlev), *sc = malloc(sizeof(double) * (nlev + 1));
    double *mf = malloc(sizeof(double) * (nlev + 1)), *om = malloc(sizeof(double) * nlev);
    double *phi = malloc(sizeof(double) * nlev);
    for (int j = 0; j < nlat; ++j) {
        const double f = 2.0 * ORC_OMEGA * sinlat[j];
        const double s2 = sinlat[j] * sinlat[j], c2 = 1.0 - s2, c4 = c2 * c2;
        for (int i = 0; i < nlon; ++i) {
            const long c = (long)j * nlon + i;
            const double ps = exp(lnps_g[c]);
            orc_vgrid_pressure(nlev, A, B, ps, ph, pf, dpl);
            orc_vgrid_alpha(nlev, ph, al, dl);
            if (ztop) dl[0] = 0.0;
            for (int k = 0; k < nlev; ++k)
                cf[k] = (ps / dpl[k]) * (B[k] * dl[k] + (B[k + 1] - B[k]) * al[k]);
            for (int k = 0; k < nlev; ++k) {
                uc[k] = u[k * ng + c];
                vc[k] = v[k * ng + c];
                tc[k] = t_g[k * ng + c];
                gc[k] = uc[k] * dlnpsdx[c] + vc[k] * dlnpsdy[c];
            }
            sc[0] = 0.0;
            for (int k = 0; k < nlev; ++k) {
                cm[k] = dpl[k] * div_g[k * ng + c] + (B[k + 1] - B[k]) * ps * gc[k];
                sc[k + 1] = sc[k] + cm[k];
            }
            dlnpsdt[c] = -sc[nlev] / ps;
            for (int k = 0; k <= nlev; ++k) mf[k] = B[k] * sc[nlev] - sc[k];
            for (int k = 0; k < nlev; ++k)
                om[k] = cf[k] * gc[k] - (sc[k] * dl[k] + al[k] * cm[k]) / dpl[k];
            orc_hydrostatic(nlev, phis[c], tc, ph, phi);
            for (int k = 0; k < nlev; ++k) {
                const long x = k * ng + c;
                const double zf = vor_g[x] + f;
                const double au = vert_adv(uc, mf, dpl, k, nlev);
                const double av = vert_adv(vc, mf, dpl, k, nlev);
                const double at = vert_adv(tc, mf, dpl, k, nlev);
                const double pgf = ORC_R_D * tc[k] * cf[k];
                ae[x] = zf * vc[k] - au - pgf * dlnpsdx[c];
                an[x] = -zf * uc[k] - av - pgf * dlnpsdy[c];
                const double kin = 0.5 * (uc[k] * uc[k] + vc[k] * vc[k]);
                ek[x] = phi[k] + kin;
                dtdt[x] = -(uc[k] * dtdx[x] + vc[k] * dtdy[x]) - at + ORC_KAPPA * tc[k] * om[k];
            }
            if (hs) {
                for (int k = 0; k < nlev; ++k) {
                    const long x = k * ng + c;
                    const double sig = pf[k] / ps;
                    const double teq = orc_hs_teq(pf[k], s2, c2);
                    const double kt = orc_hs_kt(sig, c4), kv = orc_hs_kv(sig);
                    dtdt[x] = dtdt[x] - kt * (t_g[x] - teq);
                    ae[x] = ae[x] - kv * u[x];
                    an[x] = an[x] - kv * v[x];
                }
            }
        }
    }
    free(ph); free(pf); free(dpl); free(al); free(dl); free(cf); free(uc); free(vc);
    free(tc); free(gc); free(cm); free(sc); free(mf); free(om); free(phi);
}
