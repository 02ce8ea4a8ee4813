/*
 * sht_oracle.c -- FP64 CPU restatement of the aeros transform path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library in
 * aeros_amd/: tests/ compare against it, the smoke check calls it, and the
 * CPU-baseline leg of bench.py times it. The product never links it.
 *
 * It restates, from the mathematics, what SHTns computes for the calls aeros
 * makes (src/aeros_spectral.f90:486-584; SHTns itself is not vendored in the
 * reference) plus aeros' own operators around them. A plain dense
 * Gauss-Legendre quadrature: no polar truncation (SHTns' eps = 1e-10 skip,
 * src/aeros_spectral.f90:193-199, drops terms below 1e-10 relative that are
 * ~1e-60 at T42), no hemispheric folding, the Fourier step is an own
 * mixed-radix complex FFT (or an exact O(N^2) DFT when fft_direct is set).
 *
 * What pins each function (tests/test_oracle.py, tests/test_oracle_column.py):
 *   orc_grid_size, orc_lm     the reference's own table (tests/test_grid.f90:25-37,
 *                             tests/test_spectral.f90:45-68) and a brute-force
 *                             enumeration of (l, m)
 *   orc_gauss                 numpy.polynomial.legendre.leggauss; sum(w) = 2;
 *                             exact mirror symmetry; mpmath 50-digit nodes
 *   orc_legendre (P, D)       mpmath 50-digit golden tables (tests/golden);
 *                             scipy.special.sph_harm_y with the CS phase removed;
 *                             D against a central difference of P in theta
 *   orc_rfft_*                numpy.fft.rfft / irfft; FFT vs direct DFT
 *   orc_synth / orc_anal      mpmath 50-digit direct-sum golden grids (T5, T21);
 *                             Y_00 = 1/sqrt(4 pi); round trip; Parseval
 *   orc_synth_vec / _grad     closed forms T_10 -> V_phi = +N sin(theta),
 *                             S_10 -> V_theta = -N sin(theta)
 *                             (src/dynamics/aeros_vordiv.f90:36-42); mpmath golden
 *   orc_anal_vec              round trip; solid body (tests/test_vordiv.f90:59-143)
 *   orc_laplacian             exact eigenvalue (tests/test_spectral.f90:110-124)
 *   orc_surface_integral      4 pi a^2 and int Y_10 = 0 (tests/test_spectral.f90:89-104)
 *   vordiv / gradient         solid body, pure divergence, l = 0 discarded
 *                             (tests/test_vordiv.f90:59-289)
 *   tendency transforms       the rest-state balance and the global-mass identity
 *                             through them, and a planted-mistake check
 *                             (test_oracle_column.py)
 *   orc_port_* (cpu_port.c)   the CPU-baseline arrangement: its FFT against numpy.fft,
 *                             its synthesis / analysis against orc_synth / orc_anal
 *                             at T5-T106 (tests/test_oracle.py)
 */
#define _POSIX_C_SOURCE 199309L
#include "sht_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338327950288

/* ---------------------------------------------------------------- grid size */

/* True if n is even and 2,3,5-smooth (src/aeros_spectral.f90:440-458). */
static int is_smooth_even(int n)
{
    if (n <= 0 || n % 2 != 0) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

/* Smallest quadratically unaliased grid: nlon >= 3T+1, smooth, even; nlat = nlon/2
 * (src/aeros_spectral.f90:399-438). */
int orc_grid_size(int trunc, int *nlon, int *nlat)
{
    if (trunc < 0) return 1;
    int n = 3 * trunc + 1;
    while (!is_smooth_even(n)) n++;
    *nlon = n;
    *nlat = n / 2;
    return 0;
}

/* 1-based index of (l, m), -1 outside the truncation (src/aeros_spectral.f90:460-482). */
int orc_lm(int lmax, int l, int m)
{
    if (m < 0 || m > lmax || l < m || l > lmax) return -1;
    return (m * (2 * lmax + 2 - (m + 1))) / 2 + l + 1;
}

/* ---------------------------------------------------------- Gauss-Legendre */

/* Gauss-Legendre nodes mu (descending: north -> south), weights and sin(colatitude),
 * by Newton's method on P_n(mu) with the three-term recurrence. 1 - mu^2 is formed
 * as (1-mu)(1+mu), which is exact to rounding for the represented node (1 - mu is
 * exact by Sterbenz's lemma), so neither the weights 2/((1-mu^2) P_n'(mu)^2) nor
 * sin(theta) lose accuracy near the poles. The southern half is the exact mirror
 * of the northern half, as the reference mirrors SHTns' half-hemisphere weights
 * (src/aeros_spectral.f90:246-275). */
static void gauss_theta(int n, double *mu, double *w, double *sint)
{
    int nh = (n + 1) / 2;
    for (int i = 0; i < nh; i++) {
        double x = cos(ORC_PI * (i + 0.75) / (n + 0.5));
        double p0 = 1.0, p1 = x, omx2 = 1.0, dp = 1.0;
        for (int it = 0; it <= 100; it++) {
            p0 = 1.0;
            p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            omx2 = (1.0 - x) * (1.0 + x);
            dp = n * (p0 - x * p1) / omx2;   /* P_n'(x) */
            if (it == 100) break;
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) <= 1e-17) {
                it = 99;   /* one more evaluation at the converged node */
            }
        }
        double wi = 2.0 / (omx2 * dp * dp);
        mu[i] = x;
        mu[n - 1 - i] = -x;
        w[i] = w[n - 1 - i] = wi;
        if (sint) sint[i] = sint[n - 1 - i] = sqrt(omx2);
    }
    if (n % 2 == 1) {
        mu[nh - 1] = 0.0;
        if (sint) sint[nh - 1] = 1.0;
    }
}

void orc_gauss(int n, double *mu, double *w) { gauss_theta(n, mu, w, NULL); }

/* ------------------------------------------------------ Legendre functions */

/* Orthonormal associated Legendre functions without the Condon-Shortley phase,
 * P_lm with 2 pi int_{-1}^{1} P_lm^2 dmu = 1, and their colatitude derivative
 * dP_lm/dtheta. Standard three-term recurrence in l at fixed m, seeded by
 * P_mm = sqrt((2m+1)!!/(4 pi (2m)!!)) sin^m(theta) (positive for every m). */
static void legendre_core(int lmax, int nlat, const double *mu, const double *sint, double *P, double *D)
{
    int lm = 0;
    for (int m = 0; m <= lmax; m++) {
        for (int l = m; l <= lmax; l++, lm++) {
            for (int j = 0; j < nlat; j++) {
                double x = mu[j];
                double s = sint[j];
                double p;
                if (l == m) {
                    p = 1.0 / sqrt(4.0 * ORC_PI);
                    for (int k = 1; k <= m; k++) p *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
                } else if (l == m + 1) {
                    p = sqrt(2.0 * m + 3.0) * x * P[(size_t)(lm - 1) * nlat + j];
                } else {
                    double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                    double b = sqrt((((double)l - 1) * (l - 1) - (double)m * m) /
                                    (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                    p = a * (x * P[(size_t)(lm - 1) * nlat + j] - b * P[(size_t)(lm - 2) * nlat + j]);
                }
                P[(size_t)lm * nlat + j] = p;
                /* sin(theta) dP_lm/dtheta = l mu P_lm - sqrt((2l+1)/(2l-1) (l^2-m^2)) P_{l-1,m} */
                double prev = (l > m) ? P[(size_t)(lm - 1) * nlat + j] : 0.0;
                double c = (l > m) ? sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) *
                                          ((double)l * l - (double)m * m)) : 0.0;
                D[(size_t)lm * nlat + j] = (l * x * p - c * prev) / s;
            }
        }
    }
}

void orc_legendre(int lmax, int nlat, const double *mu, double *P, double *D)
{
    double *s = (double *)malloc(sizeof(double) * nlat);
    for (int j = 0; j < nlat; j++) s[j] = sqrt((1.0 - mu[j]) * (1.0 + mu[j]));
    legendre_core(lmax, nlat, mu, s, P, D);
    free(s);
}

/* ------------------------------------------------------------------ config */

static double *make_twiddles(int n, int sign);

orc_cfg *orc_create(int trunc, int nlon, int nlat)
{
    int dlon, dlat;
    if (orc_grid_size(trunc, &dlon, &dlat)) return NULL;
    if (nlon <= 0) nlon = dlon;
    if (nlat <= 0) nlat = dlat;
    orc_cfg *c = (orc_cfg *)calloc(1, sizeof(orc_cfg));
    c->trunc = c->lmax = c->mmax = trunc;
    c->nlon = nlon;
    c->nlat = nlat;
    c->nlm = (trunc + 1) * (trunc + 2) / 2;
    c->mu = (double *)malloc(sizeof(double) * nlat);
    c->w = (double *)malloc(sizeof(double) * nlat);
    c->sint = (double *)malloc(sizeof(double) * nlat);
    c->P = (double *)malloc(sizeof(double) * (size_t)c->nlm * nlat);
    c->D = (double *)malloc(sizeof(double) * (size_t)c->nlm * nlat);
    c->l_of_lm = (int *)malloc(sizeof(int) * c->nlm);
    c->m_of_lm = (int *)malloc(sizeof(int) * c->nlm);
    gauss_theta(nlat, c->mu, c->w, c->sint);
    legendre_core(trunc, nlat, c->mu, c->sint, c->P, c->D);
    c->tw_fwd = make_twiddles(nlon, -1);
    c->tw_inv = make_twiddles(nlon, +1);
    int lm = 0;
    for (int m = 0; m <= trunc; m++)
        for (int l = m; l <= trunc; l++, lm++) {
            c->l_of_lm[lm] = l;
            c->m_of_lm[lm] = m;
        }
    return c;
}

void orc_destroy(orc_cfg *c)
{
    if (!c) return;
    free(c->mu); free(c->w); free(c->sint); free(c->P); free(c->D);
    free(c->l_of_lm); free(c->m_of_lm);
    free(c->tw_fwd); free(c->tw_inv);
    free(c);
}

void orc_set_fft_direct(orc_cfg *c, int direct) { c->fft_direct = direct; }

/* -------------------------------------------------------------------- FFT */

/* Mixed-radix decimation-in-time complex FFT, y[k] = sum_j x[j] e^{sign 2 pi i jk/n}.
 * x is read with stride xs (complex elements), y is contiguous. tw holds
 * e^{sign 2 pi i k/N} for the top-level N; tstep = N/n. */
static void fft_rec(int n, const double *x, int xs, double *y, int sign,
                    const double *tw, int tstep)
{
    if (n == 1) { y[0] = x[0]; y[1] = x[1]; return; }
    int p = (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
    int q = n / p;
    for (int r = 0; r < p; r++) fft_rec(q, x + 2 * (size_t)r * xs, xs * p, y + 2 * (size_t)r * q, sign, tw, tstep * p);
    double t[2 * 64];
    double *tt = (p <= 64) ? t : (double *)malloc(sizeof(double) * 2 * p);
    double obuf[2 * 1024];
    double *out = (n <= 1024) ? obuf : (double *)malloc(sizeof(double) * 2 * (size_t)n);
    int N = n * tstep;
    for (int k = 0; k < q; k++) {
        for (int r = 0; r < p; r++) {
            int idx = (int)(((long)r * k * tstep) % N);
            double wr = tw[2 * idx], wi = tw[2 * idx + 1];
            double ar = y[2 * (r * q + k)], ai = y[2 * (r * q + k) + 1];
            tt[2 * r] = ar * wr - ai * wi;
            tt[2 * r + 1] = ar * wi + ai * wr;
        }
        for (int s = 0; s < p; s++) {
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < p; r++) {
                int idx = (int)(((long)r * s * q * tstep) % N);
                double wr = tw[2 * idx], wi = tw[2 * idx + 1];
                sr += tt[2 * r] * wr - tt[2 * r + 1] * wi;
                si += tt[2 * r] * wi + tt[2 * r + 1] * wr;
            }
            out[2 * (k + s * q)] = sr;
            out[2 * (k + s * q) + 1] = si;
        }
    }
    memcpy(y, out, sizeof(double) * 2 * (size_t)n);
    if (out != obuf) free(out);
    if (tt != t) free(tt);
}

static double *make_twiddles(int n, int sign)
{
    double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int k = 0; k < n; k++) {
        double a = 2.0 * ORC_PI * (double)k / (double)n;
        tw[2 * k] = cos(a);
        tw[2 * k + 1] = sign * sin(a);
    }
    return tw;
}

static void cfft(int n, const double *x, double *y, int sign, const double *tw)
{
    double *own = tw ? NULL : make_twiddles(n, sign);
    fft_rec(n, x, 1, y, sign, tw ? tw : own, 1);
    free(own);
}

/* X_m = sum_i x_i e^{-i m phi_i}, m = 0..mmax (unscaled). */
static void rdft_forward(int n, int mmax, const double *x, double *X, int direct, const double *tw)
{
    if (direct) {
        for (int m = 0; m <= mmax; m++) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < n; i++) {
                double a = 2.0 * ORC_PI * (double)(((long)m * i) % n) / (double)n;
                sr += x[i] * cos(a);
                si -= x[i] * sin(a);
            }
            X[2 * m] = sr;
            X[2 * m + 1] = si;
        }
        return;
    }
    double *z = (double *)calloc(2 * (size_t)n, sizeof(double));
    double *Z = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int i = 0; i < n; i++) z[2 * i] = x[i];
    cfft(n, z, Z, -1, tw);
    for (int m = 0; m <= mmax; m++) { X[2 * m] = Z[2 * m]; X[2 * m + 1] = Z[2 * m + 1]; }
    free(z);
    free(Z);
}

/* x_i = Re X_0 + 2 sum_{m=1}^{mmax} Re(X_m e^{i m phi_i}). Requires mmax < n/2. */
static void rdft_inverse(int n, int mmax, const double *X, double *x, int direct, const double *tw)
{
    if (direct) {
        for (int i = 0; i < n; i++) {
            double s = X[0];
            for (int m = 1; m <= mmax; m++) {
                double a = 2.0 * ORC_PI * (double)(((long)m * i) % n) / (double)n;
                s += 2.0 * (X[2 * m] * cos(a) - X[2 * m + 1] * sin(a));
            }
            x[i] = s;
        }
        return;
    }
    double *Z = (double *)calloc(2 * (size_t)n, sizeof(double));
    double *z = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    Z[0] = X[0];
    for (int m = 1; m <= mmax; m++) {
        Z[2 * m] = X[2 * m];
        Z[2 * m + 1] = X[2 * m + 1];
        Z[2 * (n - m)] = X[2 * m];
        Z[2 * (n - m) + 1] = -X[2 * m + 1];
    }
    cfft(n, Z, z, +1, tw);
    for (int i = 0; i < n; i++) x[i] = z[2 * i];
    free(Z);
    free(z);
}

void orc_rfft_forward(int n, int mmax, const double *x, double *X) { rdft_forward(n, mmax, x, X, 0, NULL); }
void orc_rfft_inverse(int n, int mmax, const double *X, double *x) { rdft_inverse(n, mmax, X, x, 0, NULL); }

/* ------------------------------------------------------- scalar transforms */

#define LMOF(c, l, m) ((m) * (2 * (c)->lmax + 1 - (m)) / 2 + (l))   /* 0-based */

/* f(phi_i, theta_j) = sum_l c_l0 P_l0 + 2 Re sum_{m>=1} sum_l c_lm P_lm e^{i m phi_i}
 * (SH_to_spat as called at src/aeros_spectral.f90:502-515). */
void orc_synth(const orc_cfg *c, const double *qlm, double *grid)
{
    int nlat = c->nlat, nlon = c->nlon, M = c->mmax;
    double *F = (double *)calloc((size_t)nlat * 2 * (M + 1), sizeof(double));
    for (int m = 0; m <= M; m++)
        for (int l = m; l <= c->lmax; l++) {
            int lm = LMOF(c, l, m);
            double cr = qlm[2 * lm], ci = qlm[2 * lm + 1];
            const double *p = c->P + (size_t)lm * nlat;
            for (int j = 0; j < nlat; j++) {
                F[(size_t)j * 2 * (M + 1) + 2 * m] += cr * p[j];
                F[(size_t)j * 2 * (M + 1) + 2 * m + 1] += ci * p[j];
            }
        }
    for (int j = 0; j < nlat; j++)
        rdft_inverse(nlon, M, F + (size_t)j * 2 * (M + 1), grid + (size_t)j * nlon, c->fft_direct, c->tw_inv);
    free(F);
}

/* c_lm = sum_j w_j P_lm(mu_j) (2 pi/nlon) sum_i f_ij e^{-i m phi_i}
 * (spat_to_SH as called at src/aeros_spectral.f90:486-500). */
void orc_anal(const orc_cfg *c, const double *grid, double *qlm)
{
    int nlat = c->nlat, nlon = c->nlon, M = c->mmax;
    double *G = (double *)malloc(sizeof(double) * (size_t)nlat * 2 * (M + 1));
    for (int j = 0; j < nlat; j++) {
        rdft_forward(nlon, M, grid + (size_t)j * nlon, G + (size_t)j * 2 * (M + 1), c->fft_direct, c->tw_fwd);
        double s = c->w[j] * 2.0 * ORC_PI / nlon;
        for (int k = 0; k < 2 * (M + 1); k++) G[(size_t)j * 2 * (M + 1) + k] *= s;
    }
    for (int m = 0; m <= M; m++)
        for (int l = m; l <= c->lmax; l++) {
            int lm = LMOF(c, l, m);
            const double *p = c->P + (size_t)lm * nlat;
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < nlat; j++) {
                sr += p[j] * G[(size_t)j * 2 * (M + 1) + 2 * m];
                si += p[j] * G[(size_t)j * 2 * (M + 1) + 2 * m + 1];
            }
            qlm[2 * lm] = sr;
            qlm[2 * lm + 1] = si;
        }
    free(G);
}

/* ------------------------------------------------------- vector transforms */

/* V_theta = dS/dtheta + (1/sin) dT/dphi, V_phi = (1/sin) dS/dphi - dT/dtheta on the
 * unit sphere (SHsphtor_to_spat, src/aeros_spectral.f90:545-558; sign convention
 * src/dynamics/aeros_vordiv.f90:19-42). Fourier coefficient of order m:
 *   Vt_m = sum_l [ S D + i m T P/sin ],  Vp_m = sum_l [ i m S P/sin - T D ]. */
void orc_synth_vec(const orc_cfg *c, const double *slm, const double *tlm, double *vt, double *vp)
{
    int nlat = c->nlat, nlon = c->nlon, M = c->mmax;
    size_t rs = 2 * (size_t)(M + 1);
    double *Ft = (double *)calloc((size_t)nlat * rs, sizeof(double));
    double *Fp = (double *)calloc((size_t)nlat * rs, sizeof(double));
    for (int m = 0; m <= M; m++)
        for (int l = m; l <= c->lmax; l++) {
            int lm = LMOF(c, l, m);
            double sr = slm[2 * lm], si = slm[2 * lm + 1];
            double tr = tlm ? tlm[2 * lm] : 0.0, ti = tlm ? tlm[2 * lm + 1] : 0.0;
            const double *p = c->P + (size_t)lm * nlat;
            const double *d = c->D + (size_t)lm * nlat;
            for (int j = 0; j < nlat; j++) {
                double q = m * p[j] / c->sint[j];
                /* i*q*T = (-q ti, q tr); i*q*S = (-q si, q sr) */
                Ft[j * rs + 2 * m]     += sr * d[j] - q * ti;
                Ft[j * rs + 2 * m + 1] += si * d[j] + q * tr;
                Fp[j * rs + 2 * m]     += -q * si - tr * d[j];
                Fp[j * rs + 2 * m + 1] +=  q * sr - ti * d[j];
            }
        }
    for (int j = 0; j < nlat; j++) {
        rdft_inverse(nlon, M, Ft + j * rs, vt + (size_t)j * nlon, c->fft_direct, c->tw_inv);
        rdft_inverse(nlon, M, Fp + j * rs, vp + (size_t)j * nlon, c->fft_direct, c->tw_inv);
    }
    free(Ft);
    free(Fp);
}

/* Gradient on the unit sphere: gt = dF/dtheta, gp = (1/sin) dF/dphi
 * (SHsph_to_spat, src/aeros_spectral.f90:560-584) = vector synthesis with T = 0. */
void orc_synth_grad(const orc_cfg *c, const double *slm, double *gt, double *gp)
{
    orc_synth_vec(c, slm, NULL, gt, gp);
}

/* Inverse of orc_synth_vec (spat_to_SHsphtor, src/aeros_spectral.f90:529-543), by
 * orthogonality of the vector harmonics (norm l(l+1)):
 *   S_lm = 1/(l(l+1)) sum_j w_j [ D Gt - i Q Gp ],
 *   T_lm = 1/(l(l+1)) sum_j w_j [ -i Q Gt - D Gp ],  Q = m P/sin;  l = 0 -> 0. */
void orc_anal_vec(const orc_cfg *c, const double *vt, const double *vp, double *slm, double *tlm)
{
    int nlat = c->nlat, nlon = c->nlon, M = c->mmax;
    size_t rs = 2 * (size_t)(M + 1);
    double *Gt = (double *)malloc(sizeof(double) * (size_t)nlat * rs);
    double *Gp = (double *)malloc(sizeof(double) * (size_t)nlat * rs);
    for (int j = 0; j < nlat; j++) {
        rdft_forward(nlon, M, vt + (size_t)j * nlon, Gt + j * rs, c->fft_direct, c->tw_fwd);
        rdft_forward(nlon, M, vp + (size_t)j * nlon, Gp + j * rs, c->fft_direct, c->tw_fwd);
        double s = c->w[j] * 2.0 * ORC_PI / nlon;
        for (size_t k = 0; k < rs; k++) { Gt[j * rs + k] *= s; Gp[j * rs + k] *= s; }
    }
    for (int m = 0; m <= M; m++)
        for (int l = m; l <= c->lmax; l++) {
            int lm = LMOF(c, l, m);
            if (l == 0) {
                slm[0] = slm[1] = tlm[0] = tlm[1] = 0.0;
                continue;
            }
            const double *p = c->P + (size_t)lm * nlat;
            const double *d = c->D + (size_t)lm * nlat;
            double ar = 0, ai = 0, br = 0, bi = 0;
            for (int j = 0; j < nlat; j++) {
                double q = m * p[j] / c->sint[j];
                double gtr = Gt[j * rs + 2 * m], gti = Gt[j * rs + 2 * m + 1];
                double gpr = Gp[j * rs + 2 * m], gpi = Gp[j * rs + 2 * m + 1];
                /* -i q Gp = (q gpi, -q gpr);  -i q Gt = (q gti, -q gtr) */
                ar += d[j] * gtr + q * gpi;
                ai += d[j] * gti - q * gpr;
                br += q * gti - d[j] * gpr;
                bi += -q * gtr - d[j] * gpi;
            }
            double inv = 1.0 / ((double)l * (l + 1));
            slm[2 * lm] = ar * inv;
            slm[2 * lm + 1] = ai * inv;
            tlm[2 * lm] = br * inv;
            tlm[2 * lm + 1] = bi * inv;
        }
    free(Gt);
    free(Gp);
}

/* ------------------------------------------------------ spectral operators */

/* c *= -l(l+1)/a^2, or the inverse with l = 0 set to zero (src/aeros_spectral.f90:588-630).
 * Same arithmetic as the reference: fac formed first, then one multiply. */
void orc_laplacian(const orc_cfg *c, double *qlm, int inverse)
{
    double a2 = ORC_R_EARTH * ORC_R_EARTH;
    for (int lm = 0; lm < c->nlm; lm++) {
        int l = c->l_of_lm[lm];
        double fac;
        if (inverse) {
            if (l == 0) { qlm[2 * lm] = qlm[2 * lm + 1] = 0.0; continue; }
            fac = -a2 / (double)(l * (l + 1));
        } else {
            fac = -(double)(l * (l + 1)) / a2;
        }
        qlm[2 * lm] *= fac;
        qlm[2 * lm + 1] *= fac;
    }
}

/* a^2 sum_j w_j (2 pi/nlon) sum_i f_ij, row sums first (src/aeros_spectral.f90:632-666). */
double orc_surface_integral(const orc_cfg *c, const double *grid)
{
    double dlon = 2.0 * ORC_PI / c->nlon, total = 0.0;
    for (int j = 0; j < c->nlat; j++) {
        double row = 0.0;
        for (int i = 0; i < c->nlon; i++) row += grid[(size_t)j * c->nlon + i];
        total += c->w[j] * row * dlon;
    }
    return total * ORC_R_EARTH * ORC_R_EARTH;
}

/* -------------------------------------------------------- dynamics mapping */

/* S = -a D/(l(l+1)), T = +a zeta/(l(l+1)), l = 0 -> 0; u = V_phi, v = -V_theta
 * (src/dynamics/aeros_vordiv.f90:82-122). */
void orc_uv_from_vordiv(const orc_cfg *c, const double *vor, const double *div, double *u, double *v)
{
    size_t n2 = 2 * (size_t)c->nlm, ng = (size_t)c->nlon * c->nlat;
    double *s = (double *)malloc(sizeof(double) * n2);
    double *t = (double *)malloc(sizeof(double) * n2);
    double *vt = (double *)malloc(sizeof(double) * ng);
    for (int lm = 0; lm < c->nlm; lm++) {
        int l = c->l_of_lm[lm];
        if (l == 0) { s[2 * lm] = s[2 * lm + 1] = t[2 * lm] = t[2 * lm + 1] = 0.0; continue; }
        double fac = ORC_R_EARTH / (double)(l * (l + 1));
        s[2 * lm] = -fac * div[2 * lm];
        s[2 * lm + 1] = -fac * div[2 * lm + 1];
        t[2 * lm] = fac * vor[2 * lm];
        t[2 * lm + 1] = fac * vor[2 * lm + 1];
    }
    orc_synth_vec(c, s, t, vt, u);
    for (size_t k = 0; k < ng; k++) v[k] = -vt[k];
    free(s); free(t); free(vt);
}

/* V_theta = -v, V_phi = u; vector analysis; D = -l(l+1)/a S, zeta = +l(l+1)/a T
 * (src/dynamics/aeros_vordiv.f90:124-168). Inputs preserved. */
void orc_vordiv_from_uv(const orc_cfg *c, const double *u, const double *v, double *vor, double *div)
{
    size_t n2 = 2 * (size_t)c->nlm, ng = (size_t)c->nlon * c->nlat;
    double *vt = (double *)malloc(sizeof(double) * ng);
    double *s = (double *)malloc(sizeof(double) * n2);
    double *t = (double *)malloc(sizeof(double) * n2);
    for (size_t k = 0; k < ng; k++) vt[k] = -v[k];
    orc_anal_vec(c, vt, u, s, t);
    for (int lm = 0; lm < c->nlm; lm++) {
        int l = c->l_of_lm[lm];
        if (l == 0) { vor[2 * lm] = vor[2 * lm + 1] = div[2 * lm] = div[2 * lm + 1] = 0.0; continue; }
        double fac = (double)(l * (l + 1)) / ORC_R_EARTH;
        div[2 * lm] = -fac * s[2 * lm];
        div[2 * lm + 1] = -fac * s[2 * lm + 1];
        vor[2 * lm] = fac * t[2 * lm];
        vor[2 * lm + 1] = fac * t[2 * lm + 1];
    }
    free(vt); free(s); free(t);
}

/* dfdx = gp/a (east), dfdy = -gt/a (north) (src/dynamics/aeros_vordiv.f90:170-205). */
void orc_gradient(const orc_cfg *c, const double *f, double *dfdx, double *dfdy)
{
    size_t ng = (size_t)c->nlon * c->nlat;
    double *gt = (double *)malloc(sizeof(double) * ng);
    orc_synth_grad(c, f, gt, dfdx);
    double rinv = 1.0 / ORC_R_EARTH;
    for (size_t k = 0; k < ng; k++) {
        dfdx[k] = dfdx[k] * rinv;
        dfdy[k] = -gt[k] * rinv;
    }
    free(gt);
}

/* ------------------------------------------------- tendency transform halves */

/* Spectral -> grid half of aeros_tendency_grid (src/dynamics/aeros_tendency.f90:403-419):
 * lnps synthesis and gradient; per level (u, v), zeta, D, T, grad T. */
void orc_tendency_grid_transforms(const orc_cfg *c, int nlev,
        const double *vor, const double *div, const double *temp, const double *lnps,
        double *u, double *v, double *vor_g, double *div_g, double *t_g,
        double *dtdx, double *dtdy, double *lnps_g, double *dlnpsdx, double *dlnpsdy)
{
    size_t n2 = 2 * (size_t)c->nlm, ng = (size_t)c->nlon * c->nlat;
    orc_synth(c, lnps, lnps_g);
    orc_gradient(c, lnps, dlnpsdx, dlnpsdy);
    for (int k = 0; k < nlev; k++) {
        orc_uv_from_vordiv(c, vor + k * n2, div + k * n2, u + k * ng, v + k * ng);
        orc_synth(c, vor + k * n2, vor_g + k * ng);
        orc_synth(c, div + k * n2, div_g + k * ng);
        orc_synth(c, temp + k * n2, t_g + k * ng);
        orc_gradient(c, temp + k * n2, dtdx + k * ng, dtdy + k * ng);
    }
}

/* Grid -> spectral half, the spectral tendency assembly (src/dynamics/aeros_tendency.f90:428-473):
 * (zeta_t, D_t) = vordiv_from_uv(ae, an); D_t -= laplacian(analysis(ek)); T_t = analysis(dtdt);
 * lnps_t = analysis(dlnpsdt). */
void orc_tendency_spectral(const orc_cfg *c, int nlev,
        const double *ae, const double *an, const double *ek, const double *dtdt,
        const double *dlnpsdt, double *tvor, double *tdiv, double *ttemp, double *tlnps)
{
    size_t n2 = 2 * (size_t)c->nlm, ng = (size_t)c->nlon * c->nlat;
    double *eklm = (double *)malloc(sizeof(double) * n2);
    for (int k = 0; k < nlev; k++) {
        orc_vordiv_from_uv(c, ae + k * ng, an + k * ng, tvor + k * n2, tdiv + k * n2);
        orc_anal(c, ek + k * ng, eklm);
        orc_laplacian(c, eklm, 0);
        for (size_t q = 0; q < n2; q++) tdiv[k * n2 + q] -= eklm[q];
        orc_anal(c, dtdt + k * ng, ttemp + k * n2);
    }
    orc_anal(c, dlnpsdt, tlnps);
    free(eklm);
}

/* ----------------------------------------------------------- timing harness */

/* Mode-B harness (drivers/bench_sht.f90:142-187, drivers/bench_m0a.F90:179-236): an
 * OpenMP parallel loop, schedule(static), over nlev independent units; each unit is
 * one scalar synthesis followed by one scalar analysis. Returns seconds per unit
 * (one synthesis + one analysis = 2 scalar-equivalent transforms), averaged over
 * nrep repetitions after one warm-up synthesis pass. */
double orc_bench_set(const orc_cfg *c, int nlev, int nrep, int nthreads,
                     const double *spec_in, double *grid_work, double *spec_out)
{
    size_t n2 = 2 * (size_t)c->nlm, ng = (size_t)c->nlon * c->nlat;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < nlev; k++) orc_synth(c, spec_in + k * n2, grid_work + k * ng);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int rep = 0; rep < nrep; rep++) {
#pragma omp parallel for schedule(static)
        for (k = 0; k < nlev; k++) {
            orc_synth(c, spec_in + k * n2, grid_work + k * ng);
            orc_anal(c, grid_work + k * ng, spec_out + k * n2);
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double dt = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    return dt / ((double)nrep * nlev);
}
