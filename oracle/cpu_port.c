/*
 * cpu_port.c -- the oracle's scalar transforms arranged for CPU speed. Test and
 * measurement infrastructure only: bench.py's CPU-baseline leg times it, tests
 * check it against orc_synth / orc_anal. Nothing in the product links it.
 *
 * Same algorithm and conventions as orc_synth / orc_anal (Gauss-Legendre quadrature,
 * orthonormal P_lm without the Condon-Shortley phase, a real FFT over longitude;
 * reference src/aeros_spectral.f90:486-515 for the SHTns calls they stand for), with
 * the arrangement a tuned CPU transform uses:
 *   - hemispheric folding: P_lm(-mu) = (-1)^(l-m) P_lm(mu), so a latitude pair costs one
 *     even-degree and one odd-degree sum;
 *   - tables split by degree parity, laid out [m][degree][pair], so the synthesis inner
 *     loop is a contiguous axpy over latitude pairs and the analysis one a contiguous
 *     dot product;
 *   - the real FFT of length nlon through a complex Stockham FFT of length nlon/2.
 * The two transforms are compiled as clones for AVX2, for FMA (with AVX) and for the
 * baseline ISA, one picked at load time, so the library runs on any x86-64 host.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime */
#pragma GCC optimize("O3")

#include "sht_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PORT_PI 3.14159265358979323846264338327950288

#if defined(__x86_64__) && defined(__GNUC__) && !defined(__clang__)
#define PORT_CLONES __attribute__((target_clones("avx2,fma", "default")))
#else
#define PORT_CLONES
#endif

struct orc_port {
    int T, nlat, nlon, nh, nc, nlm;
    double *Pe, *Po;       /* parity tables, order m at Pe + oe[m] (rows of nh values) */
    size_t *oe, *oo;
    int *ne, *no;          /* even / odd degree counts of order m */
    int *lm0;              /* 0-based index of (l = m, m) */
    double *wq;            /* w_j * 2 pi / nlon for the north row of pair j */
    int nrad, rad[32];
    double *twc_f, *twc_i; /* e^{-+2 pi i k/nc}, k < nc */
    double *twr_f, *twr_i; /* e^{-+2 pi i k/nlon}, k <= nc */
};

static double *twiddles(int n, int count, int sign)
{
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)count);
    for (int k = 0; k < count; k++) {
        const double a = 2.0 * PORT_PI * (double)k / (double)n;
        t[2 * k] = cos(a);
        t[2 * k + 1] = sign * sin(a);
    }
    return t;
}

orc_port *orc_port_create(const orc_cfg *c)
{
    if (!c || c->nlat % 2 || c->nlon % 2 || c->lmax != c->mmax || 2 * c->mmax >= c->nlon) return NULL;
    orc_port *p = (orc_port *)calloc(1, sizeof(orc_port));
    const int T = c->lmax, nh = c->nlat / 2;
    p->T = T;
    p->nlat = c->nlat;
    p->nlon = c->nlon;
    p->nh = nh;
    p->nc = c->nlon / 2;
    p->nlm = c->nlm;
    /* radices of the half-length complex FFT: 4s first, then 2, 3, 5 */
    int n = p->nc;
    while (n % 4 == 0) { p->rad[p->nrad++] = 4; n /= 4; }
    while (n % 2 == 0) { p->rad[p->nrad++] = 2; n /= 2; }
    while (n % 3 == 0) { p->rad[p->nrad++] = 3; n /= 3; }
    while (n % 5 == 0) { p->rad[p->nrad++] = 5; n /= 5; }
    if (n != 1) {
        free(p);
        return NULL;
    }
    p->oe = (size_t *)malloc(sizeof(size_t) * (T + 1));
    p->oo = (size_t *)malloc(sizeof(size_t) * (T + 1));
    p->ne = (int *)malloc(sizeof(int) * (T + 1));
    p->no = (int *)malloc(sizeof(int) * (T + 1));
    p->lm0 = (int *)malloc(sizeof(int) * (T + 1));
    size_t se = 0, so = 0;
    for (int m = 0; m <= T; m++) {
        p->ne[m] = (T - m) / 2 + 1;
        p->no[m] = (T - m + 1) / 2;
        p->oe[m] = se;
        p->oo[m] = so;
        se += (size_t)p->ne[m] * nh;
        so += (size_t)p->no[m] * nh;
        p->lm0[m] = m * (2 * T + 3 - m) / 2;
    }
    p->Pe = (double *)malloc(sizeof(double) * (se ? se : 1));
    p->Po = (double *)malloc(sizeof(double) * (so ? so : 1));
    for (int m = 0; m <= T; m++)
        for (int l = m; l <= T; l++) {
            const double *src = c->P + (size_t)(p->lm0[m] + l - m) * c->nlat;
            const int i = (l - m) / 2;
            double *dst = ((l - m) % 2 == 0) ? p->Pe + p->oe[m] + (size_t)i * nh : p->Po + p->oo[m] + (size_t)i * nh;
            memcpy(dst, src, sizeof(double) * nh);   /* north rows j < nh: mu_j > 0 */
        }
    p->wq = (double *)malloc(sizeof(double) * nh);
    for (int j = 0; j < nh; j++) p->wq[j] = c->w[j] * 2.0 * PORT_PI / c->nlon;
    p->twc_f = twiddles(p->nc, p->nc, -1);
    p->twc_i = twiddles(p->nc, p->nc, +1);
    p->twr_f = twiddles(p->nlon, p->nc + 1, -1);
    p->twr_i = twiddles(p->nlon, p->nc + 1, +1);
    return p;
}

void orc_port_destroy(orc_port *p)
{
    if (!p) return;
    free(p->Pe); free(p->Po); free(p->oe); free(p->oo); free(p->ne); free(p->no); free(p->lm0);
    free(p->wq); free(p->twc_f); free(p->twc_i); free(p->twr_f); free(p->twr_i);
    free(p);
}

/* Complex FFT of length n, Y_k = sum_j x_j e^{s 2 pi i jk/n} with tw = e^{s 2 pi i t/n}:
 * self-sorting Stockham passes. A pass of radix R over the product Ns of the earlier
 * radices takes x_{j + r n/R} (r < R) for j = jb + k (jb a multiple of Ns, k < Ns),
 * multiplies by the twiddle e^{s 2 pi i r k/(Ns R)}, applies the size-R DFT and writes
 * output q to jb R + k + q Ns. Radix 4 and 2 have their own loops; 3 and 5 share one. */
static inline __attribute__((always_inline)) void pass4(int n, int Ns, const double *a, double *b, const double *tw,
                                                        double s)
{
    const int NR = n / 4, tstep = n / (Ns * 4);
    for (int jb = 0; jb < NR; jb += Ns)
        for (int k = 0; k < Ns; k++) {
            const int j = jb + k;
            const double *x0 = a + 2 * j, *x1 = a + 2 * (j + NR), *x2 = a + 2 * (j + 2 * NR), *x3 = a + 2 * (j + 3 * NR);
            const double *w1 = tw + 2 * (k * tstep), *w2 = tw + 2 * (2 * k * tstep), *w3 = tw + 2 * (3 * k * tstep);
            const double v0r = x0[0], v0i = x0[1];
            const double v1r = x1[0] * w1[0] - x1[1] * w1[1], v1i = x1[0] * w1[1] + x1[1] * w1[0];
            const double v2r = x2[0] * w2[0] - x2[1] * w2[1], v2i = x2[0] * w2[1] + x2[1] * w2[0];
            const double v3r = x3[0] * w3[0] - x3[1] * w3[1], v3i = x3[0] * w3[1] + x3[1] * w3[0];
            /* e^{s 2 pi i qr/4} = (s i)^(qr) */
            const double ar = v0r + v2r, ai = v0i + v2i, br = v0r - v2r, bi = v0i - v2i;
            const double cr = v1r + v3r, ci = v1i + v3i, dr = v1r - v3r, di = v1i - v3i;
            double *y = b + 2 * (jb * 4 + k);
            y[0] = ar + cr;
            y[1] = ai + ci;
            y[2 * Ns] = br - s * di;
            y[2 * Ns + 1] = bi + s * dr;
            y[4 * Ns] = ar - cr;
            y[4 * Ns + 1] = ai - ci;
            y[6 * Ns] = br + s * di;
            y[6 * Ns + 1] = bi - s * dr;
        }
}

static inline __attribute__((always_inline)) void pass2(int n, int Ns, const double *a, double *b, const double *tw)
{
    const int NR = n / 2, tstep = n / (Ns * 2);
    for (int jb = 0; jb < NR; jb += Ns)
        for (int k = 0; k < Ns; k++) {
            const int j = jb + k;
            const double *x0 = a + 2 * j, *x1 = a + 2 * (j + NR), *w1 = tw + 2 * (k * tstep);
            const double v1r = x1[0] * w1[0] - x1[1] * w1[1], v1i = x1[0] * w1[1] + x1[1] * w1[0];
            double *y = b + 2 * (jb * 2 + k);
            y[0] = x0[0] + v1r;
            y[1] = x0[1] + v1i;
            y[2 * Ns] = x0[0] - v1r;
            y[2 * Ns + 1] = x0[1] - v1i;
        }
}

static inline __attribute__((always_inline)) void pass_odd(int n, int R, int Ns, const double *a, double *b,
                                                           const double *tw)
{
    const int NR = n / R, tstep = n / (Ns * R);
    for (int jb = 0; jb < NR; jb += Ns)
        for (int k = 0; k < Ns; k++) {
            const int j = jb + k;
            double vr[5], vi[5];
            for (int r = 0; r < R; r++) {
                const double xr = a[2 * (j + r * NR)], xi = a[2 * (j + r * NR) + 1];
                const int t = r * k * tstep;   /* < R Ns tstep = n */
                vr[r] = xr * tw[2 * t] - xi * tw[2 * t + 1];
                vi[r] = xr * tw[2 * t + 1] + xi * tw[2 * t];
            }
            for (int q = 0; q < R; q++) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < R; r++) {
                    const int t = ((r * q) % R) * NR;
                    sr += vr[r] * tw[2 * t] - vi[r] * tw[2 * t + 1];
                    si += vr[r] * tw[2 * t + 1] + vi[r] * tw[2 * t];
                }
                b[2 * (jb * R + k + q * Ns)] = sr;
                b[2 * (jb * R + k + q * Ns) + 1] = si;
            }
        }
}

/* a and b are two buffers of n complex values; returns the one holding the result. */
static inline __attribute__((always_inline)) double *stockham(int n, const int *rad, int nrad, double *a, double *b,
                                                               const double *tw, int s)
{
    int Ns = 1;
    for (int ps = 0; ps < nrad; ps++) {
        const int R = rad[ps];
        if (R == 4) pass4(n, Ns, a, b, tw, (double)s);
        else if (R == 2) pass2(n, Ns, a, b, tw);
        else pass_odd(n, R, Ns, a, b, tw);
        double *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

/* The same passes over PV latitude rows at once, real and imaginary parts in separate
 * planes laid out [point][row]: every butterfly is a short vector loop over the rows. */
#define PV 8

static inline __attribute__((always_inline)) void vpass4(int n, int Ns, const double *ar, const double *ai, double *br,
                                                         double *bi, const double *tw, double s)
{
    const int NR = n / 4, tstep = n / (Ns * 4);
    for (int jb = 0; jb < NR; jb += Ns)
        for (int k = 0; k < Ns; k++) {
            const int j = jb + k;
            const double w1r = tw[2 * (k * tstep)], w1i = tw[2 * (k * tstep) + 1];
            const double w2r = tw[2 * (2 * k * tstep)], w2i = tw[2 * (2 * k * tstep) + 1];
            const double w3r = tw[2 * (3 * k * tstep)], w3i = tw[2 * (3 * k * tstep) + 1];
            const double *x0r = ar + (size_t)j * PV, *x1r = ar + (size_t)(j + NR) * PV;
            const double *x2r = ar + (size_t)(j + 2 * NR) * PV, *x3r = ar + (size_t)(j + 3 * NR) * PV;
            const double *x0i = ai + (size_t)j * PV, *x1i = ai + (size_t)(j + NR) * PV;
            const double *x2i = ai + (size_t)(j + 2 * NR) * PV, *x3i = ai + (size_t)(j + 3 * NR) * PV;
            const size_t d = (size_t)(jb * 4 + k) * PV, st = (size_t)Ns * PV;
            double *y0r = br + d, *y1r = y0r + st, *y2r = y1r + st, *y3r = y2r + st;
            double *y0i = bi + d, *y1i = y0i + st, *y2i = y1i + st, *y3i = y2i + st;
#pragma omp simd
            for (int v = 0; v < PV; v++) {
                const double v1r = x1r[v] * w1r - x1i[v] * w1i, v1i = x1r[v] * w1i + x1i[v] * w1r;
                const double v2r = x2r[v] * w2r - x2i[v] * w2i, v2i = x2r[v] * w2i + x2i[v] * w2r;
                const double v3r = x3r[v] * w3r - x3i[v] * w3i, v3i = x3r[v] * w3i + x3i[v] * w3r;
                const double a_r = x0r[v] + v2r, a_i = x0i[v] + v2i, b_r = x0r[v] - v2r, b_i = x0i[v] - v2i;
                const double c_r = v1r + v3r, c_i = v1i + v3i, d_r = v1r - v3r, d_i = v1i - v3i;
                y0r[v] = a_r + c_r;
                y0i[v] = a_i + c_i;
                y1r[v] = b_r - s * d_i;
                y1i[v] = b_i + s * d_r;
                y2r[v] = a_r - c_r;
                y2i[v] = a_i - c_i;
                y3r[v] = b_r + s * d_i;
                y3i[v] = b_i - s * d_r;
            }
        }
}

static inline __attribute__((always_inline)) void vpass2(int n, int Ns, const double *ar, const double *ai, double *br,
                                                         double *bi, const double *tw)
{
    const int NR = n / 2, tstep = n / (Ns * 2);
    for (int jb = 0; jb < NR; jb += Ns)
        for (int k = 0; k < Ns; k++) {
            const int j = jb + k;
            const double w1r = tw[2 * (k * tstep)], w1i = tw[2 * (k * tstep) + 1];
            const double *x0r = ar + (size_t)j * PV, *x1r = ar + (size_t)(j + NR) * PV;
            const double *x0i = ai + (size_t)j * PV, *x1i = ai + (size_t)(j + NR) * PV;
            const size_t d = (size_t)(jb * 2 + k) * PV, st = (size_t)Ns * PV;
            double *y0r = br + d, *y1r = y0r + st, *y0i = bi + d, *y1i = y0i + st;
#pragma omp simd
            for (int v = 0; v < PV; v++) {
                const double v1r = x1r[v] * w1r - x1i[v] * w1i, v1i = x1r[v] * w1i + x1i[v] * w1r;
                y0r[v] = x0r[v] + v1r;
                y0i[v] = x0i[v] + v1i;
                y1r[v] = x0r[v] - v1r;
                y1i[v] = x0i[v] - v1i;
            }
        }
}

static inline __attribute__((always_inline)) void vpass_odd(int n, int R, int Ns, const double *ar, const double *ai,
                                                            double *br, double *bi, const double *tw)
{
    const int NR = n / R, tstep = n / (Ns * R);
    for (int jb = 0; jb < NR; jb += Ns)
        for (int k = 0; k < Ns; k++) {
            const int j = jb + k;
            double vr[5][PV], vi[5][PV];
            for (int r = 0; r < R; r++) {
                const int t = r * k * tstep;
                const double wr = tw[2 * t], wi = tw[2 * t + 1];
                const double *xr = ar + (size_t)(j + r * NR) * PV, *xi = ai + (size_t)(j + r * NR) * PV;
#pragma omp simd
                for (int v = 0; v < PV; v++) {
                    vr[r][v] = xr[v] * wr - xi[v] * wi;
                    vi[r][v] = xr[v] * wi + xi[v] * wr;
                }
            }
            for (int q = 0; q < R; q++) {
                double *yr = br + (size_t)(jb * R + k + q * Ns) * PV, *yi = bi + (size_t)(jb * R + k + q * Ns) * PV;
                for (int v = 0; v < PV; v++) yr[v] = yi[v] = 0.0;
                for (int r = 0; r < R; r++) {
                    const int t = ((r * q) % R) * NR;
                    const double wr = tw[2 * t], wi = tw[2 * t + 1];
#pragma omp simd
                    for (int v = 0; v < PV; v++) {
                        yr[v] += vr[r][v] * wr - vi[r][v] * wi;
                        yi[v] += vr[r][v] * wi + vi[r][v] * wr;
                    }
                }
            }
        }
}

/* buf: four planes of n x PV values (a re, a im, b re, b im); the input is in the a
 * planes; returns the plane pair holding the result (0: a, 1: b). */
static inline __attribute__((always_inline)) int vstockham(int n, const int *rad, int nrad, double *buf,
                                                           const double *tw, int s)
{
    double *ar = buf, *ai = ar + (size_t)n * PV, *br = ai + (size_t)n * PV, *bi = br + (size_t)n * PV;
    int Ns = 1, where = 0;
    for (int ps = 0; ps < nrad; ps++) {
        const int R = rad[ps];
        if (R == 4) vpass4(n, Ns, ar, ai, br, bi, tw, (double)s);
        else if (R == 2) vpass2(n, Ns, ar, ai, br, bi, tw);
        else vpass_odd(n, R, Ns, ar, ai, br, bi, tw);
        double *t = ar; ar = br; br = t;
        t = ai; ai = bi; bi = t;
        where ^= 1;
        Ns *= R;
    }
    return where;
}

/* Work per call: nlat (T+1) complex Fourier values, the batched FFT's four planes of
 * nc x PV values, and the analysis' folded sums (4 nh). */
size_t orc_port_work_doubles(const orc_port *p)
{
    return 2 * (size_t)p->nlat * (p->T + 1) + 4 * (size_t)p->nc * PV + 4 * (size_t)p->nh;
}

/* grid(nlon, nlat) = synthesis of qlm (as orc_synth). */
PORT_CLONES
void orc_port_synth(const orc_port *p, const double *qlm, double *grid, double *work)
{
    const int T = p->T, nh = p->nh, nlat = p->nlat, nc = p->nc, M1 = T + 1;
    double *F = work;                              /* [nlat][T+1] complex */
    double *A = F + 2 * (size_t)nlat * M1;        /* batched FFT planes */
    for (int m = 0; m <= T; m++) {
        const double *q = qlm + 2 * (size_t)p->lm0[m];
        const double *pe = p->Pe + p->oe[m], *po = p->Po + p->oo[m];
        const int ne = p->ne[m], no = p->no[m];
        /* pairs in chunks of 8: the chunk's even and odd sums stay in registers over the degrees */
        for (int jc = 0; jc < nh; jc += 8) {
            double er[8] = {0}, ei[8] = {0}, or_[8] = {0}, oi[8] = {0};
            const int nj = nh - jc < 8 ? nh - jc : 8;
            if (nj == 8) {
                for (int i = 0; i < ne; i++) {
                    const double cr = q[4 * i], ci = q[4 * i + 1], *row = pe + (size_t)i * nh + jc;
                    for (int jj = 0; jj < 8; jj++) {
                        er[jj] += cr * row[jj];
                        ei[jj] += ci * row[jj];
                    }
                }
                for (int i = 0; i < no; i++) {
                    const double cr = q[4 * i + 2], ci = q[4 * i + 3], *row = po + (size_t)i * nh + jc;
                    for (int jj = 0; jj < 8; jj++) {
                        or_[jj] += cr * row[jj];
                        oi[jj] += ci * row[jj];
                    }
                }
            } else {
                for (int i = 0; i < ne; i++) {
                    const double cr = q[4 * i], ci = q[4 * i + 1], *row = pe + (size_t)i * nh + jc;
                    for (int jj = 0; jj < nj; jj++) {
                        er[jj] += cr * row[jj];
                        ei[jj] += ci * row[jj];
                    }
                }
                for (int i = 0; i < no; i++) {
                    const double cr = q[4 * i + 2], ci = q[4 * i + 3], *row = po + (size_t)i * nh + jc;
                    for (int jj = 0; jj < nj; jj++) {
                        or_[jj] += cr * row[jj];
                        oi[jj] += ci * row[jj];
                    }
                }
            }
            for (int jj = 0; jj < nj; jj++) {
                const int j = jc + jj;
                double *fn = F + 2 * ((size_t)j * M1 + m), *fs = F + 2 * ((size_t)(nlat - 1 - j) * M1 + m);
                fn[0] = er[jj] + or_[jj];
                fn[1] = ei[jj] + oi[jj];
                fs[0] = er[jj] - or_[jj];
                fs[1] = ei[jj] - oi[jj];
            }
        }
    }
    /* per latitude: x_i = Re X_0 + 2 Re sum_{m>=1} X_m e^{i m phi_i}, as the Hermitian
     * spectrum Y (Y_0 = Re X_0, Y_m = X_m, Y_{n-m} = conj X_m) folded into a half-length
     * transform: W_k = E_k + i O_k, E_k = Y_k + Y_{k+nc}, O_k = (Y_k - Y_{k+nc}) e^{2 pi i k/nlon};
     * then z = x_{2n} + i x_{2n+1}. PV rows per batched FFT. */
    for (int j0 = 0; j0 < nlat; j0 += PV) {
        const int nv = nlat - j0 < PV ? nlat - j0 : PV;
        double *wr_ = A, *wi_ = A + (size_t)nc * PV;
        for (int v = 0; v < PV; v++) {
            double *wr = wr_ + v, *wi = wi_ + v;   /* column v, stride PV */
            if (v >= nv) {
                for (int k = 0; k < nc; k++) wr[(size_t)k * PV] = wi[(size_t)k * PV] = 0.0;
                continue;
            }
            const double *X = F + 2 * (size_t)(j0 + v) * M1;
            for (int k = 0; k < nc; k++) {
                double yr = 0.0, yi = 0.0, zr = 0.0, zi = 0.0;     /* Y_k, Y_{k+nc} = conj Y_{nc-k} */
                if (k <= T) { yr = X[2 * k]; yi = k ? X[2 * k + 1] : 0.0; }
                if (k > 0 && nc - k <= T) { zr = X[2 * (nc - k)]; zi = -X[2 * (nc - k) + 1]; }
                const double tr = p->twr_i[2 * k], ti = p->twr_i[2 * k + 1];
                const double er = yr + zr, ei = yi + zi, dr = yr - zr, di = yi - zi;
                const double or_ = dr * tr - di * ti, oi = dr * ti + di * tr;
                wr[(size_t)k * PV] = er - oi;
                wi[(size_t)k * PV] = ei + or_;
            }
        }
        const int w = vstockham(nc, p->rad, p->nrad, A, p->twc_i, +1);
        const double *zr_ = A + (size_t)(2 * w) * nc * PV, *zi_ = zr_ + (size_t)nc * PV;
        for (int v = 0; v < nv; v++) {
            double *g = grid + (size_t)(j0 + v) * p->nlon;
            for (int n = 0; n < nc; n++) {
                g[2 * n] = zr_[(size_t)n * PV + v];
                g[2 * n + 1] = zi_[(size_t)n * PV + v];
            }
        }
    }
}

/* q[4 i], q[4 i + 1] = sum_j rows[i][j] (xr[j], xi[j]) for i < nrow (rows of nh values):
 * four rows at a time, so each load of x feeds eight sums. */
static inline __attribute__((always_inline)) void port_project(const double *rows, int nrow, int nh, const double *xr,
                                                               const double *xi, double *q)
{
    int i = 0;
    for (; i + 4 <= nrow; i += 4) {
        const double *r0 = rows + (size_t)i * nh, *r1 = r0 + nh, *r2 = r1 + nh, *r3 = r2 + nh;
        double a0 = 0.0, b0 = 0.0, a1 = 0.0, b1 = 0.0, a2 = 0.0, b2 = 0.0, a3 = 0.0, b3 = 0.0;
#pragma omp simd reduction(+ : a0, b0, a1, b1, a2, b2, a3, b3)
        for (int j = 0; j < nh; j++) {
            const double sr = xr[j], si = xi[j];
            a0 += r0[j] * sr; b0 += r0[j] * si;
            a1 += r1[j] * sr; b1 += r1[j] * si;
            a2 += r2[j] * sr; b2 += r2[j] * si;
            a3 += r3[j] * sr; b3 += r3[j] * si;
        }
        q[4 * i] = a0; q[4 * i + 1] = b0;
        q[4 * i + 4] = a1; q[4 * i + 5] = b1;
        q[4 * i + 8] = a2; q[4 * i + 9] = b2;
        q[4 * i + 12] = a3; q[4 * i + 13] = b3;
    }
    for (; i < nrow; i++) {
        const double *row = rows + (size_t)i * nh;
        double a = 0.0, b = 0.0;
#pragma omp simd reduction(+ : a, b)
        for (int j = 0; j < nh; j++) {
            a += row[j] * xr[j];
            b += row[j] * xi[j];
        }
        q[4 * i] = a;
        q[4 * i + 1] = b;
    }
}

/* qlm = analysis of grid(nlon, nlat) (as orc_anal). */
PORT_CLONES
void orc_port_anal(const orc_port *p, const double *grid, double *qlm, double *work)
{
    const int T = p->T, nh = p->nh, nlat = p->nlat, nc = p->nc, M1 = T + 1;
    double *G = work;                              /* [nlat][T+1] complex, weighted */
    double *A = G + 2 * (size_t)nlat * M1;         /* batched FFT planes */
    double *SD = A + 4 * (size_t)nc * PV;          /* folded sums */
    for (int j0 = 0; j0 < nlat; j0 += PV) {
        const int nv = nlat - j0 < PV ? nlat - j0 : PV;
        double *ar = A, *ai = A + (size_t)nc * PV;
        for (int n = 0; n < nc; n++)                /* z_n = x_{2n} + i x_{2n+1} */
            for (int v = 0; v < PV; v++) {
                const double *g = grid + (size_t)(j0 + (v < nv ? v : 0)) * p->nlon;
                ar[(size_t)n * PV + v] = v < nv ? g[2 * n] : 0.0;
                ai[(size_t)n * PV + v] = v < nv ? g[2 * n + 1] : 0.0;
            }
        const int w = vstockham(nc, p->rad, p->nrad, A, p->twc_f, -1);
        const double *Zr = A + (size_t)(2 * w) * nc * PV, *Zi = Zr + (size_t)nc * PV;
        for (int v = 0; v < nv; v++) {
            const int j = j0 + v;
            const double s = p->wq[j < nh ? j : nlat - 1 - j];
            double *X = G + 2 * (size_t)j * M1;
            for (int k = 0; k <= T; k++) {
                const int kc = k ? nc - k : 0;
                const double a_r = Zr[(size_t)k * PV + v], a_i = Zi[(size_t)k * PV + v];
                const double b_r = Zr[(size_t)kc * PV + v], b_i = -Zi[(size_t)kc * PV + v];
                const double er = 0.5 * (a_r + b_r), ei = 0.5 * (a_i + b_i);
                const double or_ = 0.5 * (a_i - b_i), oi = -0.5 * (a_r - b_r);   /* (Z_k - conj Z_{nc-k}) / 2i */
                const double wr = p->twr_f[2 * k], wi = p->twr_f[2 * k + 1];
                X[2 * k] = s * (er + or_ * wr - oi * wi);
                X[2 * k + 1] = s * (ei + or_ * wi + oi * wr);
            }
        }
    }
    double *Sr = SD, *Si = SD + nh, *Dr = SD + 2 * nh, *Di = SD + 3 * nh;
    for (int m = 0; m <= T; m++) {
        for (int j = 0; j < nh; j++) {
            const double *xn = G + 2 * ((size_t)j * M1 + m), *xs = G + 2 * ((size_t)(nlat - 1 - j) * M1 + m);
            Sr[j] = xn[0] + xs[0];
            Si[j] = xn[1] + xs[1];
            Dr[j] = xn[0] - xs[0];
            Di[j] = xn[1] - xs[1];
        }
        double *q = qlm + 2 * (size_t)p->lm0[m];
        port_project(p->Pe + p->oe[m], p->ne[m], nh, Sr, Si, q);
        port_project(p->Po + p->oo[m], p->no[m], nh, Dr, Di, q + 2);
    }
}

/* Complex FFT exposed for the tests: y = sum_j x_j e^{sign 2 pi i jk/n}. Returns 0, or -1
 * when n has a prime factor other than 2, 3, 5. */
int orc_port_cfft(int n, int sign, const double *x, double *y)
{
    int rad[32], nrad = 0, r = n;
    while (r % 4 == 0) { rad[nrad++] = 4; r /= 4; }
    while (r % 2 == 0) { rad[nrad++] = 2; r /= 2; }
    while (r % 3 == 0) { rad[nrad++] = 3; r /= 3; }
    while (r % 5 == 0) { rad[nrad++] = 5; r /= 5; }
    if (r != 1 || n < 1) return -1;
    double *tw = twiddles(n, n, sign >= 0 ? 1 : -1);
    double *a = (double *)malloc(sizeof(double) * 2 * (size_t)n), *b = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(a, x, sizeof(double) * 2 * (size_t)n);
    const double *z = stockham(n, rad, nrad, a, b, tw, sign >= 0 ? 1 : -1);
    memcpy(y, z, sizeof(double) * 2 * (size_t)n);
    free(a); free(b); free(tw);
    return 0;
}

/* The reference's mode-B timing shape (drivers/bench_sht.f90:142-187): an OpenMP loop
 * over nunit independent (field) units, each one synthesis + one analysis, one
 * workspace per thread; one warm-up pass, then nrep timed passes. Seconds per unit. */
double orc_port_bench_set(const orc_port *p, int nunit, int nrep, int nthreads,
                          const double *spec_in, double *grid_work, double *spec_out)
{
    const size_t n2 = 2 * (size_t)p->nlm, ng = (size_t)p->nlon * p->nlat, nw = orc_port_work_doubles(p);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp parallel
    {
        double *work = (double *)malloc(sizeof(double) * nw);
#pragma omp for schedule(static)
        for (int k = 0; k < nunit; k++) orc_port_synth(p, spec_in + k * n2, grid_work + k * ng, work);
#pragma omp single
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (int rep = 0; rep < nrep; rep++) {
#pragma omp for schedule(static)
            for (int k = 0; k < nunit; k++) {
                orc_port_synth(p, spec_in + k * n2, grid_work + k * ng, work);
                orc_port_anal(p, grid_work + k * ng, spec_out + k * n2, work);
            }
        }
        free(work);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    const double dt = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    return dt / ((double)nrep * nunit);
}
