// step_device.h -- gfx950 kernels of the spectral step completion of a
// semi-implicit leapfrog step (reference src/dynamics/aeros_timestep.f90:1043-1100,
// src/dynamics/aeros_semiimp.f90:345-571): the vorticity leapfrog, the
// semi-implicit (or explicit) update of D, T and ln p_s, implicit horizontal
// diffusion and the Robert-Asselin-Williams filter in one pass, and the mass
// fixer's FP64 global sum.
//
// The semi-implicit operators act on the levels of one spectral coefficient and
// are real, so the real and imaginary parts are independent columns. A block
// owns one degree l (its Helmholtz matrix) and 16 of that degree's 2(l+1) real
// columns, with 16 level lanes per column: every stage is a small matrix-vector
// product over LDS-resident columns. The Helmholtz solve applies the FP64
// inverse of I + (alpha h)^2 lambda_l W, which the host builds from the same
// pivoted LU factors the reference uses.
#pragma once
#include <hip/hip_runtime.h>

#define AER_SI_CB 16        // real columns per block
#define AER_SI_KL 16        // level lanes per block
#define AER_SI_MAXLEV 64

namespace aer {

enum StepFlag { ST_SEMI = 1, ST_DIFF = 2, ST_RAW = 4, ST_VOR = 8, ST_LINEAR = 16 };

struct StepArgs {
    int nlev, lmax, nlm, flags;
    double h, a1, astep, hh_old;            // step, decentering, alpha h, alpha (1-alpha) h^2
    double half_nu, raw_a, raw_b;           // RAW filter: eps/2, alpha, 1 - alpha
    const double *G, *tau, *W, *nu, *pcoef; // linear operators (row-major nlev x nlev, nlev)
    const double *Minv;                     // (lmax+1) x nlev x nlev
    const double *fvt, *fdiv;               // diffusion factors [nlev][lmax+1]
    const int *task_l, *task_c0;            // block -> degree, first real column of the degree
    const double *ovor, *odiv, *otemp, *olnps;   // X^(n-1)
    double *cvor, *cdiv, *ctemp, *clnps;         // X^n (the filter updates it in place)
    const double *tvor, *tdiv, *ttemp, *tlnps;   // R(X^n)
    double *nvor, *ndiv, *ntemp, *nlnps;         // X^(n+1); with ST_LINEAR: L(X^n) of div, temp, lnps
};

__global__ __launch_bounds__(256) void k_step(const StepArgs S)
{
    extern __shared__ double lds[];
    const int n = S.nlev;
    const int c = threadIdx.x % AER_SI_CB, kl = threadIdx.x / AER_SI_CB;
    const int l = S.task_l[blockIdx.x], cg = S.task_c0[blockIdx.x] + c;
    const bool live = cg < 2 * (l + 1);
    const int m = live ? cg / 2 : 0, part = cg & 1;
    const int lm = m * (2 * S.lmax + 3 - m) / 2 + (l - m);   // 0-based m-major index
    const long q = 2L * lm + part;                           // real index in a level
    const long ls = 2L * S.nlm;                              // level stride in doubles
    const double lam = (double)(l * (l + 1)) / (6.371e6 * 6.371e6);
    constexpr int NR = AER_SI_MAXLEV / AER_SI_KL;
    const int NC = AER_SI_CB;
    double *DOLD = lds, *DNOW = DOLD + n * NC, *TOLD = DNOW + n * NC, *TNOW = TOLD + n * NC;
    double *RD = TNOW + n * NC, *RT = RD + n * NC, *LT = RT + n * NC, *LD = LT + n * NC;
    double *TS = LD + n * NC, *RHS = TS + n * NC, *DBAR = RHS + n * NC;
    double *PC = DBAR + n * NC;   // per column: pold, pnow, rp, lpnow, pstar
    const bool semi = S.flags & ST_SEMI, lin = S.flags & ST_LINEAR;

    // 0. loads: every own level of the column, plus the column's ln p_s
    double dn[NR], tn[NR];   // X^(n+1) of this thread's levels
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_SI_KL;
        if (k >= n) break;
        const long x = k * ls + q;
        const int o = k * NC + c;
        DOLD[o] = (live && !lin) ? S.odiv[x] : 0.0;
        DNOW[o] = live ? S.cdiv[x] : 0.0;
        TOLD[o] = (live && !lin) ? S.otemp[x] : 0.0;
        TNOW[o] = live ? S.ctemp[x] : 0.0;
        RD[o] = (live && !lin) ? S.tdiv[x] : 0.0;
        RT[o] = (live && !lin) ? S.ttemp[x] : 0.0;
    }
    if (kl == 0) {
        PC[0 * NC + c] = (live && !lin) ? S.olnps[q] : 0.0;
        PC[1 * NC + c] = live ? S.clnps[q] : 0.0;
        PC[2 * NC + c] = (live && !lin) ? S.tlnps[q] : 0.0;
    }
    __syncthreads();

    if (semi || lin) {
        // 1. L(X^n): L_D = lam (pcoef P + G T), L_T = -tau D, L_P = -nu . D
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            double acc = S.pcoef[k] * PC[1 * NC + c];
            for (int j = k; j < n; ++j) acc = acc + S.G[k * n + j] * TNOW[j * NC + c];
            const double ld = lam * acc;
            double at = 0.0;
            for (int j = 0; j <= k; ++j) at = at + S.tau[k * n + j] * DNOW[j * NC + c];
            const int o = k * NC + c;
            LD[o] = ld;
            LT[o] = -at;
            TS[o] = TOLD[o] + S.astep * (RT[o] - (-at));
        }
        if (kl == 0) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = acc + S.nu[k] * DNOW[k * NC + c];
            const double lp = -acc;
            PC[3 * NC + c] = lp;
            PC[4 * NC + c] = PC[0 * NC + c] + S.astep * (PC[2 * NC + c] - lp);
        }
        __syncthreads();
        if (lin) {
            if (!live) return;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int k = kl + r * AER_SI_KL;
                if (k >= n) break;
                const long x = k * ls + q;
                S.ndiv[x] = LD[k * NC + c];
                S.ntemp[x] = LT[k * NC + c];
            }
            if (kl == 0) S.nlnps[q] = PC[3 * NC + c];
            return;
        }
        // 2. Helmholtz right-hand side
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            double acc = S.pcoef[k] * PC[4 * NC + c];
            for (int j = k; j < n; ++j) acc = acc + S.G[k * n + j] * TS[j * NC + c];
            const double ldstar = lam * acc;
            double w = 0.0;
            for (int j = 0; j < n; ++j) w = w + S.W[k * n + j] * DOLD[j * NC + c];
            const int o = k * NC + c;
            RHS[o] = DOLD[o] - S.hh_old * lam * w + S.h * (RD[o] + ldstar - LD[o]);
        }
        __syncthreads();
        // 3. solve with the degree's inverse, then the decentered average
        const double *Mi = S.Minv + (long)l * n * n;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc = acc + Mi[k * n + j] * RHS[j * NC + c];
            dn[r] = acc;
            const int o = k * NC + c;
            DBAR[o] = S.a1 * acc + (1.0 - S.a1) * DOLD[o];
        }
        __syncthreads();
        // 4. back-substitution for T and ln p_s
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            double acc = 0.0;
            for (int j = 0; j <= k; ++j) acc = acc + S.tau[k * n + j] * DBAR[j * NC + c];
            const int o = k * NC + c;
            tn[r] = TOLD[o] + S.h * (RT[o] - acc - LT[o]);
        }
    } else {
        // fully explicit leapfrog (src/dynamics/aeros_timestep.f90:1054-1064)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            const int o = k * NC + c;
            dn[r] = DOLD[o] + S.h * RD[o];
            tn[r] = TOLD[o] + S.h * RT[o];
        }
    }
    if (!live) return;
    double pn = 0.0;
    if (kl == 0) {
        if (semi) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = acc + S.nu[k] * DBAR[k * NC + c];
            pn = PC[0 * NC + c] + S.h * (PC[2 * NC + c] + (-acc) - PC[3 * NC + c]);
        } else {
            pn = PC[0 * NC + c] + S.h * PC[2 * NC + c];
        }
    }
    // 5. vorticity leapfrog, diffusion, filter, stores
    const bool vor = S.flags & ST_VOR, diff = S.flags & ST_DIFF, raw = S.flags & ST_RAW;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_SI_KL;
        if (k >= n) break;
        const long x = k * ls + q;
        const int o = k * NC + c;
        double nv = 0.0, ov = 0.0, cv = 0.0;
        if (vor) {
            ov = S.ovor[x];
            nv = ov + S.h * S.tvor[x];
        }
        double nd = dn[r], nt = tn[r];
        if (diff) {
            const double fv = S.fvt[k * (S.lmax + 1) + l], fd = S.fdiv[k * (S.lmax + 1) + l];
            nv = fv * nv;
            nd = fd * nd;
            nt = fv * nt;
        }
        if (raw) {
            double cd = DNOW[o], ct = TNOW[o];
            double d = S.half_nu * (DOLD[o] - 2.0 * cd + nd);
            cd = cd + S.raw_a * d;
            nd = nd - S.raw_b * d;
            d = S.half_nu * (TOLD[o] - 2.0 * ct + nt);
            ct = ct + S.raw_a * d;
            nt = nt - S.raw_b * d;
            S.cdiv[x] = cd;
            S.ctemp[x] = ct;
            if (vor) {
                cv = S.cvor[x];
                d = S.half_nu * (ov - 2.0 * cv + nv);
                cv = cv + S.raw_a * d;
                nv = nv - S.raw_b * d;
                S.cvor[x] = cv;
            }
        }
        if (vor) S.nvor[x] = nv;
        S.ndiv[x] = nd;
        S.ntemp[x] = nt;
    }
    if (kl == 0) {
        if (raw) {
            double cp = PC[1 * NC + c];
            const double d = S.half_nu * (PC[0 * NC + c] - 2.0 * cp + pn);
            cp = cp + S.raw_a * d;
            pn = pn - S.raw_b * d;
            S.clnps[q] = cp;
        }
        S.nlnps[q] = pn;
    }
}

// Global dry-air mass of ln p_s on the grid (src/dynamics/aeros_timestep.f90:1250-1276,
// the quadrature of src/aeros_spectral.f90:632-666: row sums first, rows summed in
// order, FP64), then, if fix, ln p_s(0,0) += ln(target/mass) c00. One block.
__global__ __launch_bounds__(256) void k_mass(const double *lnps_g, int nlon, int nlat, const double *w,
                                              double grav, int fix, double target, double c00,
                                              double *lnps, double *mass_out, double *lnr_out)
{
    extern __shared__ double rows[];   // nlat row totals, then 4 partials per row
    double *part = rows + nlat;
    const double dlon = 2.0 * 3.14159265358979323846264338327950288 / (double)nlon;
    // four threads per row, each a contiguous quarter; the quarters are added in order
    const int qn = (nlon + 3) / 4;
    for (int t = threadIdx.x; t < 4 * nlat; t += blockDim.x) {
        const int j = t >> 2, q = t & 3;
        const int i0 = q * qn, i1 = min(nlon, i0 + qn);
        double p = 0.0;
        for (int i = i0; i < i1; ++i) p = p + exp(lnps_g[(long)j * nlon + i]);
        part[t] = p;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nlat; j += blockDim.x) {
        const double row = ((part[4 * j] + part[4 * j + 1]) + part[4 * j + 2]) + part[4 * j + 3];
        rows[j] = w[j] * row * dlon;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double total = 0.0;
    for (int j = 0; j < nlat; ++j) total = total + rows[j];
    const double mass = total * 6.371e6 * 6.371e6 / grav;
    if (mass_out) mass_out[0] = mass;
    if (fix) {
        const double lnr = log(target / mass);
        lnps[0] = lnps[0] + lnr * c00;
        if (lnr_out) lnr_out[0] = lnr;
    }
}

}  // namespace aer
