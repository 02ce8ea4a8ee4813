// step_device.h -- gfx950 kernels of the spectral step completion of a
// semi-implicit leapfrog step (reference src/dynamics/aeros_timestep.f90:1043-1100,
// src/dynamics/aeros_semiimp.f90:345-571): the vorticity leapfrog, the
// semi-implicit (or explicit) update of D, T and ln p_s, implicit horizontal
// diffusion and the Robert-Asselin-Williams filter in one pass, and the mass
// fixer's FP64 global sum.
//
// The semi-implicit operators act on the levels of one spectral coefficient and
// are real, so the real and imaginary parts are independent columns. A block
// owns 16 consecutive real entries of a level row (8 consecutive lm, real and
// imaginary parts), with 16 level lanes per column, so every level-row access of
// the block is one contiguous 128-byte segment; each column applies its own
// degree's operators. Every stage is a small matrix-vector product over
// LDS-resident columns. The Helmholtz solve applies the FP64 inverse of
// I + (alpha h)^2 lambda_l W, which the host builds from the same pivoted LU
// factors the reference uses.
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
    const int *l_of_lm;                     // degree of each (m-major) spectral index
    const double *ovor, *odiv, *otemp, *olnps;   // X^(n-1)
    double *cvor, *cdiv, *ctemp, *clnps;         // X^n (the filter updates it in place)
    const double *tvor, *tdiv, *ttemp, *tlnps;   // R(X^n)
    double *nvor, *ndiv, *ntemp, *nlnps;         // X^(n+1); with ST_LINEAR: L(X^n) of div, temp, lnps
    long ms3, ms2;                               // member strides (doubles) of 3-D and ln p_s arrays
};

// Ensemble members of one size are independent copies of the same operators:
// blockIdx.y selects the member and every state array moves by its stride.
template <typename P>
__device__ __forceinline__ void member_shift(P *&p, long off)
{
    if (p) p += off;
}
__device__ __forceinline__ StepArgs member_args(StepArgs S)
{
    const long m3 = (long)blockIdx.y * S.ms3, m2 = (long)blockIdx.y * S.ms2;
    member_shift(S.ovor, m3); member_shift(S.odiv, m3); member_shift(S.otemp, m3); member_shift(S.olnps, m2);
    member_shift(S.cvor, m3); member_shift(S.cdiv, m3); member_shift(S.ctemp, m3); member_shift(S.clnps, m2);
    member_shift(S.tvor, m3); member_shift(S.tdiv, m3); member_shift(S.ttemp, m3); member_shift(S.tlnps, m2);
    member_shift(S.nvor, m3); member_shift(S.ndiv, m3); member_shift(S.ntemp, m3); member_shift(S.nlnps, m2);
    return S;
}

// OPS: the l-independent operators G, tau, W, nu, pcoef are staged in LDS (the host
// picks it when they fit beside the columns); otherwise they are read from memory.
template <bool OPS>
__global__ __launch_bounds__(256) void k_step(const StepArgs S0)
{
    const StepArgs S = member_args(S0);
    extern __shared__ double lds[];
    const int n = S.nlev;
    const int c = threadIdx.x % AER_SI_CB, kl = threadIdx.x / AER_SI_CB;
    const long q = (long)blockIdx.x * AER_SI_CB + c;         // real index in a level
    const bool live = q < 2L * S.nlm;
    const int l = S.l_of_lm[live ? (int)(q >> 1) : 0];
    const long ls = 2L * S.nlm;                              // level stride in doubles
    const double lam = (double)(l * (l + 1)) / (6.371e6 * 6.371e6);
    constexpr int NR = AER_SI_MAXLEV / AER_SI_KL;
    const int NC = AER_SI_CB;
    double *DOLD = lds, *DNOW = DOLD + n * NC, *TOLD = DNOW + n * NC, *TNOW = TOLD + n * NC;
    double *RD = TNOW + n * NC, *RT = RD + n * NC, *LT = RT + n * NC, *LD = LT + n * NC;
    double *TS = LD + n * NC, *RHS = TS + n * NC, *DBAR = RHS + n * NC;
    double *PC = DBAR + n * NC;   // per column: pold, pnow, rp, lpnow, pstar
    double *sops = PC + 5 * NC;   // with OPS: G, tau, W (n x n each), nu, pcoef
    const bool semi = S.flags & ST_SEMI, lin = S.flags & ST_LINEAR;
    if constexpr (OPS) {
        const int nn = n * n;
        for (int i = threadIdx.x; i < nn; i += blockDim.x) {
            sops[i] = S.G[i];
            sops[nn + i] = S.tau[i];
            sops[2 * nn + i] = S.W[i];
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            sops[3 * nn + i] = S.nu[i];
            sops[3 * nn + n + i] = S.pcoef[i];
        }
    }
    const double *Gp = OPS ? sops : S.G, *Tp = OPS ? sops + n * n : S.tau, *Wp = OPS ? sops + 2 * n * n : S.W;
    const double *NUp = OPS ? sops + 3 * n * n : S.nu, *PCp = OPS ? sops + 3 * n * n + n : S.pcoef;

    // 0. loads: every own level of the column, plus the column's ln p_s; the vorticity
    //    levels go to registers here so that no load waits at the end of the step
    double dn[NR], tn[NR];   // X^(n+1) of this thread's levels
    double vo[NR], vt[NR], vc[NR], fv[NR], fd[NR];
    const bool vload = live && !lin && (S.flags & ST_VOR);
    const bool vfilt = vload && (S.flags & ST_RAW);
    const bool fload = !lin && (S.flags & ST_DIFF);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_SI_KL;
        vo[r] = vt[r] = vc[r] = 0.0;
        fv[r] = fd[r] = 1.0;
        if (k >= n) break;
        const long x = k * ls + q;
        const int o = k * NC + c;
        if (fload) {
            fv[r] = S.fvt[k * (S.lmax + 1) + l];
            fd[r] = S.fdiv[k * (S.lmax + 1) + l];
        }
        if (vload) {
            vo[r] = S.ovor[x];
            vt[r] = S.tvor[x];
        }
        if (vfilt) vc[r] = S.cvor[x];
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
            double acc = PCp[k] * PC[1 * NC + c];
#pragma unroll 8
            for (int j = k; j < n; ++j) acc = acc + Gp[k * n + j] * TNOW[j * NC + c];
            const double ld = lam * acc;
            double at = 0.0;
#pragma unroll 8
            for (int j = 0; j <= k; ++j) at = at + Tp[k * n + j] * DNOW[j * NC + c];
            const int o = k * NC + c;
            LD[o] = ld;
            LT[o] = -at;
            TS[o] = TOLD[o] + S.astep * (RT[o] - (-at));
        }
        if (kl == 0) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = acc + NUp[k] * DNOW[k * NC + c];
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
            double acc = PCp[k] * PC[4 * NC + c];
#pragma unroll 8
            for (int j = k; j < n; ++j) acc = acc + Gp[k * n + j] * TS[j * NC + c];
            const double ldstar = lam * acc;
            double w = 0.0;
#pragma unroll 8
            for (int j = 0; j < n; ++j) w = w + Wp[k * n + j] * DOLD[j * NC + c];
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
#pragma unroll 8
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
#pragma unroll 8
            for (int j = 0; j <= k; ++j) acc = acc + Tp[k * n + j] * DBAR[j * NC + c];
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
            for (int k = 0; k < n; ++k) acc = acc + NUp[k] * DBAR[k * NC + c];
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
            ov = vo[r];
            nv = ov + S.h * vt[r];
        }
        double nd = dn[r], nt = tn[r];
        if (diff) {
            nv = fv[r] * nv;
            nd = fd[r] * nd;
            nt = fv[r] * nt;
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
                cv = vc[r];
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
// order, FP64), then, if fix, ln p_s(0,0) += ln(target/mass) c00. One block per
// ensemble member (blockIdx.x): member b reads grid b (stride gs) and fixes spectral
// field b (stride ss) toward targets[b], or toward `target` when targets is null.
__global__ __launch_bounds__(256) void k_mass(const double *lnps_g, int nlon, int nlat, const double *w,
                                              double grav, int fix, double target, const double *targets,
                                              double c00, double *lnps, double *mass_out, double *lnr_out,
                                              long gs, long ss)
{
    const int b = blockIdx.x;
    lnps_g += (long)b * gs;
    lnps += (long)b * ss;
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
    if (mass_out) mass_out[b] = mass;
    if (fix) {
        const double lnr = log((targets ? targets[b] : target) / mass);
        lnps[0] = lnps[0] + lnr * c00;
        if (lnr_out) lnr_out[b] = lnr;
    }
}

}  // namespace aer

namespace aer {

// Adjoint of k_step (the transpose of the linear map (X^(n-1), X^n, R) -> (X^(n+1),
// filtered X^n) that one leapfrog step applies at fixed operators). No reference
// counterpart. Same block shape as k_step: 16 consecutive real entries of a level
// row, 16 level lanes; every stage is a transposed small matrix-vector product over
// LDS columns.
struct StepAdjArgs {
    int nlev, lmax, nlm, flags;
    double h, a1, astep, hh_old, half_nu, raw_a, raw_b;
    const double *G, *tau, *W, *nu, *pcoef, *Minv, *fvt, *fdiv;
    const int *l_of_lm;
    const double *bnvor, *bndiv, *bntemp, *bnlnps;   // adjoint of X^(n+1)
    const double *bcvor, *bcdiv, *bctemp, *bclnps;   // adjoint of the filtered X^n
    double *govor, *godiv, *gotemp, *golnps;         // adjoint of X^(n-1)
    double *gcvor, *gcdiv, *gctemp, *gclnps;         // adjoint of X^n (the input, before the filter)
    double *gtvor, *gtdiv, *gttemp, *gtlnps;         // adjoint of the tendency R(X^n)
    long ms3, ms2;                                   // member strides, as in StepArgs
};

__device__ __forceinline__ StepAdjArgs member_args(StepAdjArgs S)
{
    const long m3 = (long)blockIdx.y * S.ms3, m2 = (long)blockIdx.y * S.ms2;
    member_shift(S.bnvor, m3); member_shift(S.bndiv, m3); member_shift(S.bntemp, m3); member_shift(S.bnlnps, m2);
    member_shift(S.bcvor, m3); member_shift(S.bcdiv, m3); member_shift(S.bctemp, m3); member_shift(S.bclnps, m2);
    member_shift(S.govor, m3); member_shift(S.godiv, m3); member_shift(S.gotemp, m3); member_shift(S.golnps, m2);
    member_shift(S.gcvor, m3); member_shift(S.gcdiv, m3); member_shift(S.gctemp, m3); member_shift(S.gclnps, m2);
    member_shift(S.gtvor, m3); member_shift(S.gtdiv, m3); member_shift(S.gttemp, m3); member_shift(S.gtlnps, m2);
    return S;
}

__global__ __launch_bounds__(256) void k_step_adj(const StepAdjArgs S0)
{
    const StepAdjArgs S = member_args(S0);
    extern __shared__ double lds[];
    const int n = S.nlev;
    const int c = threadIdx.x % AER_SI_CB, kl = threadIdx.x / AER_SI_CB;
    const long q = (long)blockIdx.x * AER_SI_CB + c;
    const bool live = q < 2L * S.nlm;
    const int l = S.l_of_lm[live ? (int)(q >> 1) : 0];
    const long ls = 2L * S.nlm;
    const double lam = (double)(l * (l + 1)) / (6.371e6 * 6.371e6);
    constexpr int NR = AER_SI_MAXLEV / AER_SI_KL;
    const int NC = AER_SI_CB;
    double *TRAW = lds, *DNRT = TRAW + n * NC, *RHSB = DNRT + n * NC, *LDB = RHSB + n * NC, *LTB = LDB + n * NC;
    double *PC = LTB + n * NC;   // per column: praw, ps_b, lp_b
    const bool semi = S.flags & ST_SEMI, diff = S.flags & ST_DIFF, raw = S.flags & ST_RAW;
    // per own level: adjoint pieces kept in registers
    double od[NR], ot[NR], nd[NR], nt[NR], rd[NR], rt[NR], draw[NR];
    // 0. filter and diffusion transposed, vorticity finished
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_SI_KL;
        od[r] = ot[r] = nd[r] = nt[r] = rd[r] = rt[r] = draw[r] = 0.0;
        if (k >= n) break;
        const long x = k * ls + q;
        double bnd = live ? S.bndiv[x] : 0.0, bnt = live ? S.bntemp[x] : 0.0, bnv = live ? S.bnvor[x] : 0.0;
        double bcd = live ? S.bcdiv[x] : 0.0, bct = live ? S.bctemp[x] : 0.0, bcv = live ? S.bcvor[x] : 0.0;
        double bov = 0.0;
        if (raw) {
            double d = S.raw_a * bcd - S.raw_b * bnd;
            bcd -= 2.0 * S.half_nu * d;
            bnd += S.half_nu * d;
            od[r] += S.half_nu * d;
            d = S.raw_a * bct - S.raw_b * bnt;
            bct -= 2.0 * S.half_nu * d;
            bnt += S.half_nu * d;
            ot[r] += S.half_nu * d;
            d = S.raw_a * bcv - S.raw_b * bnv;
            bcv -= 2.0 * S.half_nu * d;
            bnv += S.half_nu * d;
            bov += S.half_nu * d;
        }
        if (diff) {
            const double fv = S.fvt[k * (S.lmax + 1) + l], fd = S.fdiv[k * (S.lmax + 1) + l];
            bnv *= fv;
            bnd *= fd;
            bnt *= fv;
        }
        nd[r] = bcd;
        nt[r] = bct;
        draw[r] = bnd;
        TRAW[k * NC + c] = bnt;
        if (live) {
            S.govor[x] = bov + bnv;
            S.gtvor[x] = S.h * bnv;
            S.gcvor[x] = bcv;
        }
    }
    double bp_o = 0.0, bp_c = 0.0;
    if (kl == 0) {
        double bnp = live ? S.bnlnps[q] : 0.0, bcp = live ? S.bclnps[q] : 0.0;
        if (raw) {
            const double d = S.raw_a * bcp - S.raw_b * bnp;
            bcp -= 2.0 * S.half_nu * d;
            bnp += S.half_nu * d;
            bp_o += S.half_nu * d;
        }
        PC[0 * NC + c] = bnp;   // adjoint of the unfiltered ln p_s^(n+1)
        bp_c = bcp;
    }
    __syncthreads();
    const double praw = PC[0 * NC + c];
    if (semi) {
        // 1. T^(n+1) = T^old + h (R_T - tau DBAR - L_T), P^(n+1) likewise: adjoint of DBAR
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            const int o = k * NC + c;
            const double tr = TRAW[o];
            ot[r] += tr;
            rt[r] += S.h * tr;
            double acc = 0.0;
#pragma unroll 8
            for (int kk = k; kk < n; ++kk) acc = acc + S.tau[kk * n + k] * TRAW[kk * NC + c];
            const double dbar = -S.h * acc - S.h * S.nu[k] * praw;
            DNRT[o] = draw[r] + S.a1 * dbar;
            od[r] += (1.0 - S.a1) * dbar;
            LTB[o] = -S.h * tr;
        }
        __syncthreads();
        // 2. D^(n+1) = Minv RHS: RHS adjoint = Minv^T (adjoint of D^(n+1))
        const double *Mi = S.Minv + (long)l * n * n;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            double acc = 0.0;
#pragma unroll 8
            for (int kk = 0; kk < n; ++kk) acc = acc + Mi[kk * n + k] * DNRT[kk * NC + c];
            RHSB[k * NC + c] = acc;
        }
        __syncthreads();
        // 3. RHS = D^old - hh_old lam W D^old + h (R_D + L_D* - L_D), L_D* = lam (pcoef P* + G T*)
        if (kl == 0) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = acc + S.pcoef[k] * RHSB[k * NC + c];
            PC[1 * NC + c] = lam * S.h * acc;   // adjoint of P*
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            const int o = k * NC + c;
            double w = 0.0;
#pragma unroll 8
            for (int kk = 0; kk < n; ++kk) w = w + S.W[kk * n + k] * RHSB[kk * NC + c];
            od[r] += RHSB[o] - S.hh_old * lam * w;
            rd[r] += S.h * RHSB[o];
            LDB[o] = -S.h * RHSB[o];
            // T* = T^old + astep (R_T - L_T): adjoint of T*
            double ts = 0.0;
#pragma unroll 8
            for (int kk = 0; kk <= k; ++kk) ts = ts + S.G[kk * n + k] * RHSB[kk * NC + c];
            ts *= lam * S.h;
            ot[r] += ts;
            rt[r] += S.astep * ts;
            LTB[o] += -S.astep * ts;
        }
        __syncthreads();
        if (kl == 0) {
            const double bps = PC[1 * NC + c];
            bp_o += bps + praw;
            PC[2 * NC + c] = -S.h * praw - S.astep * bps;   // adjoint of L_P
            PC[1 * NC + c] = S.h * praw + S.astep * bps;    // adjoint of R_P
        }
        __syncthreads();
        // 4. L_D = lam (pcoef P^n + G T^n), L_T = -tau D^n, L_P = -nu . D^n
        const double blp = PC[2 * NC + c];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            double tn = 0.0;
#pragma unroll 8
            for (int kk = 0; kk <= k; ++kk) tn = tn + S.G[kk * n + k] * LDB[kk * NC + c];
            nt[r] += lam * tn;
            double dn = 0.0;
#pragma unroll 8
            for (int kk = k; kk < n; ++kk) dn = dn + S.tau[kk * n + k] * LTB[kk * NC + c];
            nd[r] += -dn - S.nu[k] * blp;
        }
        if (kl == 0) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc = acc + S.pcoef[k] * LDB[k * NC + c];
            bp_c += lam * acc;
        }
    } else {
        // explicit leapfrog: D^(n+1) = D^old + h R_D, likewise T and ln p_s
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int k = kl + r * AER_SI_KL;
            if (k >= n) break;
            const double tr = TRAW[k * NC + c];
            od[r] += draw[r];
            rd[r] += S.h * draw[r];
            ot[r] += tr;
            rt[r] += S.h * tr;
        }
        if (kl == 0) {
            bp_o += praw;
            PC[1 * NC + c] = S.h * praw;
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_SI_KL;
        if (k >= n) break;
        const long x = k * ls + q;
        S.godiv[x] = od[r];
        S.gotemp[x] = ot[r];
        S.gcdiv[x] = nd[r];
        S.gctemp[x] = nt[r];
        S.gtdiv[x] = rd[r];
        S.gttemp[x] = rt[r];
    }
    if (kl == 0) {
        S.golnps[q] = bp_o;
        S.gclnps[q] = bp_c;
        S.gtlnps[q] = PC[1 * NC + c];
    }
}

}  // namespace aer
