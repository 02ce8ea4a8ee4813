// column_device.h -- gfx950 kernel of the grid-space column work between the two
// transform halves of a dynamics step: the Simmons-Burridge column terms
// (reference src/dynamics/aeros_tendency.f90:475-630) and, optionally, the
// Held-Suarez forcing (src/physics/aeros_held_suarez.f90:177-293).
//
// Block = 32 columns x 8 level lanes (256 threads). Thread (c, kl) owns levels
// kl, kl+8, ... of column c; a wave covers 32 consecutive columns at two levels,
// so every field load and store is a 256-byte contiguous run. The column's
// winds, temperature, mass divergence and hydrostatic terms go to LDS; two LDS
// scans (the cumulative mass divergence top-down, the hydrostatic integral
// bottom-up) run one thread per column; then every thread assembles its levels.
// Each scan adds in the reference's order, so the results follow the serial
// column loop operation for operation.
#pragma once
#include <hip/hip_runtime.h>

#define AER_COL_CB 32        // columns per block
#define AER_COL_KL 8         // level lanes per block
#define AER_COL_MAXLEV 64    // nlev limit (levels per thread <= 8)

namespace aer {

constexpr double kRd = 287.05;             // src/aeros_defs.f90:88
constexpr double kCpd = 1004.64;           // src/aeros_defs.f90:90
constexpr double kKappa = kRd / kCpd;      // src/aeros_defs.f90:92
constexpr double kOmega = 7.292e-5;        // src/aeros_defs.f90:83
constexpr double kMV = -9999.0;            // src/aeros_defs.f90:67
// Held & Suarez (1994) (src/physics/aeros_held_suarez.f90:92-101)
constexpr double kHsTeq0 = 315.0, kHsTmin = 200.0, kHsDTy = 60.0, kHsDThz = 10.0;
constexpr double kHsSigB = 0.7, kHsP0 = 1.0e5;
constexpr double kHsKa = 1.0 / (40.0 * 86400.0), kHsKs = 1.0 / (4.0 * 86400.0), kHsKf = 1.0 / 86400.0;

struct ColumnArgs {
    int nlon, ncol, nlev, hs, ztop;
    const double *A, *B;            // hybrid coefficients, nlev+1 half levels, top first
    const double *sinlat;           // per latitude row, north -> south
    const double *u, *v, *vor, *div, *t, *dtdx, *dtdy;   // [nlev][ncol]
    const double *lnps, *dlnpsdx, *dlnpsdy, *phis;       // [ncol]; phis may be null
    double *ae, *an, *ek, *dtdt;    // [nlev][ncol]
    double *dlnpsdt;                // [ncol]
    const double *dteq;             // optional additive Held-Suarez T_eq [nlev][nlat], or null
    long ms3, ms2, msq;             // member strides (doubles): 3-D fields, 2-D fields, dteq (0: shared)
};

// Ensemble members of one size share the grid and the vertical coordinate:
// blockIdx.y selects the member and its fields sit at fixed strides (phis is shared).
template <typename P>
__device__ __forceinline__ void col_shift(P *&p, long off)
{
    if (p) p += off;
}
__device__ __forceinline__ ColumnArgs member_args(ColumnArgs C)
{
    const long m = blockIdx.y, m3 = m * C.ms3, m2 = m * C.ms2;
    col_shift(C.u, m3); col_shift(C.v, m3); col_shift(C.vor, m3); col_shift(C.div, m3);
    col_shift(C.t, m3); col_shift(C.dtdx, m3); col_shift(C.dtdy, m3);
    col_shift(C.lnps, m2); col_shift(C.dlnpsdx, m2); col_shift(C.dlnpsdy, m2);
    col_shift(C.ae, m3); col_shift(C.an, m3); col_shift(C.ek, m3); col_shift(C.dtdt, m3);
    col_shift(C.dlnpsdt, m2);
    col_shift(C.dteq, m * C.msq);
    return C;
}

// Per-level column geometry: dp, alpha, dlnp (with the zero-top convention of
// src/dynamics/aeros_tendency.f90:514-519) and c_k.
struct LevelGeo {
    double ph0, ph1, dp, alpha, dlnp, cfac;
};
// ln(p_{k+1/2} / p_{k-1/2}) of a level with a positive upper half level (else unused)
__device__ __forceinline__ double level_lr(const double *sA, const double *sB, int k, double ps)
{
    const double ph0 = sA[k] + sB[k] * ps, ph1 = sA[k + 1] + sB[k + 1] * ps;
    return ph0 > 0.0 ? log(ph1 / ph0) : 0.0;
}
__device__ __forceinline__ LevelGeo level_geo_lr(const ColumnArgs &C, const double *sA, const double *sB,
                                                 int k, double ps, double lr)
{
    LevelGeo g;
    g.ph0 = sA[k] + sB[k] * ps;
    g.ph1 = sA[k + 1] + sB[k + 1] * ps;
    g.dp = g.ph1 - g.ph0;
    if (g.ph0 <= 0.0) {
        g.alpha = log(2.0);
        g.dlnp = (C.ztop && k == 0) ? 0.0 : kMV;
    } else {
        g.alpha = 1.0 - (g.ph0 / g.dp) * lr;
        g.dlnp = lr;
    }
    g.cfac = (ps / g.dp) * (sB[k] * g.dlnp + (sB[k + 1] - sB[k]) * g.alpha);
    return g;
}
__device__ __forceinline__ LevelGeo level_geo(const ColumnArgs &C, const double *sA, const double *sB,
                                              int k, double ps)
{
    return level_geo_lr(C, sA, sB, k, ps, level_lr(sA, sB, k, ps));
}

// Held-Suarez T_eq(phi, p) = max(200, (315 - dT_y sin^2 - dtheta_z ln(p/p0) cos^2) (p/p0)^kappa),
// with (p/p0)^kappa as exp(kappa ln(p/p0)) so one logarithm serves both factors.
__device__ __forceinline__ double hs_teq(double pf, double s2, double c2)
{
    const double lpf = log(pf / kHsP0);
    const double teq = (kHsTeq0 - kHsDTy * s2 - kHsDThz * lpf * c2) * exp(kKappa * lpf);
    return teq > kHsTmin ? teq : kHsTmin;
}

// NR: levels per thread (nlev <= NR * AER_COL_KL), chosen by the host so that the
// per-thread level arrays hold only the levels that exist.
template <int NR>
__global__ __launch_bounds__(256) void k_column_terms(const ColumnArgs C0)
{
    const ColumnArgs C = member_args(C0);
    extern __shared__ double lds[];
    const int nlev = C.nlev;
    const int cl = threadIdx.x % AER_COL_CB, kl = threadIdx.x / AER_COL_CB;
    const int col = blockIdx.x * AER_COL_CB + cl;
    const bool live = col < C.ncol;
    const int cc = live ? col : 0;
    const long ng = C.ncol;
    // LDS: A, B (nlev+1 each), then per level x column: U, V, T, Cm, AT, H, and S
    double *sA = lds, *sB = lds + (nlev + 1);
    double *U = sB + (nlev + 1), *V = U + nlev * AER_COL_CB, *Tt = V + nlev * AER_COL_CB;
    double *Cm = Tt + nlev * AER_COL_CB, *AT = Cm + nlev * AER_COL_CB, *H = AT + nlev * AER_COL_CB;
    double *S = H + nlev * AER_COL_CB;   // (nlev+1) x CB; reused for Phi after the scans
    double *PHI = AT;

    // 1. every global load first: column 2-D fields, own levels' fields
    const double lnps = live ? C.lnps[cc] : 0.0;
    const double gx = live ? C.dlnpsdx[cc] : 0.0, gy = live ? C.dlnpsdy[cc] : 0.0;
    const double phis = (live && C.phis) ? C.phis[cc] : 0.0;
    double ru[NR], rv[NR], rt[NR], rd[NR], rz[NR], rtx[NR], rty[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_COL_KL;
        const bool ok = live && k < nlev;
        const long x = ok ? (long)k * ng + cc : 0;
        ru[r] = ok ? C.u[x] : 0.0;
        rv[r] = ok ? C.v[x] : 0.0;
        rt[r] = ok ? C.t[x] : 0.0;
        rd[r] = ok ? C.div[x] : 0.0;
        rz[r] = ok ? C.vor[x] : 0.0;
        rtx[r] = ok ? C.dtdx[x] : 0.0;
        rty[r] = ok ? C.dtdy[x] : 0.0;
    }
    for (int j = threadIdx.x; j <= nlev; j += blockDim.x) {
        sA[j] = C.A[j];
        sB[j] = C.B[j];
    }
    __syncthreads();
    const double ps = exp(lnps);

    // 2. per-level terms to LDS: winds, temperature, layer mass divergence C_k,
    //    the hydrostatic layer-point term alpha R T and the interface increment
    double lr[NR];   // each own level's ln(p_{k+1/2}/p_{k-1/2}), kept for step 4
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_COL_KL;
        lr[r] = 0.0;
        if (k >= nlev) break;
        lr[r] = level_lr(sA, sB, k, ps);
        const LevelGeo g = level_geo_lr(C, sA, sB, k, ps, lr[r]);
        const double G = ru[r] * gx + rv[r] * gy;
        const int o = k * AER_COL_CB + cl;
        U[o] = ru[r];
        V[o] = rv[r];
        Tt[o] = rt[r];
        Cm[o] = g.dp * rd[r] + (sB[k + 1] - sB[k]) * ps * G;
        AT[o] = g.alpha * kRd * rt[r];
        H[o] = g.ph0 > 0.0 ? kRd * rt[r] * lr[r] : 0.0;
    }
    __syncthreads();

    // 3. scans, one thread per column: S top-down (wave 0), Phi bottom-up (wave 1)
    if (threadIdx.x < AER_COL_CB) {
        double s = 0.0;
        S[cl] = 0.0;
        for (int k = 0; k < nlev; ++k) {
            s = s + Cm[k * AER_COL_CB + cl];
            S[(k + 1) * AER_COL_CB + cl] = s;
        }
        if (live) C.dlnpsdt[col] = -s / ps;
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + AER_COL_CB) {
        double ph = phis;
        for (int k = nlev - 1; k >= 0; --k) {
            const int o = k * AER_COL_CB + cl;
            const double h = H[o];
            PHI[o] = ph + AT[o];
            ph = ph + h;
        }
    }
    __syncthreads();

    // 4. assemble each level
    const double SN = S[nlev * AER_COL_CB + cl];
    const double f = 2.0 * kOmega * C.sinlat[cc / C.nlon];
    const double s2 = C.sinlat[cc / C.nlon] * C.sinlat[cc / C.nlon], c2 = 1.0 - s2, c4 = c2 * c2;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int k = kl + r * AER_COL_KL;
        if (k >= nlev || !live) break;
        const LevelGeo g = level_geo_lr(C, sA, sB, k, ps, lr[r]);
        const int o = k * AER_COL_CB + cl;
        const double G = ru[r] * gx + rv[r] * gy;
        const double Sk = S[o], Ck = Cm[o];
        const double om = g.cfac * G - (Sk * g.dlnp + g.alpha * Ck) / g.dp;
        // mass fluxes at the level's upper (k) and lower (k+1) half levels
        const double mfu = sB[k] * SN - Sk;
        const double mfl = sB[k + 1] * SN - S[o + AER_COL_CB];
        const bool hasl = k < nlev - 1, hasu = k > 0;
        const double uu = U[o], vv = V[o], tt = Tt[o];
        const double ul = hasl ? U[o + AER_COL_CB] : 0.0, vl = hasl ? V[o + AER_COL_CB] : 0.0, tl = hasl ? Tt[o + AER_COL_CB] : 0.0;
        const double uo = hasu ? U[o - AER_COL_CB] : 0.0, vo = hasu ? V[o - AER_COL_CB] : 0.0, to = hasu ? Tt[o - AER_COL_CB] : 0.0;
        const double au = 0.5 * ((hasl ? mfl * (ul - uu) : 0.0) + (hasu ? mfu * (uu - uo) : 0.0)) / g.dp;
        const double av = 0.5 * ((hasl ? mfl * (vl - vv) : 0.0) + (hasu ? mfu * (vv - vo) : 0.0)) / g.dp;
        const double at = 0.5 * ((hasl ? mfl * (tl - tt) : 0.0) + (hasu ? mfu * (tt - to) : 0.0)) / g.dp;
        const double zf = rz[r] + f;
        const double pgf = kRd * tt * g.cfac;
        double ae = zf * vv - au - pgf * gx;
        double an = -zf * uu - av - pgf * gy;
        const double kin = 0.5 * (uu * uu + vv * vv);
        const double ek = PHI[o] + kin;
        double dtdt = -(uu * rtx[r] + vv * rty[r]) - at + kKappa * tt * om;
        if (C.hs) {
            const double pf = 0.5 * (g.ph1 + g.ph0);
            const double sig = pf / ps;
            double teq = hs_teq(pf, s2, c2);
            if (C.dteq) teq += C.dteq[(long)k * (C.ncol / C.nlon) + cc / C.nlon];
            double hf = (sig - kHsSigB) / (1.0 - kHsSigB);
            hf = hf > 0.0 ? hf : 0.0;
            const double kt = kHsKa + (kHsKs - kHsKa) * hf * c4;
            const double kv = kHsKf * hf;
            dtdt = dtdt - kt * (tt - teq);
            ae = ae - kv * uu;
            an = an - kv * vv;
        }
        const long x = (long)k * ng + col;
        C.ae[x] = ae;
        C.an[x] = an;
        C.ek[x] = ek;
        C.dtdt[x] = dtdt;
    }
}

}  // namespace aer
