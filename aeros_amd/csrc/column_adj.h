// column_adj.h -- gfx950 kernel of the adjoint (transpose of the linearization) of the
// grid-space column work k_column_terms computes: the Simmons-Burridge column terms
// (reference src/dynamics/aeros_tendency.f90:475-630) and the Held-Suarez forcing
// (src/physics/aeros_held_suarez.f90:177-293). The reference has no adjoint; this is
// the reverse sweep of the forward column, written out by hand.
//
// One thread per column, levels in sequence: the thread recomputes the column's
// forward intermediates (geometry, mass divergence scan, hydrostatic integral) from
// the forward inputs, then runs the reverse of every assignment, so the adjoint is
// evaluated at exactly the state the forward saw. Level arrays live in the thread's
// private memory (at most AER_ADJ_MAXLEV levels).
//
// Geometry depends on p_s through the hybrid coefficients (p = A + B p_s): dp, alpha,
// dlnp and c_k are differentiated in closed form (their derivatives vanish for pure
// sigma layers below a positive top). The Held-Suarez T_eq enters as formula + dteq
// (an optional additive field per level and latitude), so the kernel also returns the
// sensitivity to dteq per column.
#pragma once
#include "column_device.h"

#define AER_ADJ_MAXLEV 64

namespace aer {

struct ColumnAdjArgs {
    ColumnArgs f;                                    // forward inputs (outputs unused)
    const double *dteq;                              // [nlev][nlat] additive T_eq, or null
    const double *bae, *ban, *bek, *bdtdt;           // output adjoints [nlev][ncol]
    const double *bdlnpsdt;                          // [ncol]
    double *gu, *gv, *gvor, *gdiv, *gt, *gdtdx, *gdtdy;   // input adjoints [nlev][ncol] (overwritten)
    double *glnps, *gdlnpsdx, *gdlnpsdy;             // [ncol]
    double *gteq;                                    // [nlev][ncol]: d/d(dteq) per column, or null
};

__device__ __forceinline__ ColumnAdjArgs member_args(ColumnAdjArgs X)
{
    const long m = blockIdx.y, m3 = m * X.f.ms3, m2 = m * X.f.ms2;
    col_shift(X.dteq, m * X.f.msq);
    X.f = member_args(X.f);
    col_shift(X.bae, m3); col_shift(X.ban, m3); col_shift(X.bek, m3); col_shift(X.bdtdt, m3);
    col_shift(X.bdlnpsdt, m2);
    col_shift(X.gu, m3); col_shift(X.gv, m3); col_shift(X.gvor, m3); col_shift(X.gdiv, m3);
    col_shift(X.gt, m3); col_shift(X.gdtdx, m3); col_shift(X.gdtdy, m3);
    col_shift(X.glnps, m2); col_shift(X.gdlnpsdx, m2); col_shift(X.gdlnpsdy, m2);
    col_shift(X.gteq, m3);
    return X;
}

__global__ __launch_bounds__(64) void k_column_terms_adj(const ColumnAdjArgs X0)
{
    const ColumnAdjArgs X = member_args(X0);
    const ColumnArgs &C = X.f;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= C.ncol) return;
    const int n = C.nlev;
    const long ng = C.ncol;
    const double *A = C.A, *B = C.B;
    const int lat = col / C.nlon;
    const double sl = C.sinlat[lat], s2 = sl * sl, c2 = 1.0 - s2, c4 = c2 * c2;
    const double f = 2.0 * kOmega * sl;
    const double lnps = C.lnps[col], gx = C.dlnpsdx[col], gy = C.dlnpsdy[col];
    const double ps = exp(lnps);

    double u[AER_ADJ_MAXLEV], v[AER_ADJ_MAXLEV], t[AER_ADJ_MAXLEV];
    double dp[AER_ADJ_MAXLEV], al[AER_ADJ_MAXLEV], dl[AER_ADJ_MAXLEV], cf[AER_ADJ_MAXLEV];
    double G[AER_ADJ_MAXLEV], Cm[AER_ADJ_MAXLEV], S[AER_ADJ_MAXLEV + 1];
    // adjoint accumulators
    double bu[AER_ADJ_MAXLEV], bv[AER_ADJ_MAXLEV], bt[AER_ADJ_MAXLEV];
    double bdp[AER_ADJ_MAXLEV], bal[AER_ADJ_MAXLEV], bdl[AER_ADJ_MAXLEV], bcf[AER_ADJ_MAXLEV];
    double bG[AER_ADJ_MAXLEV], bCm[AER_ADJ_MAXLEV], bS[AER_ADJ_MAXLEV + 1], bPHI[AER_ADJ_MAXLEV];

    // ---- forward recomputation
    for (int k = 0; k < n; ++k) {
        const long x = (long)k * ng + col;
        u[k] = C.u[x];
        v[k] = C.v[x];
        t[k] = C.t[x];
        const LevelGeo g = level_geo(C, A, B, k, ps);
        dp[k] = g.dp;
        al[k] = g.alpha;
        dl[k] = g.dlnp;
        cf[k] = g.cfac;
        G[k] = u[k] * gx + v[k] * gy;
        Cm[k] = g.dp * C.div[x] + (B[k + 1] - B[k]) * ps * G[k];
        bu[k] = bv[k] = bt[k] = 0.0;
        bdp[k] = bal[k] = bdl[k] = bcf[k] = 0.0;
        bG[k] = bCm[k] = bPHI[k] = 0.0;
    }
    S[0] = 0.0;
    for (int k = 0; k < n; ++k) S[k + 1] = S[k] + Cm[k];
    for (int k = 0; k <= n; ++k) bS[k] = 0.0;
    const double SN = S[n];
    double bSN = 0.0, bps = 0.0, bgx = 0.0, bgy = 0.0;

    // ---- reverse of the per-level assembly
    for (int k = 0; k < n; ++k) {
        const long x = (long)k * ng + col;
        const double ae_b = X.bae[x], an_b = X.ban[x], ek_b = X.bek[x], dt_b = X.bdtdt[x];
        const double uu = u[k], vv = v[k], tt = t[k];
        const bool hasl = k < n - 1, hasu = k > 0;
        const double zf = C.vor[x] + f;
        const double pgf = kRd * tt * cf[k];
        const double om = cf[k] * G[k] - (S[k] * dl[k] + al[k] * Cm[k]) / dp[k];
        double kt = 0.0, kv = 0.0, teq = 0.0;
        double dkt_dps = 0.0, dkv_dps = 0.0, dteq_dps = 0.0;
        if (C.hs) {
            const double ph0 = A[k] + B[k] * ps, ph1 = A[k + 1] + B[k + 1] * ps;
            const double pf = 0.5 * (ph1 + ph0), dpf = 0.5 * (B[k] + B[k + 1]);
            const double sig = pf / ps, dsig = (dpf * ps - pf) / (ps * ps);
            const double lpf = log(pf / kHsP0);   // as hs_teq in the forward kernel
            const double base = kHsTeq0 - kHsDTy * s2 - kHsDThz * lpf * c2;
            const double pk = exp(kKappa * lpf);
            const double teqr = base * pk;
            teq = teqr > kHsTmin ? teqr : kHsTmin;
            if (teqr > kHsTmin) dteq_dps = (-kHsDThz * c2 / pf * pk + base * kKappa * pk / pf) * dpf;
            if (X.dteq) teq += X.dteq[(long)k * (C.ncol / C.nlon) + lat];
            const double hr = (sig - kHsSigB) / (1.0 - kHsSigB);
            const double hf = hr > 0.0 ? hr : 0.0;
            const double dhf = hr > 0.0 ? dsig / (1.0 - kHsSigB) : 0.0;
            kt = kHsKa + (kHsKs - kHsKa) * hf * c4;
            kv = kHsKf * hf;
            dkt_dps = (kHsKs - kHsKa) * c4 * dhf;
            dkv_dps = kHsKf * dhf;
        }
        // ae = zf v - au - pgf gx - kv u ; an = -zf u - av - pgf gy - kv v
        double bzf = ae_b * vv - an_b * uu;
        bv[k] += ae_b * zf - an_b * kv;
        bu[k] += -an_b * zf - ae_b * kv;
        const double bau = -ae_b, bav = -an_b;
        double bpgf = -ae_b * gx - an_b * gy;
        bgx += -ae_b * pgf;
        bgy += -an_b * pgf;
        const double bkv = -ae_b * uu - an_b * vv;
        // ek = PHI + (u^2 + v^2)/2
        bPHI[k] += ek_b;
        bu[k] += ek_b * uu;
        bv[k] += ek_b * vv;
        // dtdt = -(u Tx + v Ty) - at + kappa T om - kt (T - teq)
        const long xx = x;
        X.gdtdx[xx] = -dt_b * uu;
        X.gdtdy[xx] = -dt_b * vv;
        bu[k] += -dt_b * C.dtdx[x];
        bv[k] += -dt_b * C.dtdy[x];
        const double bat = -dt_b;
        bt[k] += dt_b * (kKappa * om - kt);
        const double bom = dt_b * kKappa * tt;
        const double bkt = -dt_b * (tt - teq);
        if (X.gteq) X.gteq[xx] = C.hs ? dt_b * kt : 0.0;
        if (C.hs) bps += bkt * dkt_dps + bkv * dkv_dps + dt_b * kt * dteq_dps;
        // the vorticity enters only through zf
        X.gvor[xx] = bzf;
        // pgf = Rd T cfac
        bt[k] += bpgf * kRd * cf[k];
        bcf[k] += bpgf * kRd * tt;
        // vertical advection of u, v, T: x = 0.5 (hasl mfl (y_{k+1} - y_k) + hasu mfu (y_k - y_{k-1})) / dp
        const double mfu = B[k] * SN - S[k];
        const double mfl = B[k + 1] * SN - S[k + 1];
        double bmfu = 0.0, bmfl = 0.0;
        {
            const double qu = 0.5 * bau / dp[k], qv = 0.5 * bav / dp[k], qt = 0.5 * bat / dp[k];
            double num = 0.0;
            if (hasl) {
                const double du = u[k + 1] - uu, dv = v[k + 1] - vv, dtt = t[k + 1] - tt;
                bmfl += qu * du + qv * dv + qt * dtt;
                bu[k + 1] += qu * mfl;
                bu[k] -= qu * mfl;
                bv[k + 1] += qv * mfl;
                bv[k] -= qv * mfl;
                bt[k + 1] += qt * mfl;
                bt[k] -= qt * mfl;
                num += mfl * (bau * du + bav * dv + bat * dtt);
            }
            if (hasu) {
                const double du = uu - u[k - 1], dv = vv - v[k - 1], dtt = tt - t[k - 1];
                bmfu += qu * du + qv * dv + qt * dtt;
                bu[k] += qu * mfu;
                bu[k - 1] -= qu * mfu;
                bv[k] += qv * mfu;
                bv[k - 1] -= qv * mfu;
                bt[k] += qt * mfu;
                bt[k - 1] -= qt * mfu;
                num += mfu * (bau * du + bav * dv + bat * dtt);
            }
            bdp[k] += -0.5 * num / (dp[k] * dp[k]);
        }
        // om = cfac G - (S_k dlnp + alpha Cm) / dp
        bcf[k] += bom * G[k];
        bG[k] += bom * cf[k];
        bS[k] += -bom * dl[k] / dp[k];
        bdl[k] += -bom * S[k] / dp[k];
        bal[k] += -bom * Cm[k] / dp[k];
        bCm[k] += -bom * al[k] / dp[k];
        bdp[k] += bom * (S[k] * dl[k] + al[k] * Cm[k]) / (dp[k] * dp[k]);
        // mfu = B_k SN - S_k, mfl = B_{k+1} SN - S_{k+1}
        bSN += bmfu * B[k] + bmfl * B[k + 1];
        bS[k] -= bmfu;
        bS[k + 1] -= bmfl;
    }
    // ---- hydrostatic integral: PHI_k = phis + AT_k + sum_{j>k} H_j
    {
        double above = 0.0;   // sum of bPHI over levels above j
        for (int j = 0; j < n; ++j) {
            // H_j feeds PHI_k for every k < j
            const double bH = above;
            above += bPHI[j];
            const double ph0 = A[j] + B[j] * ps;
            if (ph0 > 0.0) {
                const double lr = log((A[j + 1] + B[j + 1] * ps) / ph0);
                bt[j] += bH * kRd * lr;
                bdl[j] += bH * kRd * t[j];   // lr is dlnp in the regular branch
            }
            // AT_j = alpha_j Rd T_j
            bal[j] += bPHI[j] * kRd * t[j];
            bt[j] += bPHI[j] * al[j] * kRd;
        }
    }
    // ---- dlnpsdt = -SN / ps
    const double bdlt = X.bdlnpsdt[col];
    bSN += -bdlt / ps;
    bps += bdlt * SN / (ps * ps);
    bS[n] += bSN;
    // ---- scan: S_{k+1} = S_k + Cm_k
    for (int k = n - 1; k >= 0; --k) {
        bCm[k] += bS[k + 1];
        bS[k] += bS[k + 1];
    }
    // ---- Cm = dp D + dB ps G ; G = u gx + v gy ; geometry(ps)
    for (int k = 0; k < n; ++k) {
        const long x = (long)k * ng + col;
        const double dB = B[k + 1] - B[k];
        bdp[k] += bCm[k] * C.div[x];
        X.gdiv[x] = bCm[k] * dp[k];
        bps += bCm[k] * dB * G[k];
        bG[k] += bCm[k] * dB * ps;
        bu[k] += bG[k] * gx;
        bv[k] += bG[k] * gy;
        bgx += bG[k] * u[k];
        bgy += bG[k] * v[k];
        // d/dps of dp, alpha, dlnp, cfac
        const double ph0 = A[k] + B[k] * ps, ph1 = A[k + 1] + B[k + 1] * ps;
        double dal = 0.0, ddl = 0.0;
        if (ph0 > 0.0) {
            const double lr = log(ph1 / ph0);
            const double dlr = B[k + 1] / ph1 - B[k] / ph0;
            dal = -((B[k] * dp[k] - ph0 * dB) / (dp[k] * dp[k]) * lr + (ph0 / dp[k]) * dlr);
            ddl = dlr;
        }
        const double Q = B[k] * dl[k] + dB * al[k];
        const double dQ = B[k] * ddl + dB * dal;
        const double dcf = (dp[k] - ps * dB) / (dp[k] * dp[k]) * Q + (ps / dp[k]) * dQ;
        bps += bdp[k] * dB + bal[k] * dal + bdl[k] * ddl + bcf[k] * dcf;
        X.gu[x] = bu[k];
        X.gv[x] = bv[k];
        X.gt[x] = bt[k];
    }
    X.glnps[col] = bps * ps;
    X.gdlnpsdx[col] = bgx;
    X.gdlnpsdy[col] = bgy;
}

}  // namespace aer
