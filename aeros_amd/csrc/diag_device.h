// diag_device.h -- gfx950 kernel of the online zonal-mean statistics (the reference's
// aeros_diag_accumulate, src/aeros_diagnostics.f90:153-233).
//
// One sample adds, for every (level, latitude) row, the instantaneous zonal means
// [u], [v], [T] and the covariances about them -- [u'v'], [v'T'], [(u'^2 + v'^2)/2],
// [T'^2], primes being departures from the INSTANTANEOUS zonal mean -- to FP64
// running sums; a last row of blocks adds [p_s] (p_s = exp(ln p_s)) and the lowest
// level's [u]. One wave per row: every lane holds its longitudes in registers, the
// two passes (means, then covariances) are wave reductions, and lane 0 adds the
// row's values to the sums, so a sample is deterministic (no atomics).
#pragma once
#include <hip/hip_runtime.h>

namespace aer {

constexpr int DG_MAXPL = 8;   // longitudes per lane (nlon <= 512)
constexpr int DG_NFIELD = 7;  // u, v, T, u'v', v'T', eke, T'^2

struct DiagArgs {
    int nlon, nlat, nlev;
    const double *u, *v, *temp, *lnps;   // [nlev][nlat][nlon], lnps [nlat][nlon]
    double *sum;                         // [7][nlev][nlat], then [2][nlat]: p_s, lowest-level u
};

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__global__ __launch_bounds__(64) void k_diag_accumulate(const DiagArgs A)
{
    const int j = blockIdx.x, k = blockIdx.y, lane = threadIdx.x;
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev;
    const double rn = 1.0 / (double)nlon;
    if (k == nlev) {
        // single-level rows: zonal-mean surface pressure and lowest-level [u]
        const double *lp = A.lnps + (long)j * nlon;
        const double *us = A.u + ((long)(nlev - 1) * nlat + j) * nlon;
        double sp = 0.0, su = 0.0;
        for (int i = lane; i < nlon; i += 64) {
            sp += exp(lp[i]);
            su += us[i];
        }
        sp = wave_sum(sp);
        su = wave_sum(su);
        if (lane == 0) {
            double *s2 = A.sum + (long)DG_NFIELD * nlev * nlat;
            s2[j] += sp * rn;
            s2[nlat + j] += su * rn;
        }
        return;
    }
    const long row = ((long)k * nlat + j) * nlon;
    double u[DG_MAXPL], v[DG_MAXPL], t[DG_MAXPL];
    double su = 0.0, sv = 0.0, st = 0.0;
#pragma unroll
    for (int q = 0; q < DG_MAXPL; ++q) {
        const int i = lane + 64 * q;
        const bool ok = i < nlon;
        u[q] = ok ? A.u[row + i] : 0.0;
        v[q] = ok ? A.v[row + i] : 0.0;
        t[q] = ok ? A.temp[row + i] : 0.0;
        su += u[q];
        sv += v[q];
        st += t[q];
    }
    const double um = wave_sum(su) * rn, vm = wave_sum(sv) * rn, tm = wave_sum(st) * rn;
    double suv = 0.0, svt = 0.0, sek = 0.0, stt = 0.0;
#pragma unroll
    for (int q = 0; q < DG_MAXPL; ++q) {
        if (lane + 64 * q >= nlon) continue;
        const double up = u[q] - um, vp = v[q] - vm, tp = t[q] - tm;
        suv += up * vp;
        svt += vp * tp;
        sek += 0.5 * (up * up + vp * vp);
        stt += tp * tp;
    }
    suv = wave_sum(suv);
    svt = wave_sum(svt);
    sek = wave_sum(sek);
    stt = wave_sum(stt);
    if (lane == 0) {
        const long o = (long)k * nlat + j, f = (long)nlev * nlat;
        A.sum[o] += um;
        A.sum[f + o] += vm;
        A.sum[2 * f + o] += tm;
        A.sum[3 * f + o] += suv * rn;
        A.sum[4 * f + o] += svt * rn;
        A.sum[5 * f + o] += sek * rn;
        A.sum[6 * f + o] += stt * rn;
    }
}

// Clears the sums (a fill kernel, so a captured graph never records a memset).
__global__ __launch_bounds__(256) void k_diag_zero(double *sum, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) sum[i] = 0.0;
}

}  // namespace aer
