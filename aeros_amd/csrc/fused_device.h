// fused_device.h -- gfx950 kernels of the tiled transform path.
//
// Synthesis is ONE launch. k_syn_fused takes a tile of 4 grid fields x 4 latitude
// pairs, stages the tile's coefficients and table slice in LDS with coalesced loads,
// forms the per-m Legendre sums on the 4x4 MFMA forms (FP32: v_mfma_f32_4x4x1_16b,
// FP64: v_mfma_f64_4x4x4_4b), leaves the Fourier coefficients in LDS and runs the
// longitude FFT of the tile there, so nothing between the two stages goes to HBM.
//
// Analysis is two launches, because its Legendre stage sums over all latitudes:
//   k_fft_anal4  tile of 4 grid fields x 4 latitude pairs: FFT, then the even/odd
//                Fourier coefficients to HBM as [grid field][m][pair][E re, E im,
//                O re, O im] (one 16- or 32-byte vector per (field, m, pair));
//   k_leg_anal4  block = (column group, the orders m and T-m): every latitude
//                pair, reading each Fourier coefficient once; the pairing of m
//                with T-m gives every block T+2 degrees of work.
//
// Lane layouts of the 4x4 MFMA forms (measured on gfx950):
//   f32 4x4x1_16b: 16 blocks b = L/4; A[b][i] at lane 4b+i, B[b][j] at lane 4b+j,
//                  D[b][i][j] in lane 4b+j, accumulator element i.
//   f64 4x4x4_4b:  4 blocks q; A[q][i][k] at lane 16k+4q+i, B[q][k][j] at lane
//                  16k+4q+j, D[q][i][j] in lane 16i+4q+j.
#pragma once
#include "sht_device.h"

namespace aer {

constexpr int FZ_THREADS = 256;  // synthesis and analysis-FFT blocks (4 waves)
constexpr int LA_THREADS = 256;  // analysis Legendre blocks (4 waves)
constexpr int FZ_ROWS = 16;      // complex FFT rows of a tile: 4 grid fields x 4 pairs (N + i S)
constexpr int FZ_UNIT = 8;       // MFMA steps whose operands are loaded together

__device__ __forceinline__ f4 mfma_f32_4x4(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ double mfma_f64_4x4(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// Blocks go to the 8 XCDs round-robin. With a grid of 8*per blocks, task
// (bid % 8) * per + bid / 8 gives each XCD a contiguous task range, so tasks
// that read the same data (the pair tiles of one field group, the column groups
// of one order m) share one L2.
__device__ __forceinline__ int xcd_task(int bid, int nblk)
{
    const int per = nblk >> 3;
    return (bid & 7) * per + (bid >> 3);
}

__host__ __device__ constexpr int fz_align(int b) { return (b + 255) / 256 * 256; }
// LDS of k_syn_fused: F[m][pair 4][column 8][E|O], then the staging area of the Legendre
// stage, reused afterwards for the two FFT buffers and the twiddles.
__host__ __device__ constexpr int fz_fbytes(int T, int csz) { return fz_align((T + 1) * 64 * csz); }
__host__ __device__ constexpr int fz_fftbytes(int N, int csz)
{
    return (2 * (FZ_ROWS * N + (FZ_ROWS * N) / 16 + 1) + N) * 2 * csz;
}
// Fourier intermediate of k_fft_anal4 per job: ngf rounded to 4 x (T+1) x pairs rounded to 4 x 4 values.
__host__ __device__ constexpr long fz_four_elems(int ngf, int T, int nh)
{
    return (long)((ngf + 3) / 4 * 4) * (T + 1) * ((nh + 3) / 4 * 4) * 4;
}
// LDS of k_leg_anal4: the block's grid fields x 2 orders x pairs x 4 values.
__host__ __device__ constexpr int la_lds(int nlgf, int nh, int csz) { return nlgf * 2 * ((nh + 3) / 4 * 4) * 4 * csz; }

// ------------------------------------------------------------------ fused synthesis
//
// A block stages, for a band of orders [m0, m1) at a time, the coefficients of its
// 4 segments (scaled by the job's per-degree factor and rounded to the compute
// type) and the rows of its 4-pair table slice (tables laid out [tile][lm][4 pairs]
// so the slice is contiguous). All loads of a band are coalesced and in flight
// together; the MFMA operands then come from LDS.
//
// Segments: scalar jobs the 4 fields of the tile; vector and gradient jobs S of
// items 0, 1 and T of items 0, 1 (zero for gradients).

// Staging area: coefficients [4][chl] complex, then two table slices [chl][4], then
// FZ_PAD rows of slack that the unmasked operand reads of a band's last steps may touch.
constexpr int FZ_PAD = 96;
__host__ __device__ constexpr int fz_stage_bytes(int chl, int csz)
{
    return fz_align(4 * chl * 2 * csz + 2 * chl * 4 * csz + FZ_PAD * 4 * csz);
}

constexpr int FZ_SU = 16;  // coefficient loads per lane per staging round (a T42 segment in one round)
constexpr int FZ_TU = 4;   // table rows per thread per staging round (256 x 4 = 64 x FZ_SU)

template <class T> struct V4;
template <> struct V4<float> {
    float4 v;
    __device__ static V4 load(const float *p) { return V4{*(const float4 *)p}; }
    __device__ void store(float *p) const { *(float4 *)p = v; }
    __device__ static V4 zero() { return V4{make_float4(0.f, 0.f, 0.f, 0.f)}; }
};
template <> struct V4<double> {
    double2 a, b;
    __device__ static V4 load(const double *p) { return V4{((const double2 *)p)[0], ((const double2 *)p)[1]}; }
    __device__ void store(double *p) const { ((double2 *)p)[0] = a; ((double2 *)p)[1] = b; }
    __device__ static V4 zero() { return V4{make_double2(0.0, 0.0), make_double2(0.0, 0.0)}; }
};

template <int PREC>
__device__ __forceinline__ void syn_stage(const Job &J, int g, int tile, int lb0, int nlmc, int chl, char *S, int wave,
                                          int lane)
{
    using T = typename Prec<PREC>::T;
    using T2 = cx<T>;
    const CfgDev &C = J.c;
    T2 *coef = (T2 *)S;
    T *tabs = (T *)(coef + 4 * chl);
    // coefficients: wave w stages segment w; the per-degree factor is read only
    // for modes other than 1 (the mode is uniform over the wave)
    const double *src = nullptr;
    int mode = M_ONE;
    if (J.kind == SYN_SCALAR) {
        const int f = g * 4 + wave;
        if (f < J.nitem) { src = J.sin0 + (long)f * J.sstride; mode = J.mode0; }
    } else {
        const int it = g * 2 + (wave & 1);
        if (it < J.nitem) {
            if (wave < 2) { src = J.sin0 + (long)it * J.sstride; mode = J.mode0; }
            else if (J.kind == SYN_VECTOR) { src = J.sin1 + (long)it * J.sstride; mode = J.mode1; }
        }
    }
    const bool scalar = J.kind == SYN_SCALAR;
    const T *tq = (const T *)(scalar ? C.tabZP : C.tabZQ) + ((long)tile * C.nlm + lb0) * 4;
    const T *td = (const T *)C.tabZD + ((long)tile * C.nlm + lb0) * 4;
    const double2 *cs = (const double2 *)src + lb0;
    const double *fs = C.lfacm + (long)mode * C.nlm + lb0;
    T2 *cd = coef + wave * chl;
    // one pass per FZ_SU * 64 coefficients: table rows of the same range ride along
    // (their stride is the block's 256 threads), every load issued before the first wait
    for (int j0 = 0; j0 < nlmc; j0 += 64 * FZ_SU) {
        double2 c[FZ_SU];
        double f[FZ_SU];
        V4<T> vq[FZ_TU], vd[FZ_TU];
        if (src != nullptr) {
#pragma unroll
            for (int u = 0; u < FZ_SU; ++u) {
                const int j = min(j0 + lane + 64 * u, nlmc - 1);
                c[u] = cs[j];
            }
            if (mode != M_ONE) {
#pragma unroll
                for (int u = 0; u < FZ_SU; ++u) f[u] = fs[min(j0 + lane + 64 * u, nlmc - 1)];
            } else {
#pragma unroll
                for (int u = 0; u < FZ_SU; ++u) f[u] = 1.0;
            }
        } else {
#pragma unroll
            for (int u = 0; u < FZ_SU; ++u) {
                c[u] = make_double2(0.0, 0.0);
                f[u] = 0.0;
            }
        }
        const int r0 = j0 + threadIdx.x;   // table rows r0 + 256 u, u < FZ_TU, cover [j0, j0 + 64 FZ_SU)
#pragma unroll
        for (int u = 0; u < FZ_TU; ++u) {
            const int r = min(r0 + FZ_THREADS * u, nlmc - 1);
            vq[u] = V4<T>::load(tq + (long)r * 4);
            if (!scalar) vd[u] = V4<T>::load(td + (long)r * 4);
        }
#pragma unroll
        for (int u = 0; u < FZ_SU; ++u) {
            const int j = j0 + lane + 64 * u;
            if (j < nlmc) cd[j] = T2{T(c[u].x * f[u]), T(c[u].y * f[u])};
        }
#pragma unroll
        for (int u = 0; u < FZ_TU; ++u) {
            const int r = r0 + FZ_THREADS * u;
            if (r < nlmc) {
                vq[u].store(tabs + (long)r * 4);
                if (!scalar) vd[u].store(tabs + ((long)chl + r) * 4);
            }
        }
    }
}

__device__ __forceinline__ int snake_wave(int k) { return ((k >> 2) & 1) ? 3 - (k & 3) : (k & 3); }

// Legendre sums of one staged band of orders [m0, m1), written to F[m][pair][col][E|O].
//
// Every MFMA block accumulates one output parity directly: its P (scalar) or
// m P/sin(theta) (vector) operand at the degrees n of that parity, and, for vector
// jobs, its dP/dtheta operand at the degrees of the other parity (dP/dtheta has the
// opposite parity, so those terms belong to this output). No cross-lane sums.
//   FP32 (4x4x1, 16 blocks): block b = L/4 = cq + 2 par + 4 ms covers column quad cq,
//        parity par and order m0' + ms of a group of 4 consecutive orders; one degree
//        per block per step (n = 2t + par).
//   FP64 (4x4x4, 4 blocks): block q = cq + 2 par, K = L/16; one order at a time, four
//        degrees per block per step (n = par + 2(4t + K)).
template <int PREC>
__device__ __forceinline__ void syn_band(const Job &J, int m0, int m1, int lb0, int chl, const char *S, int wave,
                                         int lane, void *Fv)
{
    using T = typename Prec<PREC>::T;
    const int Tr = J.c.T;
    const bool scalar = J.kind == SYN_SCALAR;
    const T *coef = (const T *)S;                          // [seg][chl][re|im]
    const T *tabQ = coef + 4 * chl * 2;                    // [chl][4]: P (scalar) or m P/sin (vector)
    const T *tabD = tabQ + chl * 4;                        // [chl][4]: dP/dtheta (vector)
    T *F = (T *)Fv;
    const int j = lane & 3, cq = (lane >> 2) & 1, par = (lane >> 3) & 1;
    const int part = j & 1, comp = j >> 1;
    // B operand segments and signs. Scalar: field 2cq + j/2. Vector: V_theta = S dP + (iT) Q,
    // V_phi = (iS) Q - T dP, with Re(iz) = -Im z, Im(iz) = +Re z.
    const int segQ = scalar ? 2 * cq + comp : (comp == 0 ? 2 + cq : cq);
    const int pQ = scalar ? part : 1 - part;
    const T sQ = scalar ? T(1) : (part ? T(1) : T(-1));
    const int segD = comp == 0 ? cq : 2 + cq;
    const T sD = comp == 0 ? T(1) : T(-1);
    // Operand addresses are affine in the step: per step the Q operand's degree moves by
    // DN (FP32 2, FP64 8) and so does the dP operand's. Reads past a lane's last degree
    // stay inside the staging area (it is padded) and are replaced by zero.
    constexpr int DN = PREC == 1 ? 2 : 8;
    if constexpr (PREC == 1) {
        const int ms = lane >> 4;
        const int ngroup = (m1 - m0 + 3) >> 2;
        for (int gi = 0; gi < ngroup; ++gi) {
            if (snake_wave(gi) != wave) continue;
            const int mg = m0 + 4 * gi;
            const int m = mg + ms;
            const bool mv = m < m1;
            const int nl = mv ? Tr - m + 1 : 0;
            const int jb = mv ? lm_base(Tr, m) - lb0 : 0;
            const int steps = (Tr - mg + 2) / 2;
            // steps valid for every lane of the group: degree 2t + 1 below the smallest nl
            const int nlmin = Tr - min(mg + 3, m1 - 1) + 1;
            const int full = nlmin / 2;
            const T *aq = tabQ + (jb + par) * 4 + j, *bq = coef + (segQ * chl + jb + par) * 2 + pQ;
            const T *ad = tabD + (jb + 1 - par) * 4 + j, *bd = coef + (segD * chl + jb + 1 - par) * 2 + part;
            f4 acc = {0.f, 0.f, 0.f, 0.f}, acc2 = {0.f, 0.f, 0.f, 0.f};   // two chains: Q and dP terms
            int t0 = 0;
            for (; t0 + FZ_UNIT <= full; t0 += FZ_UNIT) {
                T a[FZ_UNIT], b[FZ_UNIT], a2[FZ_UNIT], b2[FZ_UNIT];
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    a[u] = aq[(t0 + u) * DN * 4];
                    b[u] = sQ * bq[(t0 + u) * DN * 2];
                    if (!scalar) {
                        a2[u] = ad[(t0 + u) * DN * 4];
                        b2[u] = sD * bd[(t0 + u) * DN * 2];
                    }
                }
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    acc = mfma_f32_4x4(a[u], b[u], acc);
                    if (!scalar) acc2 = mfma_f32_4x4(a2[u], b2[u], acc2);
                }
            }
            for (; t0 < steps; t0 += FZ_UNIT) {
                T a[FZ_UNIT], b[FZ_UNIT], a2[FZ_UNIT], b2[FZ_UNIT];
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    const int n = 2 * (t0 + u);
                    const bool vq = n + par < nl, vd = n + 1 - par < nl;
                    const T xa = aq[(t0 + u) * DN * 4], xb = bq[(t0 + u) * DN * 2];
                    a[u] = vq ? xa : T(0);
                    b[u] = vq ? sQ * xb : T(0);
                    if (!scalar) {
                        const T ya = ad[(t0 + u) * DN * 4], yb = bd[(t0 + u) * DN * 2];
                        a2[u] = vd ? ya : T(0);
                        b2[u] = vd ? sD * yb : T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    if (t0 + u >= steps) break;
                    acc = mfma_f32_4x4(a[u], b[u], acc);
                    if (!scalar) acc2 = mfma_f32_4x4(a2[u], b2[u], acc2);
                }
            }
            if (mv) {
                const int c = 4 * cq + j;
#pragma unroll
                for (int i = 0; i < 4; ++i) F[((m * 4 + i) * 8 + c) * 2 + par] = scalar ? acc[i] : acc[i] + acc2[i];
            }
        }
    } else {
        const int K = lane >> 4;
        for (int m = m0; m < m1; ++m) {
            if (snake_wave(m - m0) != wave) continue;
            const int nl = Tr - m + 1;
            const int jb = lm_base(Tr, m) - lb0;
            const int steps = ((nl + 1) / 2 + 3) / 4;
            const int full = nl / 8;          // steps whose largest degree 8t + 7 is below nl
            const int o = 2 * K;
            const T *aq = tabQ + (jb + o + par) * 4 + j, *bq = coef + (segQ * chl + jb + o + par) * 2 + pQ;
            const T *ad = tabD + (jb + o + 1 - par) * 4 + j, *bd = coef + (segD * chl + jb + o + 1 - par) * 2 + part;
            double acc = 0.0, acc2 = 0.0;
            int t0 = 0;
            for (; t0 + FZ_UNIT <= full; t0 += FZ_UNIT) {
                T a[FZ_UNIT], b[FZ_UNIT], a2[FZ_UNIT], b2[FZ_UNIT];
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    a[u] = aq[(t0 + u) * DN * 4];
                    b[u] = sQ * bq[(t0 + u) * DN * 2];
                    if (!scalar) {
                        a2[u] = ad[(t0 + u) * DN * 4];
                        b2[u] = sD * bd[(t0 + u) * DN * 2];
                    }
                }
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    acc = mfma_f64_4x4(a[u], b[u], acc);
                    if (!scalar) acc2 = mfma_f64_4x4(a2[u], b2[u], acc2);
                }
            }
            for (; t0 < steps; t0 += FZ_UNIT) {
                T a[FZ_UNIT], b[FZ_UNIT], a2[FZ_UNIT], b2[FZ_UNIT];
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    const int n = 8 * (t0 + u) + o;
                    const bool vq = n + par < nl, vd = n + 1 - par < nl;
                    const T xa = aq[(t0 + u) * DN * 4], xb = bq[(t0 + u) * DN * 2];
                    a[u] = vq ? xa : T(0);
                    b[u] = vq ? sQ * xb : T(0);
                    if (!scalar) {
                        const T ya = ad[(t0 + u) * DN * 4], yb = bd[(t0 + u) * DN * 2];
                        a2[u] = vd ? ya : T(0);
                        b2[u] = vd ? sD * yb : T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < FZ_UNIT; ++u) {
                    if (t0 + u >= steps) break;
                    acc = mfma_f64_4x4(a[u], b[u], acc);
                    if (!scalar) acc2 = mfma_f64_4x4(a2[u], b2[u], acc2);
                }
            }
            F[((m * 4 + K) * 8 + 4 * cq + j) * 2 + par] = scalar ? acc : acc + acc2;
        }
    }
}

// Legendre stage of a tile: bands of orders whose coefficients fit the staging area.
template <int PREC>
__device__ __forceinline__ void syn_legendre(const Job &J, int g, int tile, int wave, int lane, void *F, char *S)
{
    const int Tr = J.c.T, chl = J.chl;
    int m0 = 0;
    while (m0 <= Tr) {
        const int lb0 = lm_base(Tr, m0);
        int m1 = m0 + 1;
        while (m1 <= Tr && lm_base(Tr, m1 + 1) - lb0 <= chl) ++m1;
        const int nlmc = lm_base(Tr, m1) - lb0;
        if (m0 > 0) __syncthreads();   // the previous band's operands are consumed
        syn_stage<PREC>(J, g, tile, lb0, nlmc, chl, S, wave, lane);
        __syncthreads();
        syn_band<PREC>(J, m0, m1, lb0, chl, S, wave, lane, F);
        m0 = m1;
    }
}

// Z_m and Z_{N-m} of one FFT row from the tile's Fourier coefficients.
template <class T, int N>
__device__ __forceinline__ void syn_pack_rows(const Job &J, const T *F, cx<T> *bufA, cx<T> *tw, const cx<T> *twg)
{
    const int M = J.c.T;
    const int adj = J.adj;
    constexpr int NT = FZ_THREADS;
    const int nz = N - 2 * M - 1;
    for (int idx = threadIdx.x; idx < FZ_ROWS * nz; idx += NT) {
        const int rp = idx % FZ_ROWS, k = M + 1 + idx / FZ_ROWS;
        bufA[pidx(rp * N + k)] = cx<T>{T(0), T(0)};
    }
    for (int idx = threadIdx.x; idx < FZ_ROWS * (M + 1); idx += NT) {
        const int rp = idx % FZ_ROWS, m = idx / FZ_ROWS;
        const int pl = rp >> 2, gfl = rp & 3;
        const T *f = F + ((m * 4 + pl) * 8 + 2 * gfl) * 2;   // [col re: E, O][col im: E, O]
        const T hf = (adj && m > 0) ? T(0.5) : T(1);
        cx<T> zm, zc;
        synth_pack(hf * f[0], hf * f[2], hf * f[1], hf * f[3], m, zm, zc);
        bufA[pidx(rp * N + m)] = zm;
        if (m > 0) bufA[pidx(rp * N + N - m)] = zc;
    }
    for (int k = threadIdx.x; k < N; k += NT) tw[k] = twg[k];
}

template <int PREC, int N>
__device__ __forceinline__ void syn_fused_body(const Job &J, int task, char *smem)
{
    using T = typename Prec<PREC>::T;
    using T2 = cx<T>;
    const CfgDev &C = J.c;
    const int ntile = (C.nh + 3) >> 2;
    const int t = task - J.ltask0;
    const int g = t / ntile, tile = t - g * ntile;
    const int p0 = tile * 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // LDS: F | union{ staging, FFT buffers }
    T *F = (T *)smem;
    char *U = smem + fz_fbytes(C.T, sizeof(T));
    T2 *bufA = (T2 *)U;
    constexpr int PL = padded_len<FZ_ROWS, N>();
    T2 *bufB = bufA + PL, *tw = bufB + PL;
    syn_legendre<PREC>(J, g, tile, wave, lane, F, U);
    __syncthreads();
    syn_pack_rows<T, N>(J, F, bufA, tw, (const T2 *)C.twN);
    __syncthreads();
    const T2 *res = fft_block_fixed<+1, T, N, FZ_ROWS, FZ_THREADS>(bufA, bufB, tw);
    const int nh = C.nh, nlat = C.nlat, ngf = J.ngf, adj = J.adj;
    const double *wpair = C.wpair;
    for (int idx = threadIdx.x; idx < FZ_ROWS * N; idx += FZ_THREADS) {
        const int rp = idx / N, i = idx - rp * N;
        const int pair = p0 + (rp >> 2), gf = g * 4 + (rp & 3);
        if (pair >= nh || gf >= ngf) continue;
        double sc;
        double *out = syn_out(J, gf, sc);
        const T2 z = res[pidx(idx)];
        const int js = nlat - 1 - pair;
        if (adj) sc *= wpair[pair] * (js == pair ? 2.0 : 1.0);
        out[(long)pair * N + i] = sc * (double)z.x;
        if (js != pair) out[(long)js * N + i] = sc * (double)z.y;
    }
}

#define AER_FZ_DISPATCH(BODY, PREC, ...)                                                    \
    switch (J.c.nlon) {                                                                     \
        case 64: BODY<PREC, 64>(__VA_ARGS__); break;                                        \
        case 96: BODY<PREC, 96>(__VA_ARGS__); break;                                        \
        case 128: BODY<PREC, 128>(__VA_ARGS__); break;                                      \
        case 160: BODY<PREC, 160>(__VA_ARGS__); break;                                      \
        case 192: BODY<PREC, 192>(__VA_ARGS__); break;                                      \
        case 256: BODY<PREC, 256>(__VA_ARGS__); break;                                      \
        case 320: BODY<PREC, 320>(__VA_ARGS__); break;                                      \
        default: break;                                                                     \
    }

template <int PM, bool PLAN>
__global__ __launch_bounds__(FZ_THREADS) void k_syn_fused(const LaunchArgs A)
{
    extern __shared__ char smem_raw[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= A.ntask) return;
    const Job &J = job_of<PLAN>(A, task, true);
    if (task - J.ltask0 >= J.ltasks) return;
    if (PM == PM_FP32 || (PM == PM_MIXED && J.prec == 1)) {
        AER_FZ_DISPATCH(syn_fused_body, 1, J, task, smem_raw)
    } else {
        AER_FZ_DISPATCH(syn_fused_body, 0, J, task, smem_raw)
    }
}

// ------------------------------------------------------------------ analysis FFT

template <int PREC, int N>
__device__ __forceinline__ void fft_anal4_body(const Job &J, int task, char *smem, void *four_base)
{
    using T = typename Prec<PREC>::T;
    using T2 = cx<T>;
    const CfgDev &C = J.c;
    constexpr int NT = FZ_THREADS;
    constexpr int PL = padded_len<FZ_ROWS, N>();
    T2 *bufA = (T2 *)smem, *bufB = bufA + PL, *tw = bufB + PL;
    const int ntile = (C.nh + 3) >> 2;
    const int t = task - J.ftask0;
    const int g = t / ntile, tile = t - g * ntile;
    const int p0 = tile * 4;
    // all grid loads of the block are issued before the first wait; row rp = 4 gfl + pair_l
    static_assert((FZ_ROWS * N) % NT == 0, "tile rows must fill the block");
    constexpr int NLD = FZ_ROWS * N / NT, NTW = (N + NT - 1) / NT;
    T2 z[NLD];
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
        const int idx = threadIdx.x + u * NT;
        const int rp = idx / N, i = idx - rp * N;
        const int pair = p0 + (rp & 3), gf = g * 4 + (rp >> 2);
        const bool v = pair < C.nh && gf < J.ngf;
        double sc = 0.0;
        const double *in = ana_in(J, v ? gf : 0, sc);
        const int pn = v ? pair : 0;
        z[u] = v ? T2{T(sc * in[(long)pn * N + i]), T(sc * in[(long)(C.nlat - 1 - pn) * N + i])} : T2{0, 0};
    }
    const T2 *twg = (const T2 *)C.twN;
    T2 twv[NTW];
#pragma unroll
    for (int u = 0; u < NTW; ++u) {
        const int k = threadIdx.x + u * NT;
        twv[u] = k < N ? twg[k] : T2{0, 0};
    }
#pragma unroll
    for (int u = 0; u < NLD; ++u) bufA[pidx(threadIdx.x + u * NT)] = z[u];
#pragma unroll
    for (int u = 0; u < NTW; ++u) {
        const int k = threadIdx.x + u * NT;
        if (k < N) tw[k] = twv[u];
    }
    __syncthreads();
    const int M = C.T, nh = C.nh, nh4 = (C.nh + 3) / 4 * 4, ngf = J.ngf, adj = J.adj, nlat = C.nlat;
    const double *wpair = C.wpair;
    T *four = (T *)((char *)four_base + J.four_off);
    const T2 *res = fft_block_fixed<-1, T, N, FZ_ROWS, NT>(bufA, bufB, tw);
    for (int idx = threadIdx.x; idx < FZ_ROWS * (M + 1); idx += NT) {
        const int rp = idx % FZ_ROWS, m = idx / FZ_ROWS;
        const int pair = p0 + (rp & 3), gf = g * 4 + (rp >> 2);
        T *o = four + (((long)gf * (M + 1) + m) * nh4 + pair) * 4;
        T e[4] = {T(0), T(0), T(0), T(0)};
        if (pair < nh && gf < ngf) {
            const double w = adj ? ((nlat - 1 - pair == pair) ? 0.5 : 1.0) * (m > 0 ? 2.0 : 1.0) : wpair[pair];
            const T2 zm = res[pidx(rp * N + m)], zc = res[pidx(rp * N + (m == 0 ? 0 : N - m))];
            T u4[18];
            anal_unpack(zm, zc, T(w), u4);
            e[0] = u4[0];
            e[1] = u4[1];
            e[2] = u4[16];
            e[3] = u4[17];
        }
        if constexpr (PREC == 1) {
            *(float4 *)o = make_float4(e[0], e[1], e[2], e[3]);
        } else {
            ((double2 *)o)[0] = make_double2(e[0], e[1]);
            ((double2 *)o)[1] = make_double2(e[2], e[3]);
        }
    }
}

template <int PM, bool PLAN>
__global__ __launch_bounds__(FZ_THREADS) void k_fft_anal4(const LaunchArgs A)
{
    extern __shared__ char smem_raw[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= A.ntask) return;
    const Job &J = job_of<PLAN>(A, task, false);
    if (PM == PM_FP32 || (PM == PM_MIXED && J.prec == 1)) {
        AER_FZ_DISPATCH(fft_anal4_body, 1, J, task, smem_raw, A.four)
    } else {
        AER_FZ_DISPATCH(fft_anal4_body, 0, J, task, smem_raw, A.four)
    }
}

// ------------------------------------------------------------------ analysis Legendre

// Column groups: scalar jobs 4 fields (column c = 4cq + j: field 2cq + j/2, part j%2);
// vector jobs 2 items (item cq, component j/2: 0 = S, 1 = T, part j%2).
__device__ __forceinline__ int la_groups(const Job &J) { return J.kind == ANA_SCALAR ? (J.nitem + 3) / 4 : (J.nitem + 1) / 2; }
__device__ __forceinline__ int la_mpairs(int T) { return (T + 2) / 2; }

// Block = (column group, orders ma = mp and mb = T - mp). The Fourier coefficients of
// the block's grid fields for both orders are staged in LDS; the tables are read
// from L2 in layouts whose values for one MFMA step are contiguous.
//   FP64 (4x4x4, 4 blocks): block q = cq + 2 par, K = L/16 = pair in a step of 4;
//        unit = (m, row quad rq): 4 degrees k = 4rq + i of each parity; tables
//        [m][parity][rq][pair][i].
//   FP32 (4x4x1, 16 blocks): block b = L/4 = cq + 2 par + 4 rs; one pair per step;
//        unit = (m, group G of 4 row quads): degrees k = 16G + 4rs + i; tables
//        [m][parity][G][pair][rs][i].
// Each block accumulates one output parity; no cross-lane sums.
template <int PREC>
__device__ __forceinline__ void leg_anal4_body(const Job &J, int task, int lane, int wave, char *smem, const void *four_base)
{
    using T = typename Prec<PREC>::T;
    using Acc = typename std::conditional<PREC == 1, f4, double>::type;
    const CfgDev &C = J.c;
    const int Tr = C.T;
    const int ncg = la_groups(J);
    const int t = task - J.ltask0;
    const int mp = t / ncg, cg = t - mp * ncg;
    const int ma = mp, mb = Tr - mp;
    const int nm = ma == mb ? 1 : 2;
    const int nh4 = (C.nh + 3) / 4 * 4;
    const bool scalar = J.kind == ANA_SCALAR;
    const bool has_ek = !scalar && J.gin2 != nullptr;
    const int gpi = scalar ? 1 : (has_ek ? 3 : 2);
    const int nlgf = scalar ? 4 : 2 * gpi;          // grid fields this block reads
    const int gf0 = scalar ? cg * 4 : cg * 2 * gpi;
    // stage Fs[msel][lgf][pair][4] for both orders: all loads in flight, then to LDS
    T *Fs = (T *)smem;
    const T *four = (const T *)((const char *)four_base + J.four_off);
    {
        const int nitem4 = nm * nlgf * nh4;   // 4-value vectors
        constexpr int U = 8;
        for (int base = 0; base < nitem4; base += LA_THREADS * U) {
            V4<T> v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = base + threadIdx.x + u * LA_THREADS;
                const int pair = idx % nh4, r = idx / nh4;
                const int lgf = r % nlgf, msel = r / nlgf;
                const int gf = gf0 + lgf;
                const bool ok = idx < nitem4 && gf < J.ngf;
                const int m = msel ? mb : ma;
                v[u] = ok ? V4<T>::load(four + (((long)gf * (Tr + 1) + m) * nh4 + pair) * 4) : V4<T>::zero();
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = base + threadIdx.x + u * LA_THREADS;
                if (idx < nitem4) v[u].store(Fs + (long)idx * 4);
            }
        }
    }
    __syncthreads();
    const int nwave = LA_THREADS / 64;
    const T *tP = (const T *)C.tabYP, *tD = (const T *)C.tabYD, *tQ = (const T *)C.tabYQ;
    const int j = lane & 3, cq = (lane >> 2) & 1, par = (lane >> 3) & 1, hi = lane >> 4;
    const int part = j & 1, comp = j >> 1;
    // B sources in Fs for this lane's column (scalar: one; vector: dP, Q and the Phi+K term)
    int lD = 0, lQ = 0, lE = 0, eD = 0, eQ = 0, eE = 0;
    T sD = T(1), sQ = T(1);
    bool useE = false;
    if (scalar) {
        lD = 2 * cq + comp;
        eD = par * 2 + part;
    } else {
        const int gt = cq * gpi, gp = gt + 1;
        lD = comp == 0 ? gt : gp;
        eD = (1 - par) * 2 + part;
        sD = comp == 0 ? T(1) : T(-1);
        lQ = comp == 0 ? gp : gt;
        eQ = par * 2 + (1 - part);
        sQ = part ? T(-1) : T(1);
        lE = has_ek ? gt + 2 : gt;
        eE = par * 2 + part;
        useE = has_ek && comp == 0;
    }
    // units of both orders: FP64 (m, rq), FP32 (m, G)
    const int nq = C.nq, nG = (C.nq + 3) / 4;
    auto nunits = [&](int m) {
        const int nk0 = (Tr - m + 2) / 2;
        const int q = (nk0 + 3) / 4;
        return PREC == 1 ? (q + 3) / 4 : q;
    };
    const int nuA = nunits(ma), nuB = nm == 2 ? nunits(mb) : 0;
    // pairs per MFMA step: FP64 4 (K = L/16), FP32 1 (block slot rs = L/16 picks degrees instead)
    constexpr int PPS = PREC == 1 ? 1 : 4;
    const int nsteps = nh4 / PPS;
    for (int u = wave; u < nuA + nuB; u += nwave) {
        const bool second = u >= nuA;
        const int m = second ? mb : ma, uq = second ? u - nuA : u;
        const T *Fm = Fs + (long)(second ? 1 : 0) * nlgf * nh4 * 4;
        const int nl = Tr - m + 1;
        const int nk = par ? nl / 2 : (nl + 1) / 2;
        // degree row of this lane's A operand and of its outputs
        long abase;
        if constexpr (PREC == 1) abase = (((long)(m * 2 + par) * nG + uq) * nh4) * 16 + 4 * hi + j;
        else abase = (((long)(m * 2 + par) * nq + uq) * nh4) * 4 + j;
        Acc acc{}, acc2{};
        // per step the pair index moves by PPS: A by PPS*4 (FP64) or 16 (FP32) values,
        // B by PPS*4 values; all pairs below nh4 exist (zero-padded), so only a partial
        // last unit needs a guard
        constexpr int ASTEP = PREC == 1 ? 16 : 16;
        const int p0 = PREC == 1 ? 0 : hi;
        const T *pa = tP + abase + (long)p0 * (PREC == 1 ? 16 : 4);
        const T *pd = tD + abase + (long)p0 * (PREC == 1 ? 16 : 4);
        const T *pq = tQ + abase + (long)p0 * (PREC == 1 ? 16 : 4);
        const T *fd = Fm + lD * nh4 * 4 + p0 * 4 + eD;
        const T *fq = Fm + lQ * nh4 * 4 + p0 * 4 + eQ;
        const T *fk = Fm + lE * nh4 * 4 + p0 * 4 + eE;
        constexpr int BSTEP = PPS * 4;
        for (int s0 = 0; s0 < nsteps; s0 += FZ_UNIT) {
            const bool whole = s0 + FZ_UNIT <= nsteps;
            T aP[FZ_UNIT], aD[FZ_UNIT], aQ[FZ_UNIT], bD[FZ_UNIT], bQ[FZ_UNIT], bE[FZ_UNIT];
#pragma unroll
            for (int uu = 0; uu < FZ_UNIT; ++uu) {
                const int st = (whole || s0 + uu < nsteps) ? s0 + uu : nsteps - 1;
                if (scalar) {
                    aP[uu] = pa[(long)st * ASTEP];
                    bD[uu] = fd[st * BSTEP];
                } else {
                    aD[uu] = pd[(long)st * ASTEP];
                    aQ[uu] = pq[(long)st * ASTEP];
                    bD[uu] = sD * fd[st * BSTEP];
                    bQ[uu] = sQ * fq[st * BSTEP];
                    if (has_ek) {
                        aP[uu] = pa[(long)st * ASTEP];
                        bE[uu] = useE ? fk[st * BSTEP] : T(0);
                    }
                }
            }
#pragma unroll
            for (int uu = 0; uu < FZ_UNIT; ++uu) {
                if (s0 + uu >= nsteps) break;
                if constexpr (PREC == 1) {
                    if (scalar) {
                        acc = mfma_f32_4x4(aP[uu], bD[uu], acc);
                    } else {
                        acc = mfma_f32_4x4(aD[uu], bD[uu], acc);
                        acc = mfma_f32_4x4(aQ[uu], bQ[uu], acc);
                        if (has_ek) acc2 = mfma_f32_4x4(aP[uu], bE[uu], acc2);
                    }
                } else {
                    if (scalar) {
                        acc = mfma_f64_4x4(aP[uu], bD[uu], acc);
                    } else {
                        acc = mfma_f64_4x4(aD[uu], bD[uu], acc);
                        acc = mfma_f64_4x4(aQ[uu], bQ[uu], acc);
                        if (has_ek) acc2 = mfma_f64_4x4(aP[uu], bE[uu], acc2);
                    }
                }
            }
        }
        // epilogue: output column c = 4cq + j
        const int c = 4 * cq + j;
        int item, ocomp;
        if (scalar) { item = cg * 4 + (c >> 1); ocomp = 0; }
        else { item = cg * 2 + cq; ocomp = comp; }
        double *out = ocomp ? J.sout1 : J.sout0;
        if (item >= J.nitem || out == nullptr) continue;
        double *o = out + (long)item * J.sstride;
        const double *fa = C.lfac + (ocomp ? J.mode1 : J.mode0) * (Tr + 1);
        const double *fe = C.lfac + J.mode2 * (Tr + 1);
        const int lmb = lm_base(Tr, m);
        if constexpr (PREC == 1) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kr = 16 * uq + 4 * hi + i;
                if (kr >= nk) continue;
                const int l = m + par + 2 * kr;
                double v = fa[l] * (double)acc[i];
                if (useE) v += fe[l] * (double)acc2[i];
                const long lm = lmb + (l - m);
                if (J.accumulate) o[2 * lm + part] += v;
                else o[2 * lm + part] = v;
            }
        } else {
            const int kr = 4 * uq + hi;
            if (kr < nk) {
                const int l = m + par + 2 * kr;
                double v = fa[l] * acc;
                if (useE) v += fe[l] * acc2;
                const long lm = lmb + (l - m);
                if (J.accumulate) o[2 * lm + part] += v;
                else o[2 * lm + part] = v;
            }
        }
    }
}

template <int PM, bool PLAN>
__global__ __launch_bounds__(LA_THREADS) void k_leg_anal4(const LaunchArgs A)
{
    extern __shared__ char smem_raw[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= A.ntask) return;
    const Job &J = job_of<PLAN>(A, task, true);
    if (task - J.ltask0 >= J.ltasks) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (PM == PM_FP32 || (PM == PM_MIXED && J.prec == 1)) leg_anal4_body<1>(J, task, lane, wave, smem_raw, A.four);
    else leg_anal4_body<0>(J, task, lane, wave, smem_raw, A.four);
}

}  // namespace aer
