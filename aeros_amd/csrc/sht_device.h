// sht_device.h -- gfx950 kernels of the spherical-harmonic transform.
//
// A transform is two stages with a Fourier intermediate in HBM between them:
//
//   synthesis:  k_leg_synth  (per-m Legendre GEMM on MFMA, hemispheric folding)
//               k_fft_synth  (LDS Stockham complex FFT of (north + i south) rows)
//   analysis:   k_fft_anal   (the same FFT, forward, weights folded in)
//               k_leg_anal   (per-m Legendre GEMM on MFMA)
//
// Hemispheric folding: P_lm(-mu) = (-1)^(l-m) P_lm(mu), so with E and O the
// even- and odd-(l-m) partial sums at a northern latitude, the northern and
// southern Fourier coefficients are E+O and E-O. The Legendre GEMM for one m
// therefore has rows = northern latitudes (padded to 16), K = the degrees l of
// one parity, columns = 8 complex fields (16 real columns). dP/dtheta has the
// opposite parity, m P/sin(theta) the same.
//
// Every launch takes a table of up to AER_MAXJOB jobs; each job is one batched
// transform kind on one truncation, so a whole transform set (or an ensemble
// of truncations) is one launch per stage.
//
// Intermediate layout, per job: [group of 8 grid fields][m][pair][E|O][field(8)][re|im].
// For a fixed (group, m, pair) the 32 values are contiguous, which makes both
// the MFMA output stores and the FFT-side loads 128-byte coalesced.
#pragma once
#include <hip/hip_runtime.h>

#define AER_MAXJOB 8
#define AER_MAXPASS 12
#define AER_NMODE 8
#define AER_FFT_THREADS 128

namespace aer {

enum JobKind { SYN_SCALAR = 0, SYN_VECTOR = 1, SYN_GRAD = 2, ANA_SCALAR = 3, ANA_VECTOR = 4 };

// Per-degree factors applied to spectral inputs (synthesis) or outputs (analysis).
enum Mode {
    M_ONE = 0,         // 1
    M_S_FROM_DIV = 1,  // -a/(l(l+1)), 0 at l=0    (src/dynamics/aeros_vordiv.f90:100-107)
    M_T_FROM_VOR = 2,  // +a/(l(l+1)), 0 at l=0
    M_INV_LL1 = 3,     // 1/(l(l+1)), 0 at l=0     (raw spheroidal/toroidal analysis)
    M_DIV_FROM_S = 4,  // -1/a, 0 at l=0  (= -l(l+1)/a * 1/(l(l+1)), src/dynamics/aeros_vordiv.f90:155-164)
    M_VOR_FROM_T = 5,  // +1/a, 0 at l=0
    M_NEG_LAP = 6,     // +l(l+1)/a^2   (-laplacian, src/dynamics/aeros_tendency.f90:458-460)
    M_ZERO = 7         // 0
};

struct CfgDev {
    int T, nlat, nlon, nh, nhp, npt, nlm;  // nh latitude pairs, nhp padded to 16, npt = nhp/16
    int pb, fb;                            // FFT block: fb grid fields (divides 8) x pb latitude pairs
    const void *tabP, *tabD, *tabQ;        // [nlm][nhp] in the compute type
    const void *tabPT, *tabDT, *tabQT;     // [m][parity][nhp][kp]: degree l = m + parity + 2k fastest
    int kp;                                // k extent of the transposed tables (multiple of 16)
    const void *tabZP, *tabZD, *tabZQ;     // tiled synthesis: [tile of 4 pairs][lm][4]
    const void *tabYP, *tabYD, *tabYQ;     // tiled analysis: FP64 [m][parity][row quad][pair (nh rounded to 4)][4],
                                           // FP32 [m][parity][4 row quads][pair][16]
    int nq;                                // row quads of the tiled analysis tables: ((T+2)/2 + 3)/4
    const int *expP, *expD, *expQ;         // split form: row exponents [m][parity][nhp]
    const int *expPT, *expDT, *expQT;      // split form: row exponents [m][parity][kp]
    const double *lfac;                    // [AER_NMODE][T+1]
    const double *lfacm;                   // [AER_NMODE][nlm]: the same factors per coefficient
    const double *wpair;                   // [nhp] w_j * 2pi/nlon (x1/2 on a self-mirrored equator row)
    const void *tw;                        // per-pass forward twiddles (generic FFT path)
    const void *twN;                       // e^{-2 pi i k/nlon}, k < nlon (compile-time FFT paths)
    int npass;
    int radix[AER_MAXPASS];
    int twoff[AER_MAXPASS];
};

struct Job {
    CfgDev c;
    int kind, nitem, ngf, ngroup, lgroups;
    int ltask0, ftask0;
    long sstride, gstride;
    const double *sin0, *sin1;        // synthesis spectral inputs
    double *sout0, *sout1;            // analysis spectral outputs
    double *gout0, *gout1;            // synthesis grid outputs
    const double *gin0, *gin1, *gin2; // analysis grid inputs (gin2: optional Phi+K for the assembly)
    int mode0, mode1, mode2;
    double gsc0, gsc1;                // grid-side scale of component 0 / 1
    int accumulate;
    int prec;                         // compute type of this job: 0 = FP64, 1 = FP32, 2 = split FP32
    int adj;                          // transpose of the opposite transform: synthesis jobs halve
                                      // m > 0 and weight rows by w_j 2pi/nlon; analysis jobs use
                                      // unit row weights and double m > 0
    int ltasks;                       // Legendre tasks of the job (its range is padded to whole blocks)
    int chl;                          // tiled synthesis: coefficients per staging band
    long four_off;                    // byte offset of the job's intermediate region
};

// Kernel precision modes: every job FP64, every job FP32, every job split-FP32, or
// per job (the kernel branches on Job::prec; a block never spans two jobs, so the
// branch is uniform).
enum PrecMode { PM_FP64 = 0, PM_FP32 = 1, PM_SPLIT = 2, PM_MIXED = 3 };

// Job table of one launch: inline in the kernel arguments (up to AER_MAXJOB jobs),
// or in device memory (djobs != nullptr, any number of jobs; compiled plans).
struct LaunchArgs {
    int njob, ntask;
    int lstart[AER_MAXJOB], fstart[AER_MAXJOB];   // first Legendre / FFT task of each inline job
    int kmax;               // max T+1 over the jobs (LDS staging of the Legendre synthesis)
    void *four;
    const Job *djobs;       // device job table (compiled plans)
    const int *task_job;    // device task -> job index (compiled plans)
    Job job[AER_MAXJOB];
};

template <class T> struct __attribute__((aligned(2 * sizeof(T)))) cx {
    T x, y;
};

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Operand layouts of the 16x16x4 MFMAs (lane L): A[L%16][L/16], B[L/16][L%16] for
// both types. Accumulator element r of lane L holds D[row(L/16, r)][L%16], where
// row is 4r + L/16 for f64 and 4(L/16) + r for f32 (measured on gfx950).
template <int PREC> struct Prec;
template <> struct Prec<0> {
    using T = double;
    using T2 = cx<double>;
    using Acc = d4;
    __device__ static inline Acc mfma(T a, T b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static inline int row(int kq, int r) { return 4 * r + kq; }
};
template <> struct Prec<1> {
    using T = float;
    using T2 = cx<float>;
    using Acc = f4;
    __device__ static inline Acc mfma(T a, T b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int row(int kq, int r) { return 4 * kq + r; }
};

// Split-FP32 form: FP64 operands carried as five 10-bit integer slices,
// x = 2^e (a1 2^-10 + a2 2^-20 + ... + a5 2^-50), with e a power-of-two scale
// shared by every product that one MFMA step adds into one output element: per
// row of A (over all K) and per column of B (over the step's 4 K values). With
// |a1| <= 2^10 and |a_s| <= 2^9 after it, a level (s + t - 2 = 0..4; the products
// with s + t > 6 are below 2^-51) sums at most 7 * 2^20 over a step, so FP32
// accumulates it exactly; the five levels are promoted to FP64 after every step.
// The A slices are built on the host: four as float4, the fifth in a float array
// that follows the float4 region of the same table.
template <> struct Prec<2> {
    using T = double;          // B operands, staging and the Fourier intermediate
    using Acc = d4;
    __device__ static inline int row(int kq, int r) { return 4 * kq + r; }   // FP32 MFMA layout
};

constexpr int SPLIT_BITS = 10;

__device__ __forceinline__ f4 mfma32(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// One 16x16x4 step of A (this lane's slices a, rows' exponents ea of the
// accumulator rows 4(L/16)+r) times B (this lane's value b at [L/16][L%16]).
// Returns the step's FP64 contribution to the lane's 4 accumulator elements.
struct Exp4 {
    int r0, r1, r2, r3;
};

__device__ __forceinline__ d4 split_step(const float4 a, const float a5, const Exp4 ea, double b)
{
    double mb = fabs(b);
    mb = fmax(mb, __shfl_xor(mb, 16));
    mb = fmax(mb, __shfl_xor(mb, 32));   // the column's 4 K values live in lanes L ^ 16, L ^ 32
    const int fb = mb > 0.0 ? ilogb(mb) + 1 : 0;
    double y = ldexp(b, SPLIT_BITS - fb);
    const double b1 = rint(y);
    y = (y - b1) * 1024.0;
    const double b2 = rint(y);
    y = (y - b2) * 1024.0;
    const double b3 = rint(y);
    y = (y - b3) * 1024.0;
    const double b4 = rint(y);
    y = (y - b4) * 1024.0;
    const double b5 = rint(y);
    const float s1 = (float)b1, s2 = (float)b2, s3 = (float)b3, s4 = (float)b4, s5 = (float)b5;
    const f4 z = {0.f, 0.f, 0.f, 0.f};
    const f4 c0 = mfma32(a.x, s1, z);
    f4 c1 = mfma32(a.x, s2, z);
    c1 = mfma32(a.y, s1, c1);
    f4 c2 = mfma32(a.x, s3, z);
    c2 = mfma32(a.y, s2, c2);
    c2 = mfma32(a.z, s1, c2);
    f4 c3 = mfma32(a.x, s4, z);
    c3 = mfma32(a.y, s3, c3);
    c3 = mfma32(a.z, s2, c3);
    c3 = mfma32(a.w, s1, c3);
    f4 c4 = mfma32(a.x, s5, z);
    c4 = mfma32(a.y, s4, c4);
    c4 = mfma32(a.z, s3, c4);
    c4 = mfma32(a.w, s2, c4);
    c4 = mfma32(a5, s1, c4);
    const double k = 1.0 / 1024.0;
    const int e[4] = {ea.r0, ea.r1, ea.r2, ea.r3};
    d4 r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double t =
            ((((double)c4[q] * k + (double)c3[q]) * k + (double)c2[q]) * k + (double)c1[q]) * k + (double)c0[q];
        r[q] = ldexp(t, e[q] + fb - 2 * SPLIT_BITS);
    }
    return r;
}

__device__ __forceinline__ Exp4 load_exp4(const int *p) { return Exp4{p[0], p[1], p[2], p[3]}; }

// A slices of element idx, zero where !v: the load always reads a valid element
// (index 0 when masked) and the mask is applied to the values, so no zero
// aggregate is ever addressed.
__device__ __forceinline__ float4 ld_slices(const float4 *__restrict__ p, long idx, bool v)
{
    const float4 a = p[v ? idx : 0];
    return make_float4(v ? a.x : 0.f, v ? a.y : 0.f, v ? a.z : 0.f, v ? a.w : 0.f);
}
__device__ __forceinline__ float ld_slice5(const float *__restrict__ p, long idx, bool v)
{
    const float a = p[v ? idx : 0];
    return v ? a : 0.f;
}
__device__ __forceinline__ double ld_masked(const double *__restrict__ p, long idx, bool v)
{
    const double x = p[v ? idx : 0];
    return v ? x : 0.0;
}

// The job a task belongs to: the last job whose first task is <= task. The task
// index is uniform over a wave, so the job index is too (readfirstlane lets the
// compiler read the job's fields with scalar loads).
// PLAN (a template parameter, so that one kernel never mixes a pointer into the
// kernel arguments with one into device memory): jobs in a device table.
template <bool PLAN>
__device__ __forceinline__ const Job &job_of(const LaunchArgs &A, int task, bool legendre)
{
    if constexpr (PLAN) {
        const int jb = __builtin_amdgcn_readfirstlane(A.task_job[task]);
        return A.djobs[jb];
    }
    // Inline table: indexed straight in the kernel-argument segment (LaunchArgs is
    // the only kernel argument, at offset 0), not through the by-value parameter,
    // whose dynamically indexed use would keep a private copy of all of it.
    // The task starts sit together at the front of the arguments, so they load at
    // once and the job index is a count, not a chain of dependent loads.
    const LaunchArgs *K = (const LaunchArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    const int *st = legendre ? K->lstart : K->fstart;
    const int nj = K->njob;
    int jb = 0;
#pragma unroll
    for (int k = 1; k < AER_MAXJOB; ++k) jb += (k < nj && task >= st[k]) ? 1 : 0;
    jb = __builtin_amdgcn_readfirstlane(jb);
    return K->job[jb];
}

// 0-based index of (l = m, m) in the m-major, l-fastest layout: the orders m' < m
// hold sum (T+1-m') = m(2T+3-m)/2 coefficients (reference src/aeros_spectral.f90:478).
__device__ inline int lm_base(int T, int m) { return m * (2 * T + 3 - m) / 2; }

// MFMA steps are processed in chunks: the operands of a whole chunk are loaded
// before its MFMAs issue, so one load latency covers the chunk. KCH_S steps of 4
// degrees cover both parities up to T = 47 in one chunk; KCH_A steps of 4
// latitude pairs cover 32 pairs (T42) in one chunk.
constexpr int KCH_S = 12;
constexpr int KCH_A = 8;
constexpr int KCH = KCH_A;

// ------------------------------------------------------------------ Legendre synthesis
// One wavefront per task (job, group of 8 grid fields, m, tile of 16 latitude pairs).
// The wave first stages the coefficients of order m of its fields in LDS (contiguous
// in memory, so the loads coalesce), scaled by the job's per-degree factor; the MFMA
// B operands are then read from LDS. Segment s of a wave holds field g*8+s (scalar
// jobs) or S (s < 4) / T (s >= 4) of item g*4 + s%4 (vector and gradient jobs).
template <int PREC>
__device__ __forceinline__ void leg_synth_body(const LaunchArgs &A, const Job &J, int task, bool active,
                                               int wave, int lane, char *smem_raw)
{
    using P = Prec<PREC>;
    using T = typename P::T;
    using Acc = typename P::Acc;
    cx<T> *sb = (cx<T> *)smem_raw + (long)wave * 8 * A.kmax;
    const CfgDev &C = J.c;
    int t = task - J.ltask0;
    const int g = t % J.lgroups;
    t /= J.lgroups;
    const int pt = t % C.npt;
    const int m = t / C.npt;
    const int lmb = lm_base(C.T, m);
    if (active) {
        const int nl = C.T - m + 1;
        const bool scalar = J.kind == SYN_SCALAR;
        // 64 degrees per pass; within a pass the 8 segments' loads are all issued
        // before the first wait (registers, then LDS).
        for (int n0 = 0; n0 < nl; n0 += 64) {
            const int n = n0 + lane;
            double re[8], im[8], fv[8];
#pragma unroll
            for (int seg = 0; seg < 8; ++seg) {
                const double *src = nullptr, *fac = nullptr;
                if (scalar) {
                    const int f = g * 8 + seg;
                    if (f < J.nitem) { src = J.sin0 + (long)f * J.sstride; fac = C.lfac + J.mode0 * (C.T + 1); }
                } else {
                    const int it = g * 4 + (seg & 3);
                    if (it < J.nitem) {
                        if (seg < 4) { src = J.sin0 + (long)it * J.sstride; fac = C.lfac + J.mode0 * (C.T + 1); }
                        else if (J.kind == SYN_VECTOR) { src = J.sin1 + (long)it * J.sstride; fac = C.lfac + J.mode1 * (C.T + 1); }
                    }
                }
                const bool v = src != nullptr && n < nl;
                re[seg] = v ? src[2 * (lmb + n)] : 0.0;
                im[seg] = v ? src[2 * (lmb + n) + 1] : 0.0;
                fv[seg] = v ? fac[m + n] : 0.0;
            }
            if (n < nl) {
#pragma unroll
                for (int seg = 0; seg < 8; ++seg)
                    sb[seg * A.kmax + n] = cx<T>{T(fv[seg] * re[seg]), T(fv[seg] * im[seg])};
            }
        }
    }
    __syncthreads();
    if (!active) return;
    const int i = lane & 15, kq = lane >> 4;
    const int part = i & 1;
    const int comp = (i >> 1) & 1;   // vector jobs: 0 = V_theta column, 1 = V_phi column
    // B operands as one LDS element each: b = sign * seg[2 (l-m) + offset]. Scalar: the
    // field's real or imaginary part. Vector, with V_theta = S*D + (iT)*Q and
    // V_phi = (iS)*Q + (-T)*D: bD = +-S or -T component, bQ = the other component of
    // iT or iS, whose real part is -Im and imaginary part +Re.
    const T *segS = (const T *)(sb + (J.kind == SYN_SCALAR ? (i >> 1) : (i >> 2)) * A.kmax);
    const T *segT = (const T *)(sb + (4 + (i >> 2)) * A.kmax);
    // Scalar jobs have no component: every lane reads its own field's segment.
    const bool useS = J.kind == SYN_SCALAR || comp == 0;
    const T *pD = (useS ? segS : segT) + part;
    const T sD = comp == 0 ? T(1) : T(-1);
    const T *pQ = (comp == 0 ? segT : segS) + (1 - part);
    const T sQ = part ? T(1) : T(-1);
    const int nhp = C.nhp;
    const T *__restrict__ tP = (const T *)C.tabP + (long)lmb * nhp + pt * 16 + i;
    const T *__restrict__ tD = (const T *)C.tabD + (long)lmb * nhp + pt * 16 + i;
    const T *__restrict__ tQ = (const T *)C.tabQ + (long)lmb * nhp + pt * 16 + i;
    Acc E = {0, 0, 0, 0}, O = {0, 0, 0, 0};
    // MFMA steps of 4 degrees: the even-(l-m) steps first, then the odd ones.
    const int nk0 = (C.T - m) / 2 + 1;
    const int nk1 = (C.T - m - 1 >= 0) ? (C.T - m - 1) / 2 + 1 : 0;
    const int s0n = (nk0 + 3) / 4, S = s0n + (nk1 + 3) / 4;
    if constexpr (PREC == 2) {
        // split form: A = float4 slices, row exponents per (table, parity, pair)
        const long to = (long)lmb * nhp + pt * 16 + i;
        const long ntab = (long)C.nlm * nhp;   // the fifth slices follow the float4 region
        const float4 *__restrict__ qP = (const float4 *)C.tabP + to;
        const float4 *__restrict__ qD = (const float4 *)C.tabD + to;
        const float4 *__restrict__ qQ = (const float4 *)C.tabQ + to;
        const float *__restrict__ rP = (const float *)((const float4 *)C.tabP + ntab) + to;
        const float *__restrict__ rD = (const float *)((const float4 *)C.tabD + ntab) + to;
        const float *__restrict__ rQ = (const float *)((const float4 *)C.tabQ + ntab) + to;
        const long xo = (long)m * 2 * nhp + pt * 16 + 4 * kq;
        const bool scal = J.kind == SYN_SCALAR;
        // (every split table exists for a split job, so the loads need no guard)
        const Exp4 eP0 = load_exp4(C.expP + xo), eP1 = load_exp4(C.expP + xo + nhp);
        const Exp4 eD0 = load_exp4(C.expD + xo), eD1 = load_exp4(C.expD + xo + nhp);
        const Exp4 eQ0 = load_exp4(C.expQ + xo), eQ1 = load_exp4(C.expQ + xo + nhp);
        constexpr int KS = 4;
        for (int c0 = 0; c0 < S; c0 += KS) {
            float4 aA[KS], aB[KS];
            float a5A[KS], a5B[KS];
            double bA[KS], bB[KS];
#pragma unroll
            for (int u = 0; u < KS; ++u) {
                const int st = c0 + u;
                const int par = st < s0n ? 0 : 1;
                const int k = (st < s0n ? st : st - s0n) * 4 + kq;
                const bool kv = st < S && k < (par ? nk1 : nk0);
                const int n = par + 2 * k;
                if (scal) {
                    aA[u] = ld_slices(qP, (long)n * nhp, kv);
                    a5A[u] = ld_slice5(rP, (long)n * nhp, kv);
                    bA[u] = kv ? pD[2 * n] : 0.0;
                } else {
                    aA[u] = ld_slices(qQ, (long)n * nhp, kv);
                    a5A[u] = ld_slice5(rQ, (long)n * nhp, kv);
                    bA[u] = kv ? sQ * pQ[2 * n] : 0.0;
                    aB[u] = ld_slices(qD, (long)n * nhp, kv);
                    a5B[u] = ld_slice5(rD, (long)n * nhp, kv);
                    bB[u] = kv ? sD * pD[2 * n] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < KS; ++u) {
                const int st = c0 + u;
                if (st >= S) continue;
                const int par = st < s0n ? 0 : 1;
                if (scal) {
                    const d4 c = split_step(aA[u], a5A[u], par ? eP1 : eP0, bA[u]);
                    if (par == 0) E += c;
                    else O += c;
                } else {
                    // P and Q have the parity of l-m, dP/dtheta the opposite one
                    const d4 cq = split_step(aA[u], a5A[u], par ? eQ1 : eQ0, bA[u]);
                    const d4 cd = split_step(aB[u], a5B[u], par ? eD1 : eD0, bB[u]);
                    if (par == 0) {
                        E += cq;
                        O += cd;
                    } else {
                        O += cq;
                        E += cd;
                    }
                }
            }
        }
    } else
    for (int c0 = 0; c0 < S; c0 += KCH_S) {
        if (J.kind == SYN_SCALAR) {
            T a[KCH_S], b[KCH_S];
#pragma unroll
            for (int u = 0; u < KCH_S; ++u) {
                const int st = c0 + u;
                const int par = st < s0n ? 0 : 1;
                const int k = (st < s0n ? st : st - s0n) * 4 + kq;
                const bool kv = st < S && k < (par ? nk1 : nk0);
                const int n = par + 2 * k;
                a[u] = kv ? tP[n * nhp] : T(0);
                b[u] = kv ? pD[2 * n] : T(0);
            }
#pragma unroll
            for (int u = 0; u < KCH_S; ++u) {
                const int st = c0 + u;
                if (st < s0n) E = P::mfma(a[u], b[u], E);
                else if (st < S) O = P::mfma(a[u], b[u], O);
            }
        } else {
            T aD[KCH_S], aQ[KCH_S], bD[KCH_S], bQ[KCH_S];
#pragma unroll
            for (int u = 0; u < KCH_S; ++u) {
                const int st = c0 + u;
                const int par = st < s0n ? 0 : 1;
                const int k = (st < s0n ? st : st - s0n) * 4 + kq;
                const bool kv = st < S && k < (par ? nk1 : nk0);
                const int n = par + 2 * k;
                aD[u] = kv ? tD[n * nhp] : T(0);
                aQ[u] = kv ? tQ[n * nhp] : T(0);
                bD[u] = kv ? sD * pD[2 * n] : T(0);
                bQ[u] = kv ? sQ * pQ[2 * n] : T(0);
            }
#pragma unroll
            for (int u = 0; u < KCH_S; ++u) {
                const int st = c0 + u;
                // P and Q have the parity of l-m, dP/dtheta the opposite one
                if (st < s0n) {
                    E = P::mfma(aQ[u], bQ[u], E);
                    O = P::mfma(aD[u], bD[u], O);
                } else if (st < S) {
                    O = P::mfma(aQ[u], bQ[u], O);
                    E = P::mfma(aD[u], bD[u], E);
                }
            }
        }
    }
    T *four = (T *)((char *)A.four + J.four_off);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int pair = pt * 16 + P::row(kq, r);
        const long base = (((long)g * (C.T + 1) + m) * C.nhp + pair) * 32;
        four[base + i] = E[r];
        four[base + 16 + i] = O[r];
    }
}

// The block's 4 waves: tasks blockIdx*4 .. +3, all of one job (job task ranges are
// padded to multiples of 4), so the precision branch and the staging barrier are
// block-uniform.
template <int PM, bool PLAN>
__global__ __launch_bounds__(256) void k_leg_synth(const LaunchArgs A)
{
    extern __shared__ char smem_raw[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + wave;
    const Job &J = job_of<PLAN>(A, task < A.ntask ? task : A.ntask - 1, true);
    const bool active = task < A.ntask && task - J.ltask0 < J.ltasks;
    if constexpr (PM == PM_MIXED) {
        if (J.prec == 1) leg_synth_body<1>(A, J, task, active, wave, lane, smem_raw);
        else if (J.prec == 2) leg_synth_body<2>(A, J, task, active, wave, lane, smem_raw);
        else leg_synth_body<0>(A, J, task, active, wave, lane, smem_raw);
    } else {
        leg_synth_body<PM>(A, J, task, active, wave, lane, smem_raw);
    }
}

// ------------------------------------------------------------------ Legendre analysis
// One wavefront per task (job, column group, m, parity x tile of 16 degrees).
// Rows = degrees l of one parity, K = latitude pairs, columns = 8 complex outputs
// (scalar) or 4 (S, T) pairs (vector).
template <int PREC>
__device__ __forceinline__ void leg_anal_body(const LaunchArgs &A, const Job &J, int task, int lane)
{
    using P = Prec<PREC>;
    using T = typename P::T;
    using Acc = typename P::Acc;
    const CfgDev &C = J.c;
    const int ntpp = ((C.T + 2) / 2 + 15) / 16;  // tiles per parity
    // Task order: (parity, degree tile) fastest, then field group, then m. The tasks of one
    // (group, m) read the same intermediate rows (both parities share each 128-byte pair
    // record), so they run in one block, or in neighbouring blocks, and share L1/L2.
    int t = task - J.ltask0;
    const int tt = t % (2 * ntpp);
    t /= 2 * ntpp;
    const int g = t % J.lgroups;
    const int m = t / J.lgroups;
    const int par = tt / ntpp, tile = tt % ntpp;
    const int nk = (C.T - m - par >= 0) ? (C.T - m - par) / 2 + 1 : 0;
    if (tile * 16 >= nk) return;
    const int i = lane & 15, kq = lane >> 4;
    const int lmb = lm_base(C.T, m);
    const int kra = tile * 16 + i;
    const bool rv = kra < nk;
    // A[l][pair] from the transposed tables: 16 consecutive degrees per row of lanes
    const long abase = (long)(m * 2 + par) * C.nhp * C.kp + kra;
    const T *__restrict__ tP = (const T *)C.tabPT;
    const T *__restrict__ tD = (const T *)C.tabDT;
    const T *__restrict__ tQ = (const T *)C.tabQT;
    const T *__restrict__ four = (const T *)((const char *)A.four + J.four_off);
    const long mstride = (long)(C.T + 1) * C.nhp * 32;  // one intermediate group
    const long mo = (long)m * C.nhp * 32;
    Acc acc = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};
    int item, comp = 0;
    const int part = i & 1;
    if (J.kind == ANA_SCALAR) item = g * 8 + (i >> 1);
    else { item = g * 4 + (i >> 2); comp = (i >> 1) & 1; }
    const bool cv = item < J.nitem;
    const bool has_ek = (J.kind == ANA_VECTOR) && J.gin2 != nullptr;
    const int gpi = has_ek ? 3 : 2;  // grid fields per vector item
    // the epilogue's per-degree factors, loaded now so that no load waits after the MFMAs
    double fav[4], fev[4];
    {
        const double *fa = C.lfac + ((J.kind == ANA_VECTOR && comp == 1) ? J.mode1 : J.mode0) * (C.T + 1);
        const double *fe = C.lfac + J.mode2 * (C.T + 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int kr = tile * 16 + P::row(kq, r);
            const int l = m + par + 2 * kr;
            fav[r] = kr < nk ? fa[l] : 0.0;
            fev[r] = (has_ek && comp == 0 && kr < nk) ? fe[l] : 0.0;
        }
    }
    if constexpr (PREC == 2) {
        // split form: A = float4 slices, row exponents per (table, degree row)
        const float4 *__restrict__ qP = (const float4 *)C.tabPT;
        const float4 *__restrict__ qD = (const float4 *)C.tabDT;
        const float4 *__restrict__ qQ = (const float4 *)C.tabQT;
        const long ntab = (long)(C.T + 1) * 2 * C.nhp * C.kp;   // the fifth slices follow
        const float *__restrict__ rP = (const float *)(qP + ntab);
        const float *__restrict__ rD = (const float *)(qD + ntab);
        const float *__restrict__ rQ = (const float *)(qQ + ntab);
        const long xo = (long)(m * 2 + par) * C.kp + tile * 16 + 4 * kq;
        const bool scal = J.kind == ANA_SCALAR;
        const Exp4 eP = load_exp4(C.expPT + xo), eD = load_exp4(C.expDT + xo), eQ = load_exp4(C.expQT + xo);
        if (scal) {
            const long cb = (long)(item >> 3) * mstride + mo + par * 16 + (item & 7) * 2 + part;
            for (int p0 = 0; p0 < C.nhp; p0 += 4) {
                const int pair = p0 + kq;
                const long ao = abase + (long)pair * C.kp;
                const float4 a = ld_slices(qP, ao, rv);
                const float a5 = ld_slice5(rP, ao, rv);
                const double b = ld_masked(four, cb + (long)pair * 32, cv);
                acc += split_step(a, a5, eP, b);
            }
        } else {
            const int gt = item * gpi, gp = gt + 1, ge = gt + 2;
            const int eoD = 1 - par, eoQ = par;
            const long bt = (long)(gt >> 3) * mstride + mo + (gt & 7) * 2;
            const long bp = (long)(gp >> 3) * mstride + mo + (gp & 7) * 2;
            const long be = (long)(ge >> 3) * mstride + mo + (ge & 7) * 2;
            const long offD = (comp == 0 ? bt : bp) + eoD * 16 + part;
            const double sD = comp == 0 ? 1.0 : -1.0;
            const long offQ = (comp == 0 ? bp : bt) + eoQ * 16 + (1 - part);
            const double sQ = part ? -1.0 : 1.0;
            const long offE = be + par * 16 + part;
            const bool useE = has_ek && comp == 0;
            for (int p0 = 0; p0 < C.nhp; p0 += 4) {
                const int pair = p0 + kq;
                const long po = (long)pair * 32, ao = abase + (long)pair * C.kp;
                const float4 aD = ld_slices(qD, ao, rv), aQ = ld_slices(qQ, ao, rv);
                const float aD5 = ld_slice5(rD, ao, rv), aQ5 = ld_slice5(rQ, ao, rv);
                const double bD = sD * ld_masked(four, offD + po, cv);
                const double bQ = sQ * ld_masked(four, offQ + po, cv);
                acc += split_step(aD, aD5, eD, bD);
                acc += split_step(aQ, aQ5, eQ, bQ);
                if (has_ek) {
                    const float4 aP = ld_slices(qP, ao, rv);
                    const float aP5 = ld_slice5(rP, ao, rv);
                    const double bE = ld_masked(four, offE + po, useE && cv);
                    acc2 += split_step(aP, aP5, eP, bE);
                }
            }
        }
    } else
    if (J.kind == ANA_SCALAR) {
        const long cb = (long)(item >> 3) * mstride + mo + par * 16 + (item & 7) * 2 + part;
        for (int p0 = 0; p0 < C.nhp; p0 += 4 * KCH) {
            T a[KCH], b[KCH];
#pragma unroll
            for (int u = 0; u < KCH; ++u) {
                const int pair = p0 + 4 * u + kq;
                const bool pv = pair < C.nhp;
                a[u] = (rv && pv) ? tP[abase + (long)pair * C.kp] : T(0);
                b[u] = (cv && pv) ? four[cb + (long)pair * 32] : T(0);
            }
#pragma unroll
            for (int u = 0; u < KCH; ++u) {
                if (p0 + 4 * u < C.nhp) acc = P::mfma(a[u], b[u], acc);
            }
        }
    } else {
        // S = D*Gt - i Q*Gp ; T = -i Q*Gt - D*Gp (times 1/(l(l+1)) in the epilogue), with
        // -i z = (Im z, -Re z): every B operand is one intermediate element times a sign.
        const int gt = item * gpi, gp = gt + 1, ge = gt + 2;
        const int eoD = 1 - par, eoQ = par;
        const long bt = (long)(gt >> 3) * mstride + mo + (gt & 7) * 2;
        const long bp = (long)(gp >> 3) * mstride + mo + (gp & 7) * 2;
        const long be = (long)(ge >> 3) * mstride + mo + (ge & 7) * 2;
        const long offD = (comp == 0 ? bt : bp) + eoD * 16 + part;
        const T sD = comp == 0 ? T(1) : T(-1);
        const long offQ = (comp == 0 ? bp : bt) + eoQ * 16 + (1 - part);
        const T sQ = part ? T(-1) : T(1);
        const long offE = be + par * 16 + part;
        const bool useE = has_ek && comp == 0;
        for (int p0 = 0; p0 < C.nhp; p0 += 4 * KCH) {
            T aD[KCH], aQ[KCH], aP[KCH], bD[KCH], bQ[KCH], bE[KCH];
#pragma unroll
            for (int u = 0; u < KCH; ++u) {
                const int pair = p0 + 4 * u + kq;
                const bool pv = pair < C.nhp;
                const long po = (long)pair * 32;
                aD[u] = (rv && pv) ? tD[abase + (long)pair * C.kp] : T(0);
                aQ[u] = (rv && pv) ? tQ[abase + (long)pair * C.kp] : T(0);
                aP[u] = (has_ek && rv && pv) ? tP[abase + (long)pair * C.kp] : T(0);
                bD[u] = (cv && pv) ? sD * four[offD + po] : T(0);
                bQ[u] = (cv && pv) ? sQ * four[offQ + po] : T(0);
                bE[u] = (useE && cv && pv) ? four[offE + po] : T(0);
            }
#pragma unroll
            for (int u = 0; u < KCH; ++u) {
                if (p0 + 4 * u < C.nhp) {
                    acc = P::mfma(aD[u], bD[u], acc);
                    acc = P::mfma(aQ[u], bQ[u], acc);
                    if (has_ek) acc2 = P::mfma(aP[u], bE[u], acc2);
                }
            }
        }
    }
    if (!cv) return;
    double *out = (J.kind == ANA_VECTOR && comp == 1) ? J.sout1 : J.sout0;
    if (!out) return;   // a vector job may leave its T output out (adjoint of the gradient)
    out += (long)item * J.sstride;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int kr = tile * 16 + P::row(kq, r);
        if (kr >= nk) continue;
        const int l = m + par + 2 * kr;
        const int lm = lmb + (l - m);
        double v = fav[r] * (double)acc[r];
        if (has_ek && comp == 0) v += fev[r] * (double)acc2[r];
        if (J.accumulate) out[2 * lm + part] += v;
        else out[2 * lm + part] = v;
    }
}

template <int PM, bool PLAN>
__global__ __launch_bounds__(256) void k_leg_anal(const LaunchArgs A)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + wave;
    if (task >= A.ntask) return;
    const Job &J = job_of<PLAN>(A, task, true);
    if (task - J.ltask0 >= J.ltasks) return;
    if constexpr (PM == PM_MIXED) {
        if (J.prec == 1) leg_anal_body<1>(A, J, task, lane);
        else if (J.prec == 2) leg_anal_body<2>(A, J, task, lane);
        else leg_anal_body<0>(A, J, task, lane);
    } else {
        leg_anal_body<PM>(A, J, task, lane);
    }
}

// ------------------------------------------------------------------ FFT
// Stockham autosort FFT (the formulation of Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC'08): pass p with radix R
// and Ns = product of earlier radices reads x[j + r N/R], multiplies by
// e^{sign 2 pi i r (j mod Ns)/(Ns R)}, does a radix-R DFT and writes to
// (j div Ns) Ns R + (j mod Ns) + r Ns. Natural order in and out, ping-pong in LDS.
//
// The grid lengths of the standard truncations (64, 96, 128, 160, 192, 256, 320)
// have compile-time radix plans (8 first, then 4, 2, 3, 5) and take their twiddles
// from an LDS copy of e^{-2 pi i k/N}; any other 2/3/5-smooth length runs the
// runtime-plan path.

template <class T> __device__ inline cx<T> cmul(cx<T> a, cx<T> b) { return cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <class T> __device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return cx<T>{a.x + b.x, a.y + b.y}; }
template <class T> __device__ inline cx<T> csub(cx<T> a, cx<T> b) { return cx<T>{a.x - b.x, a.y - b.y}; }
// s * i * z for s = +-1
template <int S, class T> __device__ inline cx<T> sixz(cx<T> z) { return S > 0 ? cx<T>{-z.y, z.x} : cx<T>{z.y, -z.x}; }

template <int S, class T> __device__ inline void dft2(cx<T> *v)
{
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <int S, class T> __device__ inline void dft4(cx<T> *v)
{
    cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]), c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, sixz<S>(d));
    v[3] = csub(b, sixz<S>(d));
}
template <int S, class T> __device__ inline void dft8(cx<T> *v)
{
    cx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft4<S>(e);
    dft4<S>(o);
    const T h = T(0.70710678118654752440084436210484904);  // cos(pi/4)
    // w^1 = h(1 + S i), w^2 = S i, w^3 = h(-1 + S i) with w = e^{S 2 pi i/8}
    const cx<T> t1 = cx<T>{h * (o[1].x - S * o[1].y), h * (o[1].y + S * o[1].x)};
    const cx<T> t2 = sixz<S>(o[2]);
    const cx<T> t3 = cx<T>{h * (-o[3].x - S * o[3].y), h * (-o[3].y + S * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
}
template <int S, class T> __device__ inline void dft3(cx<T> *v)
{
    using T2 = cx<T>;
    const T h = T(0.86602540378443864676372317075293618);  // sin(2 pi/3)
    T2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    T2 m = T2{v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
    T2 r = sixz<S>(T2{h * d.x, h * d.y});
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}
template <int S, class T> __device__ inline void dft5(cx<T> *v)
{
    using T2 = cx<T>;
    const T c1 = T(0.30901699437494742410229341718281906);   // cos(2 pi/5)
    const T c2 = T(-0.80901699437494742410229341718281906);  // cos(4 pi/5)
    const T s1 = T(0.95105651629515357211643933337938214);   // sin(2 pi/5)
    const T s2 = T(0.58778525229247312916870595463907277);   // sin(4 pi/5)
    T2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]), a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    T2 x0 = v[0];
    T2 r1 = T2{x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    T2 r2 = T2{x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    T2 i1 = sixz<S>(T2{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
    T2 i2 = sixz<S>(T2{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
    v[0] = T2{x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
}
template <int S, int R, class T> __device__ inline void dft(cx<T> *v)
{
    if constexpr (R == 2) dft2<S>(v);
    else if constexpr (R == 3) dft3<S>(v);
    else if constexpr (R == 4) dft4<S>(v);
    else if constexpr (R == 5) dft5<S>(v);
    else dft8<S>(v);
}

struct RadixPlan {
    int npass;
    int r[AER_MAXPASS];
};
__host__ __device__ constexpr RadixPlan radix_plan(int n)
{
    RadixPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    while (n % 8 == 0 && p.npass < AER_MAXPASS) { p.r[p.npass++] = 8; n /= 8; }
    while (n % 4 == 0 && p.npass < AER_MAXPASS) { p.r[p.npass++] = 4; n /= 4; }
    while (n % 2 == 0 && p.npass < AER_MAXPASS) { p.r[p.npass++] = 2; n /= 2; }
    while (n % 3 == 0 && p.npass < AER_MAXPASS) { p.r[p.npass++] = 3; n /= 3; }
    while (n % 5 == 0 && p.npass < AER_MAXPASS) { p.r[p.npass++] = 5; n /= 5; }
    return p;
}
__host__ __device__ constexpr int ns_before(int n, int pass)
{
    const RadixPlan p = radix_plan(n);
    int s = 1;
    for (int i = 0; i < pass; ++i) s *= p.r[i];
    return s;
}

// Grid lengths with a compile-time FFT path, and the rows (complex sequences =
// latitude-pair x field) one block of AER_FFT_THREADS transforms.
__host__ __device__ constexpr bool fft_fixed(int n)
{
    return n == 64 || n == 96 || n == 128 || n == 160 || n == 192 || n == 256 || n == 320;
}
// Rows of a fixed-length FFT block: 16 at nlon 64, 8 otherwise. A launch with few rows in
// all (one model's transform set) takes blocks of 4 rows at nlon 96 and 128 instead, so
// the launch has twice the blocks and each a shorter chain (fft_small_rows).
__host__ __device__ constexpr int fft_fixed_rows(int n) { return n == 64 ? 16 : 8; }
__host__ __device__ constexpr bool fft_small_rows_ok(int n) { return n == 96 || n == 128; }
constexpr int AER_FFT_SMALL_ROWS = 4;
constexpr long AER_FFT_SMALL_LAUNCH = 16384;   // launches with at most this many rows take 4-row blocks

// Row rp of a fixed-length FFT block. ROWS >= 8: the group's 8 fields x ROWS/8 latitude
// pairs, field fastest; ROWS < 8: ROWS fields (the sub-th slice of the group) x one pair.
template <int ROWS>
__device__ __forceinline__ void fixed_row(int rp, int sub, int p0, int &gfl, int &pair)
{
    if constexpr (ROWS >= 8) {
        gfl = rp & 7;
        pair = p0 + (rp >> 3);
    } else {
        gfl = sub * ROWS + rp;
        pair = p0;
    }
}

// LDS index with one padding element per 16: the Stockham stores of a radix-R pass
// stride R elements across lanes, which without padding hits a few banks only.
__device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }
template <int ROWS, int N> constexpr int padded_len() { return ROWS * N + (ROWS * N) / 16 + 1; }

template <int S, class T, int N, int ROWS, int P, int NT = AER_FFT_THREADS>
__device__ inline void fft_pass_fixed(const cx<T> *__restrict__ src, cx<T> *__restrict__ dst,
                                      const cx<T> *__restrict__ tw)
{
    constexpr RadixPlan PL = radix_plan(N);
    constexpr int R = PL.r[P];
    constexpr int Ns = ns_before(N, P);
    constexpr int NR = N / R;
    constexpr int TS = N / (Ns * R);
    constexpr int TOT = ROWS * NR;
#pragma unroll
    for (int it = 0; it < (TOT + NT - 1) / NT; ++it) {
        const int bi = threadIdx.x + it * NT;
        if (TOT % NT != 0 && bi >= TOT) break;
        const int rp = bi / NR, j = bi - rp * NR;
        const int k = j % Ns;
        cx<T> v[R];
        const int sb = rp * N + j;
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[pidx(sb + r * NR)];
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cx<T> w = tw[r * k * TS];
                if (S > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft<S, R>(v);
        const int db = rp * N + (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[pidx(db + r * Ns)] = v[r];
    }
}

template <int S, class T, int N, int ROWS, int NT = AER_FFT_THREADS, int P = 0>
__device__ inline cx<T> *fft_block_fixed(cx<T> *a, cx<T> *b, const cx<T> *tw)
{
    constexpr RadixPlan PL = radix_plan(N);
    if constexpr (P >= PL.npass) {
        return a;
    } else {
        fft_pass_fixed<S, T, N, ROWS, P, NT>(a, b, tw);
        __syncthreads();
        return fft_block_fixed<S, T, N, ROWS, NT, P + 1>(b, a, tw);
    }
}

// Runtime-plan path: rows complex sequences of length C.nlon. Returns the result buffer.
template <int S, class T>
__device__ __forceinline__ cx<T> *fft_rows(cx<T> *a, cx<T> *b, int rows, const CfgDev &C)
{
    using T2 = cx<T>;
    const int N = C.nlon;
    const T2 *tw = (const T2 *)C.tw;
    int Ns = 1;
    for (int p = 0; p < C.npass; ++p) {
        const int R = C.radix[p], NR = N / R;
        const T2 *twp = tw + C.twoff[p];
        for (int bi = threadIdx.x; bi < rows * NR; bi += blockDim.x) {
            const int rp = bi / NR, j = bi - rp * NR;
            const T2 *src = a + rp * N;
            T2 *dst = b + rp * N;
            const int k = j % Ns;
            T2 v[8];
            for (int r = 0; r < R; ++r) v[r] = src[j + r * NR];
            for (int r = 1; r < R; ++r) {
                T2 w = twp[k * (R - 1) + r - 1];
                if (S > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
            switch (R) {
                case 2: dft2<S>(v); break;
                case 3: dft3<S>(v); break;
                case 4: dft4<S>(v); break;
                case 5: dft5<S>(v); break;
                default: dft8<S>(v); break;
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
        }
        __syncthreads();
        T2 *tmp = a; a = b; b = tmp;
        Ns *= R;
    }
    return a;
}

// Where a synthesis grid field goes: output pointer and scale.
__device__ __forceinline__ double *syn_out(const Job &J, int gf, double &sc)
{
    if (J.kind == SYN_SCALAR) { sc = J.gsc0; return J.gout0 + (long)gf * J.gstride; }
    const int comp = gf & 1;
    sc = comp ? J.gsc1 : J.gsc0;
    return (comp ? J.gout1 : J.gout0) + (long)(gf >> 1) * J.gstride;
}

// Where an analysis grid field comes from.
__device__ __forceinline__ const double *ana_in(const Job &J, int gf, double &sc)
{
    const int gpi = (J.kind == ANA_VECTOR) ? (J.gin2 ? 3 : 2) : 1;
    const int comp = gf % gpi, item = gf / gpi;
    sc = comp == 0 ? J.gsc0 : (comp == 1 ? J.gsc1 : 1.0);
    return (comp == 0 ? J.gin0 : (comp == 1 ? J.gin1 : J.gin2)) + (long)item * J.gstride;
}

// Z_m = F_N + i F_S for 0 < m <= M, Z_{N-m} = conj(F_N) + i conj(F_S), Z_0 = Re F_N + i Re F_S,
// with F_N = E + O, F_S = E - O read from the intermediate.
template <class T>
__device__ __forceinline__ void synth_pack(T er, T ei, T or_, T oi, int m, cx<T> &zm, cx<T> &zc)
{
    const T nr = er + or_, ni = ei + oi, sr = er - or_, si = ei - oi;
    if (m == 0) {
        zm = cx<T>{nr, sr};
        zc = zm;
    } else {
        zm = cx<T>{nr - si, ni + sr};
        zc = cx<T>{nr + si, sr - ni};
    }
}

// G_N = (Z_m + conj Z_{N-m})/2, G_S = -i (Z_m - conj Z_{N-m})/2; E = w (G_N + G_S), O = w (G_N - G_S).
template <class T>
__device__ inline void anal_unpack(cx<T> zm, cx<T> zc, T w, T *__restrict__ out)
{
    const T gnr = T(0.5) * (zm.x + zc.x), gni = T(0.5) * (zm.y - zc.y);
    const T gsr = T(0.5) * (zm.y + zc.y), gsi = -T(0.5) * (zm.x - zc.x);
    out[0] = w * (gnr + gsr);
    out[1] = w * (gni + gsi);
    out[16] = w * (gnr - gsr);
    out[17] = w * (gni - gsi);
}

// ---- synthesis FFT, compile-time length: block = 8 grid fields x ROWS/8 latitude pairs,
// row rp = pair_local * 8 + field_local.
template <class T, int N, int ROWS>
__device__ __forceinline__ void fft_synth_fixed(const Job &J, const CfgDev &C, int g, int sub, int p0, cx<T> *smem,
                                                const T *__restrict__ four)
{
    constexpr int NT = AER_FFT_THREADS;
    constexpr int PL = padded_len<ROWS, N>();
    using T2 = cx<T>;
    T2 *bufA = smem, *bufB = smem + PL, *tw = smem + 2 * PL;
    const T2 *twg = (const T2 *)C.twN;
    const int M = C.T;
    const int adj = J.adj;
    const double *wpair = C.wpair;
    // Every global load of the block is issued before the first wait: the
    // intermediate values (T + 1 <= N/2 since the grid has nlon > 2T) and the
    // twiddles go to registers, then to LDS.
    constexpr int NLD = (ROWS * (N / 2) + NT - 1) / NT, NTW = (N + NT - 1) / NT;
    T er[NLD], ei[NLD], orr[NLD], oi[NLD];
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
        const int idx = threadIdx.x + u * NT;
        const int rp = idx % ROWS, m = idx / ROWS;
        int gfl, pair;
        fixed_row<ROWS>(rp, sub, p0, gfl, pair);
        const bool v = m <= M && pair < C.nh && g * 8 + gfl < J.ngf;
        const T *b = four + ((long)(v ? m : 0) * C.nhp + (v ? pair : 0)) * 32 + gfl * 2;
        er[u] = v ? b[0] : T(0);
        ei[u] = v ? b[1] : T(0);
        orr[u] = v ? b[16] : T(0);
        oi[u] = v ? b[17] : T(0);
    }
    T2 twv[NTW];
#pragma unroll
    for (int u = 0; u < NTW; ++u) {
        const int k = threadIdx.x + u * NT;
        twv[u] = k < N ? twg[k] : T2{0, 0};
    }
    const int nz = N - 2 * M - 1;
    for (int idx = threadIdx.x; idx < ROWS * nz; idx += NT) {
        const int rp = idx % ROWS, k = M + 1 + idx / ROWS;
        bufA[pidx(rp * N + k)] = T2{0, 0};
    }
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
        const int idx = threadIdx.x + u * NT;
        const int rp = idx % ROWS, m = idx / ROWS;
        if (m <= M) {
            T2 zm, zc;
            const T hf = (adj && m > 0) ? T(0.5) : T(1);
            synth_pack(hf * er[u], hf * ei[u], hf * orr[u], hf * oi[u], m, zm, zc);
            bufA[pidx(rp * N + m)] = zm;
            if (m > 0) bufA[pidx(rp * N + N - m)] = zc;
        }
    }
#pragma unroll
    for (int u = 0; u < NTW; ++u) {
        const int k = threadIdx.x + u * NT;
        if (k < N) tw[k] = twv[u];
    }
    __syncthreads();
    const T2 *res = fft_block_fixed<+1, T, N, ROWS>(bufA, bufB, tw);
    for (int idx = threadIdx.x; idx < ROWS * N; idx += NT) {
        const int rp = idx / N, i = idx - rp * N;
        int gfl, pair;
        fixed_row<ROWS>(rp, sub, p0, gfl, pair);
        const int gf = g * 8 + gfl;
        if (pair >= C.nh || gf >= J.ngf) continue;
        double sc;
        double *out = syn_out(J, gf, sc);
        const T2 z = res[pidx(idx)];
        const int js = C.nlat - 1 - pair;
        if (adj) sc *= wpair[pair] * (js == pair ? 2.0 : 1.0);   // w_j 2pi/nlon (the equator's half undone)
        out[(long)pair * N + i] = sc * (double)z.x;
        if (js != pair) out[(long)js * N + i] = sc * (double)z.y;
    }
}

template <class T, int N, int ROWS>
__device__ __forceinline__ void fft_anal_fixed(const Job &J, const CfgDev &C, int g, int sub, int p0, cx<T> *smem,
                                               T *__restrict__ four)
{
    constexpr int NT = AER_FFT_THREADS;
    constexpr int PL = padded_len<ROWS, N>();
    using T2 = cx<T>;
    T2 *bufA = smem, *bufB = smem + PL, *tw = smem + 2 * PL;
    const T2 *twg = (const T2 *)C.twN;
    // All grid loads of the block are issued before the first wait (registers, then LDS).
    static_assert((ROWS * N) % NT == 0, "fixed FFT rows must tile the block");
    constexpr int NLD = ROWS * N / NT, NTW = (N + NT - 1) / NT;
    T2 z[NLD];
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
        const int idx = threadIdx.x + u * NT;
        const int rp = idx / N, i = idx - rp * N;
        int gfl, pair;
        fixed_row<ROWS>(rp, sub, p0, gfl, pair);
        const int gf = g * 8 + gfl;
        const bool v = pair < C.nh && gf < J.ngf;
        double sc = 0.0;
        const double *in = ana_in(J, v ? gf : 0, sc);
        const int pn = v ? pair : 0;
        z[u] = v ? T2{T(sc * in[(long)pn * N + i]), T(sc * in[(long)(C.nlat - 1 - pn) * N + i])} : T2{0, 0};
    }
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
    // job fields read before the global stores below (kernel-argument memory is
    // then never re-read after a store that might alias it)
    const int M = C.T, nh = C.nh, nhp = C.nhp, ngf = J.ngf, adj = J.adj, nlat = C.nlat;
    // the block's one or two latitude-pair weights, loaded before the FFT so that no
    // global load waits after it
    static_assert(ROWS <= 16, "a fixed FFT block spans at most two latitude pairs");
    const double wp0 = p0 < nh ? C.wpair[p0] : 0.0;
    const double wp1 = (ROWS > 8 && p0 + 1 < nh) ? C.wpair[p0 + 1] : 0.0;
    const T2 *res = fft_block_fixed<-1, T, N, ROWS>(bufA, bufB, tw);
    for (int idx = threadIdx.x; idx < ROWS * (M + 1); idx += NT) {
        const int rp = idx % ROWS, m = idx / ROWS;
        int gfl, pair;
        fixed_row<ROWS>(rp, sub, p0, gfl, pair);
        if (pair >= nhp) continue;
        T *o = four + ((long)m * nhp + pair) * 32 + gfl * 2;
        if (pair < nh && g * 8 + gfl < ngf) {
            const double w = adj ? ((nlat - 1 - pair == pair) ? 0.5 : 1.0) * (m > 0 ? 2.0 : 1.0)
                                 : ((rp >> 3) ? wp1 : wp0);
            anal_unpack(res[pidx(rp * N + m)], res[pidx(rp * N + (m == 0 ? 0 : N - m))], T(w), o);
        } else {
            o[0] = o[1] = o[16] = o[17] = T(0);
        }
    }
}

// ---- runtime-length paths (block = fb fields x pb pairs, row rp = field_local * pb + pair_local)
template <class T>
__device__ __forceinline__ void fft_synth_generic(const Job &J, const CfgDev &C, int g, int sub, int p0, cx<T> *smem, const T *__restrict__ four)
{
    using T2 = cx<T>;
    const int N = C.nlon, M = C.T, rows = C.fb * C.pb;
    T2 *bufA = smem, *bufB = smem + rows * N;
    for (int idx = threadIdx.x; idx < rows * N; idx += blockDim.x) {
        const int rp = idx / N, k = idx - rp * N;
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb;
        T2 z{0, 0};
        if (pair < C.nh && g * 8 + gfl < J.ngf) {
            int m = -1;
            if (k <= M) m = k;
            else if (k >= N - M) m = N - k;
            if (m >= 0) {
                T2 zm, zc;
                const T *b = four + ((long)m * C.nhp + pair) * 32 + gfl * 2;
                const T hf = (J.adj && m > 0) ? T(0.5) : T(1);
                synth_pack(hf * b[0], hf * b[1], hf * b[16], hf * b[17], m, zm, zc);
                z = (k <= M) ? zm : zc;
            }
        }
        bufA[idx] = z;
    }
    __syncthreads();
    const T2 *res = fft_rows<+1, T>(bufA, bufB, rows, C);
    const int adj = J.adj;
    const double *wpair = C.wpair;
    for (int idx = threadIdx.x; idx < rows * N; idx += blockDim.x) {
        const int rp = idx / N, i = idx - rp * N;
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb, gf = g * 8 + gfl;
        if (pair >= C.nh || gf >= J.ngf) continue;
        double sc;
        double *out = syn_out(J, gf, sc);
        const T2 z = res[idx];
        const int js = C.nlat - 1 - pair;
        if (adj) sc *= wpair[pair] * (js == pair ? 2.0 : 1.0);
        out[(long)pair * N + i] = sc * (double)z.x;
        if (js != pair) out[(long)js * N + i] = sc * (double)z.y;
    }
}

template <class T>
__device__ __forceinline__ void fft_anal_generic(const Job &J, const CfgDev &C, int g, int sub, int p0, cx<T> *smem, T *__restrict__ four)
{
    using T2 = cx<T>;
    const int N = C.nlon, M = C.T, rows = C.fb * C.pb;
    T2 *bufA = smem, *bufB = smem + rows * N;
    for (int idx = threadIdx.x; idx < rows * N; idx += blockDim.x) {
        const int rp = idx / N, i = idx - rp * N;
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb, gf = g * 8 + gfl;
        T2 z{0, 0};
        if (pair < C.nh && gf < J.ngf) {
            double sc;
            const double *in = ana_in(J, gf, sc);
            z = T2{T(sc * in[(long)pair * N + i]), T(sc * in[(long)(C.nlat - 1 - pair) * N + i])};
        }
        bufA[idx] = z;
    }
    __syncthreads();
    const T2 *res = fft_rows<-1, T>(bufA, bufB, rows, C);
    const int nh = C.nh, nhp = C.nhp, ngf = J.ngf, fbk = C.fb, pbk = C.pb, adj = J.adj, nlat = C.nlat;
    const double *wpair = C.wpair;
    for (int idx = threadIdx.x; idx < rows * (M + 1); idx += blockDim.x) {
        const int rp = idx / (M + 1), m = idx - rp * (M + 1);
        const int gfl = sub * fbk + rp / pbk, pair = p0 + rp % pbk;
        if (pair >= nhp) continue;
        T *o = four + ((long)m * nhp + pair) * 32 + gfl * 2;
        if (pair < nh && g * 8 + gfl < ngf)
            anal_unpack(res[rp * N + m], res[rp * N + (m == 0 ? 0 : N - m)],
                        T(adj ? ((nlat - 1 - pair == pair) ? 0.5 : 1.0) * (m > 0 ? 2.0 : 1.0) : wpair[pair]), o);
        else
            o[0] = o[1] = o[16] = o[17] = T(0);
    }
}

// Block -> (job, intermediate group, field subgroup, first latitude pair).
template <bool PLAN>
__device__ __forceinline__ const Job &fft_task(const LaunchArgs &A, int task, int &g, int &sub, int &p0)
{
    const Job &J = job_of<PLAN>(A, task, false);
    const int nsub = 8 / J.c.fb;
    int t = task - J.ftask0;
    g = t % J.ngroup;
    t /= J.ngroup;
    sub = t % nsub;
    p0 = (t / nsub) * J.c.pb;
    return J;
}

#define AER_FFT_DISPATCH(BODY, ...)                                                        \
    switch (C.nlon) {                                                                      \
        case 64: BODY##_fixed<T, 64, fft_fixed_rows(64)>(__VA_ARGS__); break;              \
        case 96:                                                                           \
            if (C.fb == AER_FFT_SMALL_ROWS) BODY##_fixed<T, 96, AER_FFT_SMALL_ROWS>(__VA_ARGS__); \
            else BODY##_fixed<T, 96, fft_fixed_rows(96)>(__VA_ARGS__);                     \
            break;                                                                         \
        case 128:                                                                          \
            if (C.fb == AER_FFT_SMALL_ROWS) BODY##_fixed<T, 128, AER_FFT_SMALL_ROWS>(__VA_ARGS__); \
            else BODY##_fixed<T, 128, fft_fixed_rows(128)>(__VA_ARGS__);                   \
            break;                                                                         \
        case 160: BODY##_fixed<T, 160, fft_fixed_rows(160)>(__VA_ARGS__); break;           \
        case 192: BODY##_fixed<T, 192, fft_fixed_rows(192)>(__VA_ARGS__); break;           \
        case 256: BODY##_fixed<T, 256, fft_fixed_rows(256)>(__VA_ARGS__); break;           \
        case 320: BODY##_fixed<T, 320, fft_fixed_rows(320)>(__VA_ARGS__); break;           \
        default: break;                                                                    \
    }

// GENERIC = false: only the compile-time lengths (no scratch); true: also the
// runtime-plan path (the host picks it when some job has another grid length).
template <int PREC, bool GENERIC>
__device__ __forceinline__ void fft_synth_body(const Job &J, int g, int sub, int p0, char *smem_raw, const void *four_base)
{
    using T = typename Prec<PREC>::T;
    const CfgDev &C = J.c;
    cx<T> *smem = (cx<T> *)smem_raw;
    const T *four = (const T *)((const char *)four_base + J.four_off) + (long)g * (C.T + 1) * C.nhp * 32;
    if (fft_fixed(C.nlon)) {
        AER_FFT_DISPATCH(fft_synth, J, C, g, sub, p0, smem, four)
    } else if constexpr (GENERIC) {
        fft_synth_generic<T>(J, C, g, sub, p0, smem, four);
    }
}

template <int PREC, bool GENERIC>
__device__ __forceinline__ void fft_anal_body(const Job &J, int g, int sub, int p0, char *smem_raw, void *four_base)
{
    using T = typename Prec<PREC>::T;
    const CfgDev &C = J.c;
    cx<T> *smem = (cx<T> *)smem_raw;
    T *four = (T *)((char *)four_base + J.four_off) + (long)g * (C.T + 1) * C.nhp * 32;
    if (fft_fixed(C.nlon)) {
        AER_FFT_DISPATCH(fft_anal, J, C, g, sub, p0, smem, four)
    } else if constexpr (GENERIC) {
        fft_anal_generic<T>(J, C, g, sub, p0, smem, four);
    }
}

template <int PM, bool GENERIC, bool PLAN>
__global__ __launch_bounds__(AER_FFT_THREADS) void k_fft_synth(const LaunchArgs A)
{
    extern __shared__ char smem_raw[];
    const int task = blockIdx.x;
    if (task >= A.ntask) return;
    int g, sub, p0;
    const Job &J = fft_task<PLAN>(A, task, g, sub, p0);
    if constexpr (PM == PM_MIXED) {
        if (J.prec == 1) fft_synth_body<1, GENERIC>(J, g, sub, p0, smem_raw, A.four);
        else fft_synth_body<0, GENERIC>(J, g, sub, p0, smem_raw, A.four);
    } else {
        fft_synth_body<PM == PM_FP32 ? 1 : 0, GENERIC>(J, g, sub, p0, smem_raw, A.four);
    }
}

template <int PM, bool GENERIC, bool PLAN>
__global__ __launch_bounds__(AER_FFT_THREADS) void k_fft_anal(const LaunchArgs A)
{
    extern __shared__ char smem_raw[];
    const int task = blockIdx.x;
    if (task >= A.ntask) return;
    int g, sub, p0;
    const Job &J = fft_task<PLAN>(A, task, g, sub, p0);
    if constexpr (PM == PM_MIXED) {
        if (J.prec == 1) fft_anal_body<1, GENERIC>(J, g, sub, p0, smem_raw, A.four);
        else fft_anal_body<0, GENERIC>(J, g, sub, p0, smem_raw, A.four);
    } else {
        fft_anal_body<PM == PM_FP32 ? 1 : 0, GENERIC>(J, g, sub, p0, smem_raw, A.four);
    }
}

}  // namespace aer
