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
    const void *tabP, *tabD, *tabQ;            // [nlm][nhp] in the compute type
    const double *lfac;                        // [AER_NMODE][T+1]
    const double *wpair;                       // [nhp] w_j * 2pi/nlon (x1/2 on a self-mirrored equator row)
    const void *tw;                            // forward twiddles, complex, compute type
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
    long four_off;                    // element offset of the job's intermediate region
};

struct LaunchArgs {
    int njob, ntask;
    void *four;
    Job job[AER_MAXJOB];
};

template <class T> struct __attribute__((aligned(2 * sizeof(T)))) cx {
    T x, y;
};

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <int PREC> struct Prec;
// Operand layouts of the 16x16x4 MFMAs (lane L): A[L%16][L/16], B[L/16][L%16] for
// both types. Accumulator element r of lane L holds D[row(L/16, r)][L%16], where
// row is 4r + L/16 for f64 and 4(L/16) + r for f32 (measured on gfx950).
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

__device__ inline int find_job(const LaunchArgs &A, int task, bool legendre)
{
    int jb = 0;
    while (jb + 1 < A.njob && task >= (legendre ? A.job[jb + 1].ltask0 : A.job[jb + 1].ftask0)) ++jb;
    return jb;
}

// 0-based index of (l = m, m) in the m-major, l-fastest layout: the orders m' < m
// hold sum (T+1-m') = m(2T+3-m)/2 coefficients (reference src/aeros_spectral.f90:478).
__device__ inline int lm_base(int T, int m) { return m * (2 * T + 3 - m) / 2; }

// ------------------------------------------------------------------ Legendre synthesis
// One wavefront per task (job, group of 8 grid fields, m, tile of 16 latitude pairs).
template <int PREC>
__global__ __launch_bounds__(256) void k_leg_synth(const LaunchArgs A)
{
    using P = Prec<PREC>;
    using T = typename P::T;
    using Acc = typename P::Acc;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + wave;
    if (task >= A.ntask) return;
    const Job &J = A.job[find_job(A, task, true)];
    const CfgDev &C = J.c;
    int t = task - J.ltask0;
    const int g = t % J.lgroups;
    t /= J.lgroups;
    const int pt = t % C.npt;
    const int m = t / C.npt;
    const int i = lane & 15, kq = lane >> 4;
    const int gf = g * 8 + (i >> 1), part = i & 1;
    const bool cv = gf < J.ngf;
    int item, comp;
    if (J.kind == SYN_SCALAR) { item = gf; comp = 0; }
    else { item = gf >> 1; comp = gf & 1; }
    const double *s0 = J.sin0 + (long)item * J.sstride;
    const double *s1 = (J.kind == SYN_VECTOR) ? J.sin1 + (long)item * J.sstride : nullptr;
    const int lmb = lm_base(C.T, m);
    const T *tP = (const T *)C.tabP, *tD = (const T *)C.tabD, *tQ = (const T *)C.tabQ;
    const double *f0 = C.lfac + J.mode0 * (C.T + 1), *f1 = C.lfac + J.mode1 * (C.T + 1);
    Acc E = {0, 0, 0, 0}, O = {0, 0, 0, 0};
    for (int par = 0; par < 2; ++par) {
        const int nk = (C.T - m - par >= 0) ? (C.T - m - par) / 2 + 1 : 0;
        for (int k0 = 0; k0 < nk; k0 += 4) {
            const int k = k0 + kq;
            const bool kv = k < nk;
            const int l = m + par + 2 * k;
            const int lm = lmb + (l - m);
            const long trow = (long)lm * C.nhp + pt * 16 + i;
            if (J.kind == SYN_SCALAR) {
                const T a = kv ? tP[trow] : T(0);
                const T b = (kv && cv) ? T(f0[l] * s0[2 * lm + part]) : T(0);
                if (par == 0) E = P::mfma(a, b, E);
                else O = P::mfma(a, b, O);
            } else {
                const T aD = kv ? tD[trow] : T(0);
                const T aQ = kv ? tQ[trow] : T(0);
                double sr = 0, si = 0, tr = 0, ti = 0;
                if (kv && cv) {
                    sr = f0[l] * s0[2 * lm];
                    si = f0[l] * s0[2 * lm + 1];
                    if (s1) {
                        tr = f1[l] * s1[2 * lm];
                        ti = f1[l] * s1[2 * lm + 1];
                    }
                }
                // V_theta = S*D + (iT)*Q ; V_phi = (iS)*Q + (-T)*D
                const double bD = comp == 0 ? (part ? si : sr) : (part ? -ti : -tr);
                const double bQ = comp == 0 ? (part ? tr : -ti) : (part ? sr : -si);
                if (par == 0) {
                    E = P::mfma(aQ, T(bQ), E);
                    O = P::mfma(aD, T(bD), O);
                } else {
                    O = P::mfma(aQ, T(bQ), O);
                    E = P::mfma(aD, T(bD), E);
                }
            }
        }
    }
    T *four = (T *)A.four + J.four_off;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int pair = pt * 16 + P::row(kq, r);
        const long base = (((long)g * (C.T + 1) + m) * C.nhp + pair) * 32;
        four[base + i] = E[r];
        four[base + 16 + i] = O[r];
    }
}

// ------------------------------------------------------------------ Legendre analysis
// One wavefront per task (job, column group, m, parity x tile of 16 degrees).
// Rows = degrees l of one parity, K = latitude pairs, columns = 8 complex outputs
// (scalar) or 4 (S, T) pairs (vector).
template <int PREC>
__global__ __launch_bounds__(256) void k_leg_anal(const LaunchArgs A)
{
    using P = Prec<PREC>;
    using T = typename P::T;
    using Acc = typename P::Acc;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + wave;
    if (task >= A.ntask) return;
    const Job &J = A.job[find_job(A, task, true)];
    const CfgDev &C = J.c;
    const int ntpp = ((C.T + 2) / 2 + 15) / 16;  // tiles per parity
    int t = task - J.ltask0;
    const int g = t % J.lgroups;
    t /= J.lgroups;
    const int tt = t % (2 * ntpp);
    const int m = t / (2 * ntpp);
    const int par = tt / ntpp, tile = tt % ntpp;
    const int nk = (C.T - m - par >= 0) ? (C.T - m - par) / 2 + 1 : 0;
    if (tile * 16 >= nk) return;
    const int i = lane & 15, kq = lane >> 4;
    const int lmb = lm_base(C.T, m);
    const int kra = tile * 16 + i;
    const bool rv = kra < nk;
    const long arow = (long)(lmb + par + 2 * kra) * C.nhp;
    const T *tP = (const T *)C.tabP, *tD = (const T *)C.tabD, *tQ = (const T *)C.tabQ;
    const T *four = (const T *)A.four + J.four_off;
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
    if (J.kind == ANA_SCALAR) {
        const long cb = (long)(item >> 3) * mstride + mo + par * 16 + (item & 7) * 2 + part;
        for (int p0 = 0; p0 < C.nhp; p0 += 4) {
            const int pair = p0 + kq;
            const T a = rv ? tP[arow + pair] : T(0);
            const T b = cv ? four[cb + (long)pair * 32] : T(0);
            acc = P::mfma(a, b, acc);
        }
    } else {
        const int gt = item * gpi, gp = gt + 1, ge = gt + 2;
        const int eoD = 1 - par, eoQ = par;
        const long bt = (long)(gt >> 3) * mstride + mo + (gt & 7) * 2;
        const long bp = (long)(gp >> 3) * mstride + mo + (gp & 7) * 2;
        const long be = (long)(ge >> 3) * mstride + mo + (ge & 7) * 2;
        for (int p0 = 0; p0 < C.nhp; p0 += 4) {
            const int pair = p0 + kq;
            const long po = (long)pair * 32;
            const T aD = rv ? tD[arow + pair] : T(0);
            const T aQ = rv ? tQ[arow + pair] : T(0);
            T bD = 0, bQ = 0, bE = 0;
            if (cv) {
                // S = D*Gt - i Q*Gp ; T = -i Q*Gt - D*Gp   (times 1/(l(l+1)) in the epilogue)
                if (comp == 0) {
                    bD = four[bt + po + eoD * 16 + part];
                    bQ = part ? -four[bp + po + eoQ * 16] : four[bp + po + eoQ * 16 + 1];
                    if (has_ek) bE = four[be + po + par * 16 + part];
                } else {
                    bD = -four[bp + po + eoD * 16 + part];
                    bQ = part ? -four[bt + po + eoQ * 16] : four[bt + po + eoQ * 16 + 1];
                }
            }
            acc = P::mfma(aD, bD, acc);
            acc = P::mfma(aQ, bQ, acc);
            if (has_ek) {
                const T aP = rv ? tP[arow + pair] : T(0);
                acc2 = P::mfma(aP, bE, acc2);
            }
        }
    }
    if (!cv) return;
    double *out = (J.kind == ANA_VECTOR && comp == 1) ? J.sout1 : J.sout0;
    out += (long)item * J.sstride;
    const double *fa = C.lfac + ((J.kind == ANA_VECTOR && comp == 1) ? J.mode1 : J.mode0) * (C.T + 1);
    const double *fe = C.lfac + J.mode2 * (C.T + 1);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int kr = tile * 16 + P::row(kq, r);
        if (kr >= nk) continue;
        const int l = m + par + 2 * kr;
        const int lm = lmb + (l - m);
        double v = fa[l] * (double)acc[r];
        if (has_ek && comp == 0) v += fe[l] * (double)acc2[r];
        if (J.accumulate) out[2 * lm + part] += v;
        else out[2 * lm + part] = v;
    }
}

// ------------------------------------------------------------------ FFT
// Stockham autosort FFT (the formulation of Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC'08): pass p with radix R
// and Ns = product of earlier radices reads x[j + r N/R], multiplies by
// e^{sign 2 pi i r (j mod Ns)/(Ns R)}, does a radix-R DFT and writes to
// (j div Ns) Ns R + (j mod Ns) + r Ns. Natural order in and out, ping-pong in LDS.

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

// rows complex sequences of length N, row r at a + r*N. Returns the buffer holding the result.
template <int S, class T>
__device__ cx<T> *fft_rows(cx<T> *a, cx<T> *b, int rows, const CfgDev &C)
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
            T2 v[5];
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
                default: dft5<S>(v); break;
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

// Synthesis FFT: one block per (job, group of 8 grid fields, fb of them, block of pb latitude pairs).
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_synth(const LaunchArgs A)
{
    using P = Prec<PREC>;
    using T = typename P::T;
    using T2 = typename P::T2;
    extern __shared__ char smem_raw[];
    const int task = blockIdx.x;
    if (task >= A.ntask) return;
    const Job &J = A.job[find_job(A, task, false)];
    const CfgDev &C = J.c;
    const int nsub = 8 / C.fb;
    int t = task - J.ftask0;
    const int g = t % J.ngroup;
    t /= J.ngroup;
    const int sub = t % nsub;
    const int p0 = (t / nsub) * C.pb;
    const int N = C.nlon, M = C.T, rows = C.fb * C.pb;
    T2 *bufA = (T2 *)smem_raw, *bufB = bufA + rows * N;
    const T *four = (const T *)A.four + J.four_off + (long)g * (M + 1) * C.nhp * 32;
    for (int idx = threadIdx.x; idx < rows * N; idx += blockDim.x) {
        const int rp = idx / N, k = idx - rp * N;
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb;
        T2 z = T2{0, 0};
        if (pair < C.nh && g * 8 + gfl < J.ngf) {
            int m = -1;
            bool cj = false;
            if (k <= M) m = k;
            else if (k >= N - M) { m = N - k; cj = true; }
            if (m >= 0) {
                const long b = ((long)m * C.nhp + pair) * 32 + gfl * 2;
                const T er = four[b], ei = four[b + 1], or_ = four[b + 16], oi = four[b + 17];
                const T nr = er + or_, ni = ei + oi, sr = er - or_, si = ei - oi;
                if (k == 0) z = T2{nr, sr};
                else if (!cj) z = T2{nr - si, ni + sr};   // F_N + i F_S
                else z = T2{nr + si, sr - ni};            // conj(F_N) + i conj(F_S)
            }
        }
        bufA[idx] = z;
    }
    __syncthreads();
    const T2 *res = fft_rows<+1, T>(bufA, bufB, rows, C);
    for (int idx = threadIdx.x; idx < rows * N; idx += blockDim.x) {
        const int rp = idx / N, i = idx - rp * N;
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb;
        const int gf = g * 8 + gfl;
        if (pair >= C.nh || gf >= J.ngf) continue;
        double *out;
        double sc;
        if (J.kind == SYN_SCALAR) { out = J.gout0 + (long)gf * J.gstride; sc = J.gsc0; }
        else {
            const int comp = gf & 1;
            out = (comp ? J.gout1 : J.gout0) + (long)(gf >> 1) * J.gstride;
            sc = comp ? J.gsc1 : J.gsc0;
        }
        const T2 z = res[idx];
        const int js = C.nlat - 1 - pair;
        out[(long)pair * N + i] = sc * (double)z.x;
        if (js != pair) out[(long)js * N + i] = sc * (double)z.y;
    }
}

// Analysis FFT: rows north + i south, forward transform, unpack, fold to E/O, weight.
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_anal(const LaunchArgs A)
{
    using P = Prec<PREC>;
    using T = typename P::T;
    using T2 = typename P::T2;
    extern __shared__ char smem_raw[];
    const int task = blockIdx.x;
    if (task >= A.ntask) return;
    const Job &J = A.job[find_job(A, task, false)];
    const CfgDev &C = J.c;
    const int nsub = 8 / C.fb;
    int t = task - J.ftask0;
    const int g = t % J.ngroup;
    t /= J.ngroup;
    const int sub = t % nsub;
    const int p0 = (t / nsub) * C.pb;
    const int N = C.nlon, M = C.T, rows = C.fb * C.pb;
    T2 *bufA = (T2 *)smem_raw, *bufB = bufA + rows * N;
    const int gpi = (J.kind == ANA_VECTOR) ? (J.gin2 ? 3 : 2) : 1;
    for (int idx = threadIdx.x; idx < rows * N; idx += blockDim.x) {
        const int rp = idx / N, i = idx - rp * N;
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb;
        const int gf = g * 8 + gfl;
        T2 z = T2{0, 0};
        if (pair < C.nh && gf < J.ngf) {
            const int comp = gf % gpi, item = gf / gpi;
            const double *in = comp == 0 ? J.gin0 : (comp == 1 ? J.gin1 : J.gin2);
            const double sc = comp == 0 ? J.gsc0 : (comp == 1 ? J.gsc1 : 1.0);
            in += (long)item * J.gstride;
            z = T2{T(sc * in[(long)pair * N + i]), T(sc * in[(long)(C.nlat - 1 - pair) * N + i])};
        }
        bufA[idx] = z;
    }
    __syncthreads();
    const T2 *res = fft_rows<-1, T>(bufA, bufB, rows, C);
    T *four = (T *)A.four + J.four_off + (long)g * (M + 1) * C.nhp * 32;
    for (int idx = threadIdx.x; idx < rows * (M + 1); idx += blockDim.x) {
        const int rp = idx / (M + 1), m = idx - rp * (M + 1);
        const int gfl = sub * C.fb + rp / C.pb, pair = p0 + rp % C.pb;
        if (pair >= C.nhp) continue;
        T er = 0, ei = 0, or_ = 0, oi = 0;
        if (pair < C.nh && g * 8 + gfl < J.ngf) {
            const T2 zm = res[rp * N + m];
            const T2 zc = res[rp * N + (m == 0 ? 0 : N - m)];
            // G_N = (Z_m + conj Z_{N-m})/2 ; G_S = -i (Z_m - conj Z_{N-m})/2
            const T gnr = T(0.5) * (zm.x + zc.x), gni = T(0.5) * (zm.y - zc.y);
            const T gsr = T(0.5) * (zm.y + zc.y), gsi = -T(0.5) * (zm.x - zc.x);
            const T w = T(C.wpair[pair]);
            er = w * (gnr + gsr);
            ei = w * (gni + gsi);
            or_ = w * (gnr - gsr);
            oi = w * (gni - gsi);
        }
        const long b = ((long)m * C.nhp + pair) * 32 + gfl * 2;
        four[b] = er;
        four[b + 1] = ei;
        four[b + 16] = or_;
        four[b + 17] = oi;
    }
}

}  // namespace aer
