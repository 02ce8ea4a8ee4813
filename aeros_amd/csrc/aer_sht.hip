// aer_sht.hip -- host side of the gfx950 spherical-harmonic transform library:
// the C ABI declared in include/aer_sht.h, FP64 geometry and Legendre tables,
// FFT plans, and the job builder that turns each API call into one launch per
// stage (see sht_device.h for the kernels).
#include "sht_device.h"
#include "fused_device.h"
#include "column_device.h"
#include "column_adj.h"
#include "step_device.h"
#include "diag_device.h"
#include "../../include/aer_sht.h"

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kREarth = 6.371e6;  // reference src/aeros_defs.f90:84

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(AER_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// A device buffer that grows on demand. Growing never frees the previous allocation:
// a HIP graph captured earlier holds that address in its kernel nodes (a handle's
// Fourier intermediate, the mass grids), and a later, larger eager call on the same
// handle must not pull it from under the graph's next replay. Superseded allocations
// are kept until release() (the owner's destruction). A regrowth takes at least 1.5x
// the current size, so the kept allocations sum to at most twice the live one however
// the requests creep upwards.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    std::vector<void *> retired;
    int ensure(size_t n)
    {
        if (n <= bytes) return AER_OK;
        const size_t want = bytes ? std::max(n, bytes + bytes / 2) : n;
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, want);
        if (e != hipSuccess) return fail(AER_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        if (p) retired.push_back(p);
        p = q;
        bytes = want;
        return AER_OK;
    }
    // Retires the current allocation (kept until release()), so the next ensure()
    // allocates afresh: for contents that a captured graph must keep seeing unchanged.
    void renew()
    {
        if (p) retired.push_back(p);
        p = nullptr;
        bytes = 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        for (void *q : retired) (void)hipFree(q);
        retired.clear();
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace

struct aer_sht {
    int lmax = 0, mmax = 0, mres = 1, norm = 0, layout = 0;
    int nlat = 0, nlon = 0, nlm = 0, nh = 0, nhp = 0, npt = 0, pb = 1, fb = 8, kp = 16;
    double eps = 0.0;
    bool grid_set = false;
    int prec = AER_PREC_FP64;
    int path_syn = AER_PATH_STAGED;        // kernel path of synthesis launches (aer_sht_set_path, autotune)
    int path_ana = AER_PATH_STAGED;        // kernel path of analysis launches
    int band_kb = 64;                      // tiled synthesis: staging band budget in KiB (autotune)
    hipStream_t stream = nullptr;
    std::vector<double> ct, w;             // nodes (cos colatitude) N->S, weights
    std::vector<int> li, mi;
    std::vector<int> radix, twoff;
    // device data
    DevBuf tab64[3], tab32[3], tabT64[3], tabT32[3], lfac, lfacm, wpair, tw64, tw32, twN64, twN32, four, stage_s, stage_g;
    DevBuf tabS[3], tabTS[3], expS[3], expTS[3];   // split form: float4 slices and row exponents
    DevBuf tabZ64[3], tabZ32[3], tabY64[3], tabY32[3];   // tiled path layouts (tile_table; quad_table / group_table)
    int nq = 1;                             // row quads of the tiled analysis tables
    DevBuf sinlat;                          // cos(colatitude) per row, for the column kernel
    DevBuf wfull, massg;                    // Gauss weights (nlat) and a grid for the mass sums
    double c00 = 0.0;                       // (0,0) coefficient of the constant field 1
    std::vector<double> tw_host, twN_host;  // per-pass twiddles; e^{-2 pi i k/nlon}
    // per-kernel event timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    struct Rec { int kind; hipEvent_t a, b; };
    std::vector<Rec> recs;
};

namespace {

// ------------------------------------------------------------------ host math

// 2/3/5-smooth and even (reference src/aeros_spectral.f90:440-458).
bool smooth_even(int n)
{
    if (n <= 0 || n % 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

// Gauss-Legendre nodes (descending) and weights: Newton on P_n(x) from the
// three-term recurrence; 1 - x^2 formed as (1-x)(1+x), exact to rounding for the
// represented node, so the polar weights keep full accuracy. The southern half
// is the exact mirror of the northern half.
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const int nh = (n + 1) / 2;
    for (int i = 0; i < nh; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
        double omz2 = 1.0, dp = 1.0;
        for (int it = 0; it <= 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            omz2 = (1.0 - z) * (1.0 + z);
            dp = n * (p0 - z * p1) / omz2;
            if (it == 100) break;
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) <= 1e-17) it = 99;  // one more evaluation at the converged node
        }
        x[i] = z;
        x[n - 1 - i] = -z;
        w[i] = w[n - 1 - i] = 2.0 / (omz2 * dp * dp);
    }
    if (n % 2) x[nh - 1] = 0.0;
}

// Orthonormal, no Condon-Shortley phase: P_lm, dP_lm/dtheta and m P_lm/sin(theta)
// at the northern nodes, [lm][nhp] with zero padding.
void legendre_tables(int T, int nh, int nhp, const std::vector<double> &ct,
                     std::vector<double> &P, std::vector<double> &D, std::vector<double> &Q)
{
    const int nlm = (T + 1) * (T + 2) / 2;
    P.assign((size_t)nlm * nhp, 0.0);
    D.assign((size_t)nlm * nhp, 0.0);
    Q.assign((size_t)nlm * nhp, 0.0);
    for (int j = 0; j < nh; ++j) {
        const double x = ct[j];
        const double s = std::sqrt((1.0 - x) * (1.0 + x));
        int lm = 0;
        double pmm = 1.0 / std::sqrt(4.0 * kPi);
        for (int m = 0; m <= T; ++m) {
            if (m > 0) pmm *= std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s;
            double pm2 = 0.0, pm1 = 0.0;
            for (int l = m; l <= T; ++l, ++lm) {
                double p;
                if (l == m) p = pmm;
                else if (l == m + 1) p = std::sqrt(2.0 * m + 3.0) * x * pm1;
                else {
                    const double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                    const double b = std::sqrt((((double)l - 1) * (l - 1) - (double)m * m) /
                                               (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                    p = a * (x * pm1 - b * pm2);
                }
                const double c = (l > m) ? std::sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) *
                                                     ((double)l * l - (double)m * m)) : 0.0;
                P[(size_t)lm * nhp + j] = p;
                D[(size_t)lm * nhp + j] = (l * x * p - c * pm1) / s;
                Q[(size_t)lm * nhp + j] = m * p / s;
                pm2 = pm1;
                pm1 = p;
            }
        }
    }
}

// Analysis copy of a [lm][nhp] table: [m][parity][nhp][kp] with l = m + parity + 2k
// fastest, zero where l > T.
std::vector<double> transpose_table(const std::vector<double> &X, int T, int nhp, int kp)
{
    std::vector<double> Y((size_t)(T + 1) * 2 * nhp * kp, 0.0);
    for (int m = 0; m <= T; ++m) {
        const int lmb = m * (2 * T + 3 - m) / 2;
        for (int par = 0; par < 2; ++par)
            for (int k = 0; m + par + 2 * k <= T; ++k) {
                const int lm = lmb + par + 2 * k;
                for (int j = 0; j < nhp; ++j)
                    Y[(((size_t)m * 2 + par) * nhp + j) * kp + k] = X[(size_t)lm * nhp + j];
            }
    }
    return Y;
}

// Tiled-synthesis copy of a [lm][nhp] table: [tile][lm][4], 4 latitude pairs per tile,
// so the slice of one tile is contiguous.
std::vector<double> tile_table(const std::vector<double> &X, int nlm, int nhp, int ntile)
{
    std::vector<double> Y((size_t)ntile * nlm * 4, 0.0);
    for (int t = 0; t < ntile; ++t)
        for (int lm = 0; lm < nlm; ++lm)
            for (int i = 0; i < 4; ++i) Y[((size_t)t * nlm + lm) * 4 + i] = X[(size_t)lm * nhp + 4 * t + i];
    return Y;
}

// Tiled-analysis copy: [m][parity][row quad rq][pair < nh4][4] holding the degree
// l = m + parity + 2(4 rq + i) in slot i, zero where l > T.
std::vector<double> quad_table(const std::vector<double> &X, int T, int nhp, int nh4, int nq)
{
    std::vector<double> Y((size_t)(T + 1) * 2 * nq * nh4 * 4, 0.0);
    for (int m = 0; m <= T; ++m) {
        const int lmb = m * (2 * T + 3 - m) / 2;
        for (int par = 0; par < 2; ++par)
            for (int rq = 0; rq < nq; ++rq)
                for (int p = 0; p < nh4; ++p)
                    for (int i = 0; i < 4; ++i) {
                        const int l = m + par + 2 * (4 * rq + i);
                        if (l > T) continue;
                        Y[((((size_t)m * 2 + par) * nq + rq) * nh4 + p) * 4 + i] = X[(size_t)(lmb + l - m) * nhp + p];
                    }
    }
    return Y;
}

// FP32 tiled-analysis copy: [m][parity][group G of 4 row quads][pair < nh4][16] holding
// the degree l = m + parity + 2(16 G + s) in slot s, zero where l > T.
std::vector<double> group_table(const std::vector<double> &X, int T, int nhp, int nh4, int nq)
{
    const int nG = (nq + 3) / 4;
    std::vector<double> Y((size_t)(T + 1) * 2 * nG * nh4 * 16, 0.0);
    for (int m = 0; m <= T; ++m) {
        const int lmb = m * (2 * T + 3 - m) / 2;
        for (int par = 0; par < 2; ++par)
            for (int G = 0; G < nG; ++G)
                for (int p = 0; p < nh4; ++p)
                    for (int sl = 0; sl < 16; ++sl) {
                        const int l = m + par + 2 * (16 * G + sl);
                        if (l > T) continue;
                        Y[((((size_t)m * 2 + par) * nG + G) * nh4 + p) * 16 + sl] = X[(size_t)(lmb + l - m) * nhp + p];
                    }
    }
    return Y;
}

int upload(DevBuf &b, const void *src, size_t bytes)
{
    int rc = b.ensure(bytes);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return AER_OK;
}

template <class T> std::vector<T> convert(const std::vector<double> &v)
{
    return std::vector<T>(v.begin(), v.end());
}

int check_device()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(AER_ENODEV, std::string("no HIP device: ") + hipGetErrorString(e));
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return fail(AER_ENODEV, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(AER_ENODEV, std::string("device is ") + prop.gcnArchName + ", the library is built for gfx950 only");
    return AER_OK;
}

int ensure_tables32(aer_sht *h)
{
    if (h->tab32[0].p) return AER_OK;
    std::vector<double> P, D, Q;
    legendre_tables(h->lmax, h->nh, h->nhp, h->ct, P, D, Q);
    const std::vector<double> *src[3] = {&P, &D, &Q};
    for (int k = 0; k < 3; ++k) {
        auto f = convert<float>(*src[k]);
        int rc = upload(h->tab32[k], f.data(), f.size() * sizeof(float));
        if (rc) return rc;
        auto ft = convert<float>(transpose_table(*src[k], h->lmax, h->nhp, h->kp));
        rc = upload(h->tabT32[k], ft.data(), ft.size() * sizeof(float));
        if (rc) return rc;
        auto fz = convert<float>(tile_table(*src[k], h->nlm, h->nhp, (h->nh + 3) / 4));
        if ((rc = upload(h->tabZ32[k], fz.data(), fz.size() * sizeof(float)))) return rc;
        auto fy = convert<float>(group_table(*src[k], h->lmax, h->nhp, (h->nh + 3) / 4 * 4, h->nq));
        if ((rc = upload(h->tabY32[k], fy.data(), fy.size() * sizeof(float)))) return rc;
    }
    auto tw32 = convert<float>(h->tw_host);
    int rc = upload(h->tw32, tw32.data(), tw32.size() * sizeof(float));
    if (rc) return rc;
    auto twN32 = convert<float>(h->twN_host);
    return upload(h->twN32, twN32.data(), twN32.size() * sizeof(float));
}

// Split-FP32 copies of the tables: each element as five 10-bit integer slices
// (four as float4, then the fifth in a float array after the float4 region)
// under a power-of-two scale 2^e shared by the row of A it belongs to --
// synthesis rows are latitude pairs (over the degrees of one parity of one m),
// analysis rows are degrees (over all latitude pairs) -- and the row exponents.
void split_slices(double x, int e, float *s4, float *s5)
{
    double y = std::ldexp(x, aer::SPLIT_BITS - e);
    for (int q = 0; q < 5; ++q) {
        const double a = std::rint(y);
        if (q < 4) s4[q] = (float)a;
        else *s5 = (float)a;
        y = (y - a) * 1024.0;
    }
}

int row_exponent(double amax) { return amax > 0.0 ? std::ilogb(amax) + 1 : 0; }

int ensure_tables_split(aer_sht *h)
{
    if (h->tabS[0].p) return AER_OK;
    const int T = h->lmax, nhp = h->nhp, kp = h->kp, nlm = h->nlm;
    std::vector<double> P, D, Q;
    legendre_tables(T, h->nh, nhp, h->ct, P, D, Q);
    const std::vector<double> *src[3] = {&P, &D, &Q};
    for (int t = 0; t < 3; ++t) {
        const std::vector<double> &X = *src[t];
        const size_t ns = (size_t)nlm * nhp;
        std::vector<float> sl(ns * 5, 0.f);   // ns float4, then ns fifth slices
        std::vector<int> ex((size_t)(T + 1) * 2 * nhp, 0);
        for (int m = 0; m <= T; ++m) {
            const int lmb = m * (2 * T + 3 - m) / 2;
            for (int par = 0; par < 2; ++par)
                for (int j = 0; j < nhp; ++j) {
                    double amax = 0.0;
                    for (int l = m + par; l <= T; l += 2) amax = std::max(amax, std::fabs(X[(size_t)(lmb + l - m) * nhp + j]));
                    const int e = row_exponent(amax);
                    ex[((size_t)m * 2 + par) * nhp + j] = e;
                    for (int l = m + par; l <= T; l += 2) {
                        const size_t o = (size_t)(lmb + l - m) * nhp + j;
                        split_slices(X[o], e, &sl[o * 4], &sl[ns * 4 + o]);
                    }
                }
        }
        int rc = upload(h->tabS[t], sl.data(), sl.size() * sizeof(float));
        if (!rc) rc = upload(h->expS[t], ex.data(), ex.size() * sizeof(int));
        if (rc) return rc;
        const std::vector<double> XT = transpose_table(X, T, nhp, kp);
        const size_t nt = XT.size();
        std::vector<float> slt(nt * 5, 0.f);
        std::vector<int> ext((size_t)(T + 1) * 2 * kp, 0);
        for (int mp = 0; mp < (T + 1) * 2; ++mp)
            for (int k = 0; k < kp; ++k) {
                double amax = 0.0;
                for (int j = 0; j < nhp; ++j) amax = std::max(amax, std::fabs(XT[((size_t)mp * nhp + j) * kp + k]));
                const int e = row_exponent(amax);
                ext[(size_t)mp * kp + k] = e;
                for (int j = 0; j < nhp; ++j) {
                    const size_t o = ((size_t)mp * nhp + j) * kp + k;
                    split_slices(XT[o], e, &slt[o * 4], &slt[nt * 4 + o]);
                }
            }
        rc = upload(h->tabTS[t], slt.data(), slt.size() * sizeof(float));
        if (!rc) rc = upload(h->expTS[t], ext.data(), ext.size() * sizeof(int));
        if (rc) return rc;
    }
    return AER_OK;
}

aer::CfgDev cfg_dev(const aer_sht *h, int prec)
{
    aer::CfgDev c{};
    c.T = h->lmax;
    c.nlat = h->nlat;
    c.nlon = h->nlon;
    c.nh = h->nh;
    c.nhp = h->nhp;
    c.npt = h->npt;
    c.nlm = h->nlm;
    c.pb = h->pb;
    c.fb = h->fb;
    const DevBuf *tab = (prec == AER_PREC_FP32) ? h->tab32 : (prec == AER_PREC_SPLIT ? h->tabS : h->tab64);
    c.tabP = tab[0].p;
    c.tabD = tab[1].p;
    c.tabQ = tab[2].p;
    if (prec == AER_PREC_SPLIT) {
        c.expP = (const int *)h->expS[0].p;
        c.expD = (const int *)h->expS[1].p;
        c.expQ = (const int *)h->expS[2].p;
        c.expPT = (const int *)h->expTS[0].p;
        c.expDT = (const int *)h->expTS[1].p;
        c.expQT = (const int *)h->expTS[2].p;
    }
    const DevBuf *tabT = (prec == AER_PREC_FP32) ? h->tabT32 : (prec == AER_PREC_SPLIT ? h->tabTS : h->tabT64);
    c.tabPT = tabT[0].p;
    c.tabDT = tabT[1].p;
    c.tabQT = tabT[2].p;
    c.kp = h->kp;
    const DevBuf *tz = (prec == AER_PREC_FP32) ? h->tabZ32 : h->tabZ64;
    const DevBuf *ty = (prec == AER_PREC_FP32) ? h->tabY32 : h->tabY64;
    c.tabZP = tz[0].p;
    c.tabZD = tz[1].p;
    c.tabZQ = tz[2].p;
    c.tabYP = ty[0].p;
    c.tabYD = ty[1].p;
    c.tabYQ = ty[2].p;
    c.nq = h->nq;
    c.lfac = (const double *)h->lfac.p;
    c.lfacm = (const double *)h->lfacm.p;
    c.wpair = (const double *)h->wpair.p;
    c.tw = (prec == AER_PREC_FP32) ? h->tw32.p : h->tw64.p;
    c.twN = (prec == AER_PREC_FP32) ? h->twN32.p : h->twN64.p;
    c.npass = (int)h->radix.size();
    for (int p = 0; p < c.npass; ++p) {
        c.radix[p] = h->radix[p];
        c.twoff[p] = h->twoff[p];
    }
    return c;
}

// ------------------------------------------------------------------ job launch

hipEvent_t take_event(aer_sht *h)
{
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch one kernel, bracketed by events when timing is on.
template <class F>
int timed_launch(aer_sht *h, int kind, F &&launch)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (h->timing) {
        a = take_event(h);
        b = take_event(h);
        if (!a || !b) return fail(AER_EHIP, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(a, h->stream));
    }
    launch();
    HIP_TRY(hipGetLastError());
    if (h->timing) {
        HIP_TRY(hipEventRecord(b, h->stream));
        h->recs.push_back({kind, a, b});
    }
    return AER_OK;
}

struct JobSpec {
    int kind, nitem;
    const double *sin0 = nullptr, *sin1 = nullptr;
    double *sout0 = nullptr, *sout1 = nullptr;
    double *gout0 = nullptr, *gout1 = nullptr;
    const double *gin0 = nullptr, *gin1 = nullptr, *gin2 = nullptr;
    long sstride = 0, gstride = 0;
    int mode0 = aer::M_ONE, mode1 = aer::M_ONE, mode2 = aer::M_ZERO;
    double gsc0 = 1.0, gsc1 = 1.0;
    int accumulate = 0;
    int prec = -1;          // compute type: 0 = FP64, 1 = FP32; -1 = the handle's (resolved at launch)
    int adj = 0;            // transpose of the opposite transform (see aer::Job::adj)
};

// Job table of one direction: jobs on handles hs[k] (one truncation each), their
// task ranges in the Legendre and FFT launches, the intermediate size and the FFT
// kernel's dynamic LDS. Each job's Legendre range is padded to whole 4-wave blocks.
struct Table {
    std::vector<aer::Job> jobs;
    int ltask = 0, ftask = 0;
    size_t four_bytes = 0;
    size_t lds = 0;
    bool generic = false;   // some job needs the runtime-plan FFT path
    int kmax = 1;           // max T+1 over the jobs
    int pm = aer::PM_FP64;  // kernel precision mode
    size_t cxmax = 16;      // largest complex element size among the jobs
    bool tiled = false;     // tiled path (fused synthesis; 4x4-pair analysis FFT + per-m-pair Legendre)
    size_t lds_leg = 0;     // tiled analysis Legendre kernel's dynamic LDS
    bool overflow = false;  // a task count left the 32-bit range (the launch is refused)
};

// Fields of one job: 3 grid fields per item (a vector analysis with its Laplacian term)
// stay within 32 bits, as do the tables' task counts at any size that fits in memory.
constexpr int kMaxFields = 1 << 26;

// Adds n tasks to a 32-bit launch total; flags the table when the total (and the grid
// rounded up to whole XCD rounds) would leave the int range.
inline void add_tasks(Table &B, int &total, long n)
{
    const long t = (long)total + n;
    if (n < 0 || t > 0x7fffffffL - 64) {
        B.overflow = true;
        return;
    }
    total = (int)t;
}

void build_table(aer_sht *const *hs, const JobSpec *specs, int njob, bool synth, Table &B)
{
    B = Table{};
    bool any64 = false, any32 = false, anysp = false;
    size_t cxmax = 0;
    // FFT rows of the whole launch: few rows (one model's set) take the small blocks
    long rows_all = 0;
    for (int k = 0; k < njob; ++k) {
        const JobSpec &s = specs[k];
        if (s.nitem <= 0) continue;
        const int ngf = (s.kind == aer::SYN_SCALAR || s.kind == aer::ANA_SCALAR) ? s.nitem
                        : (s.kind == aer::SYN_VECTOR || s.kind == aer::SYN_GRAD) ? 2 * s.nitem
                        : (s.gin2 ? 3 : 2) * s.nitem;
        rows_all += (long)(ngf + 7) / 8 * 8 * hs[k]->nhp;
    }
    const bool small = rows_all <= aer::AER_FFT_SMALL_LAUNCH;
    for (int k = 0; k < njob; ++k) {
        const JobSpec &s = specs[k];
        if (s.nitem <= 0) continue;
        const aer_sht *h = hs[k];
        aer::CfgDev C = cfg_dev(h, s.prec == 1 ? AER_PREC_FP32 : (s.prec == 2 ? AER_PREC_SPLIT : AER_PREC_FP64));
        int frows = aer::fft_fixed_rows(h->nlon);
        if (small && aer::fft_small_rows_ok(h->nlon)) {
            frows = aer::AER_FFT_SMALL_ROWS;
            C.fb = frows;   // fields per block: a slice of each group of 8
            C.pb = 1;
        }
        const size_t cxb = s.prec == 1 ? 2 * sizeof(float) : 2 * sizeof(double);
        (s.prec == 1 ? any32 : (s.prec == 2 ? anysp : any64)) = true;
        cxmax = std::max(cxmax, cxb);
        const int ntpp = ((C.T + 2) / 2 + 15) / 16;
        const int nsub = 8 / C.fb;
        aer::Job J{};
        J.c = C;
        J.kind = s.kind;
        J.nitem = s.nitem;
        J.prec = s.prec;
        switch (s.kind) {
            case aer::SYN_SCALAR: case aer::ANA_SCALAR: J.ngf = s.nitem; break;
            case aer::SYN_VECTOR: case aer::SYN_GRAD: J.ngf = 2 * s.nitem; break;
            default: J.ngf = (s.gin2 ? 3 : 2) * s.nitem; break;
        }
        J.ngroup = (J.ngf + 7) / 8;
        if (synth) J.lgroups = J.ngroup;
        else J.lgroups = (s.kind == aer::ANA_SCALAR) ? (s.nitem + 7) / 8 : (s.nitem + 3) / 4;
        J.sstride = s.sstride;
        J.gstride = s.gstride;
        J.sin0 = s.sin0; J.sin1 = s.sin1;
        J.sout0 = s.sout0; J.sout1 = s.sout1;
        J.gout0 = s.gout0; J.gout1 = s.gout1;
        J.gin0 = s.gin0; J.gin1 = s.gin1; J.gin2 = s.gin2;
        J.mode0 = s.mode0; J.mode1 = s.mode1; J.mode2 = s.mode2;
        J.gsc0 = s.gsc0; J.gsc1 = s.gsc1;
        J.accumulate = s.accumulate;
        J.adj = s.adj;
        J.four_off = (long)B.four_bytes;
        B.four_bytes += ((size_t)J.ngroup * (C.T + 1) * C.nhp * 32 * (cxb / 2) + 255) / 256 * 256;
        J.ltask0 = B.ltask;
        J.ftask0 = B.ftask;
        const long lt = synth ? (long)J.lgroups * (C.T + 1) * C.npt : (long)J.lgroups * (C.T + 1) * 2 * ntpp;
        J.ltasks = lt > 0x7fffffffL ? 0 : (int)lt;
        add_tasks(B, B.ltask, (lt + 3) / 4 * 4);
        add_tasks(B, B.ftask, (long)J.ngroup * nsub * (((synth ? C.nh : C.nhp) + C.pb - 1) / C.pb));
        const size_t rn = (size_t)frows * h->nlon;
        const size_t lds = aer::fft_fixed(h->nlon)
                               ? (2 * (rn + rn / 16 + 1) + h->nlon) * cxb
                               : (size_t)2 * h->fb * h->pb * h->nlon * cxb;
        B.lds = std::max(B.lds, lds);
        B.generic = B.generic || !aer::fft_fixed(h->nlon);
        B.kmax = std::max(B.kmax, C.T + 1);
        B.jobs.push_back(J);
    }
    const int kinds = (int)any64 + (int)any32 + (int)anysp;
    B.pm = kinds > 1 ? aer::PM_MIXED : (any32 ? aer::PM_FP32 : (anysp ? aer::PM_SPLIT : aer::PM_FP64));
    B.cxmax = cxmax ? cxmax : 16;
}

// Job table of the tiled path. Returns false when some job cannot take it (split
// precision, a grid length without a compile-time FFT, or LDS beyond the CU's).
bool build_table_tiled(aer_sht *const *hs, const JobSpec *specs, int njob, bool synth, Table &B)
{
    B = Table{};
    B.tiled = true;
    bool any64 = false, any32 = false;
    for (int k = 0; k < njob; ++k) {
        const JobSpec &s = specs[k];
        if (s.nitem <= 0) continue;
        const aer_sht *h = hs[k];
        if (s.prec != 0 && s.prec != 1) return false;
        if ((synth ? h->path_syn : h->path_ana) != AER_PATH_TILED || !aer::fft_fixed(h->nlon)) return false;
        // the synthesis stages coefficients as 16-byte complex loads
        if (synth && ((s.sstride & 1) || ((uintptr_t)s.sin0 & 15) || ((uintptr_t)s.sin1 & 15))) return false;
        const aer::CfgDev C = cfg_dev(h, s.prec == 1 ? AER_PREC_FP32 : AER_PREC_FP64);
        const int csz = s.prec == 1 ? (int)sizeof(float) : (int)sizeof(double);
        (s.prec == 1 ? any32 : any64) = true;
        aer::Job J{};
        J.c = C;
        J.kind = s.kind;
        J.nitem = s.nitem;
        J.prec = s.prec;
        switch (s.kind) {
            case aer::SYN_SCALAR: case aer::ANA_SCALAR: J.ngf = s.nitem; break;
            case aer::SYN_VECTOR: case aer::SYN_GRAD: J.ngf = 2 * s.nitem; break;
            default: J.ngf = (s.gin2 ? 3 : 2) * s.nitem; break;
        }
        J.sstride = s.sstride;
        J.gstride = s.gstride;
        J.sin0 = s.sin0; J.sin1 = s.sin1;
        J.sout0 = s.sout0; J.sout1 = s.sout1;
        J.gout0 = s.gout0; J.gout1 = s.gout1;
        J.gin0 = s.gin0; J.gin1 = s.gin1; J.gin2 = s.gin2;
        J.mode0 = s.mode0; J.mode1 = s.mode1; J.mode2 = s.mode2;
        J.gsc0 = s.gsc0; J.gsc1 = s.gsc1;
        J.accumulate = s.accumulate;
        J.adj = s.adj;
        const int ntile = (C.nh + 3) / 4, ng4 = (J.ngf + 3) / 4;
        if (synth) {
            const long lt = (long)ng4 * ntile;
            J.ltask0 = B.ltask;
            J.ltasks = lt > 0x7fffffffL ? 0 : (int)lt;
            add_tasks(B, B.ltask, lt);
            // staging band: every coefficient when that fits beside the FFT buffers' LDS,
            // else bands of whole orders of at most chl coefficients (chl >= T+1)
            const int fft = aer::fz_fftbytes(h->nlon, csz);
            const int per = 4 * 2 * csz + 2 * 4 * csz;   // 4 complex coefficients + 2 table rows of 4
            const int budget = std::max(fft, h->band_kb * 1024);
            J.chl = std::max(C.T + 1, std::min(C.nlm, budget / per));
            // segment stride = 8 mod 32 coefficients: the four segments' operand reads of
            // one MFMA step then start 16 banks apart
            J.chl = (J.chl + 23) / 32 * 32 + 8;
            B.lds = std::max(B.lds, (size_t)(aer::fz_fbytes(C.T, csz) + std::max(fft, aer::fz_stage_bytes(J.chl, csz))));
        } else {
            J.ftask0 = B.ftask;
            add_tasks(B, B.ftask, (long)ng4 * ntile);
            J.four_off = (long)B.four_bytes;
            B.four_bytes += ((size_t)aer::fz_four_elems(J.ngf, C.T, C.nh) * csz + 255) / 256 * 256;
            const bool scalar = s.kind == aer::ANA_SCALAR;
            const int ncg = scalar ? (s.nitem + 3) / 4 : (s.nitem + 1) / 2;
            const long lt = (long)ncg * ((C.T + 2) / 2);
            J.ltask0 = B.ltask;
            J.ltasks = lt > 0x7fffffffL ? 0 : (int)lt;
            add_tasks(B, B.ltask, lt);
            B.lds = std::max(B.lds, (size_t)aer::fz_fftbytes(h->nlon, csz));
            const int nlgf = scalar ? 4 : 2 * (s.gin2 ? 3 : 2);
            B.lds_leg = std::max(B.lds_leg, (size_t)aer::la_lds(nlgf, C.nh, csz));
        }
        B.kmax = std::max(B.kmax, C.T + 1);
        B.jobs.push_back(J);
    }
    if (B.lds > 160 * 1024 || B.lds_leg > 160 * 1024) return false;
    B.pm = (any64 && any32) ? aer::PM_MIXED : (any32 ? aer::PM_FP32 : aer::PM_FP64);
    return true;
}

// Grid of a tiled launch: whole rounds of the 8 XCDs (aer::xcd_task).
int xcd_grid(int ntask) { return (ntask + 7) / 8 * 8; }

template <int PM, bool PLAN>
int launch_tiled_pm(aer_sht *th, hipStream_t st, const aer::LaunchArgs &L, const Table &B, bool synth,
                    const int *tj_leg, const int *tj_fft)
{
    aer::LaunchArgs LL = L, LF = L;
    LL.task_job = tj_leg;
    LF.task_job = tj_fft;
    LL.kmax = LF.kmax = B.kmax;
    LL.ntask = B.ltask;
    LF.ntask = B.ftask;
    if (synth) {
        return timed_launch(th, 4, [&] {
            hipLaunchKernelGGL((aer::k_syn_fused<PM, PLAN>), dim3(xcd_grid(B.ltask)), dim3(aer::FZ_THREADS), B.lds, st, LL);
        });
    }
    int rc = timed_launch(th, 2, [&] {
        hipLaunchKernelGGL((aer::k_fft_anal4<PM, PLAN>), dim3(xcd_grid(B.ftask)), dim3(aer::FZ_THREADS), B.lds, st, LF);
    });
    if (rc) return rc;
    return timed_launch(th, 3, [&] {
        hipLaunchKernelGGL((aer::k_leg_anal4<PM, PLAN>), dim3(xcd_grid(B.ltask)), dim3(aer::LA_THREADS), B.lds_leg, st, LL);
    });
}

template <int PM, bool PLAN>
int launch_table_pm(aer_sht *th, hipStream_t st, const aer::LaunchArgs &L, const Table &B, bool synth,
                     const int *tj_leg, const int *tj_fft)
{
    aer::LaunchArgs LL = L, LF = L;
    LL.task_job = tj_leg;
    LF.task_job = tj_fft;
    LL.kmax = LF.kmax = B.kmax;
    const size_t lds_leg = synth ? (size_t)4 * 8 * B.kmax * B.cxmax : 0;
    LL.ntask = B.ltask;
    LF.ntask = B.ftask;
    const int lb = (B.ltask + 3) / 4, fb = B.ftask;
    const size_t lds = B.lds;
    int rc;
    if (synth) {
        rc = timed_launch(th, 0, [&] {
            hipLaunchKernelGGL((aer::k_leg_synth<PM, PLAN>), dim3(lb), dim3(256), lds_leg, st, LL);
        });
        if (rc) return rc;
        return timed_launch(th, 1, [&] {
            if (B.generic)
                hipLaunchKernelGGL((aer::k_fft_synth<PM, true, PLAN>), dim3(fb), dim3(AER_FFT_THREADS), lds, st, LF);
            else
                hipLaunchKernelGGL((aer::k_fft_synth<PM, false, PLAN>), dim3(fb), dim3(AER_FFT_THREADS), lds, st, LF);
        });
    }
    rc = timed_launch(th, 2, [&] {
        if (B.generic)
            hipLaunchKernelGGL((aer::k_fft_anal<PM, true, PLAN>), dim3(fb), dim3(AER_FFT_THREADS), lds, st, LF);
        else
            hipLaunchKernelGGL((aer::k_fft_anal<PM, false, PLAN>), dim3(fb), dim3(AER_FFT_THREADS), lds, st, LF);
    });
    if (rc) return rc;
    return timed_launch(th, 3, [&] {
        hipLaunchKernelGGL((aer::k_leg_anal<PM, PLAN>), dim3(lb), dim3(256), 0, st, LL);
    });
}

// The two launches of one direction. th: the handle whose timing hook records them.
int launch_table(aer_sht *th, hipStream_t st, const aer::LaunchArgs &L, const Table &B, bool synth,
                 const int *tj_leg = nullptr, const int *tj_fft = nullptr)
{
    const bool plan = L.djobs != nullptr;
    if (B.tiled) {
        switch (B.pm) {
            case aer::PM_FP64:
                return plan ? launch_tiled_pm<aer::PM_FP64, true>(th, st, L, B, synth, tj_leg, tj_fft)
                            : launch_tiled_pm<aer::PM_FP64, false>(th, st, L, B, synth, tj_leg, tj_fft);
            case aer::PM_FP32:
                return plan ? launch_tiled_pm<aer::PM_FP32, true>(th, st, L, B, synth, tj_leg, tj_fft)
                            : launch_tiled_pm<aer::PM_FP32, false>(th, st, L, B, synth, tj_leg, tj_fft);
            default:
                return plan ? launch_tiled_pm<aer::PM_MIXED, true>(th, st, L, B, synth, tj_leg, tj_fft)
                            : launch_tiled_pm<aer::PM_MIXED, false>(th, st, L, B, synth, tj_leg, tj_fft);
        }
    }
    switch (B.pm) {
        case aer::PM_FP64:
            return plan ? launch_table_pm<aer::PM_FP64, true>(th, st, L, B, synth, tj_leg, tj_fft)
                        : launch_table_pm<aer::PM_FP64, false>(th, st, L, B, synth, tj_leg, tj_fft);
        case aer::PM_FP32:
            return plan ? launch_table_pm<aer::PM_FP32, true>(th, st, L, B, synth, tj_leg, tj_fft)
                        : launch_table_pm<aer::PM_FP32, false>(th, st, L, B, synth, tj_leg, tj_fft);
        case aer::PM_SPLIT:
            return plan ? launch_table_pm<aer::PM_SPLIT, true>(th, st, L, B, synth, tj_leg, tj_fft)
                        : launch_table_pm<aer::PM_SPLIT, false>(th, st, L, B, synth, tj_leg, tj_fft);
        default:
            return plan ? launch_table_pm<aer::PM_MIXED, true>(th, st, L, B, synth, tj_leg, tj_fft)
                        : launch_table_pm<aer::PM_MIXED, false>(th, st, L, B, synth, tj_leg, tj_fft);
    }
}

// Resolve a job's compute type (its own, else the handle's setting) and make sure
// the FP32 tables exist when it needs them.
int resolve_prec(aer_sht *h, JobSpec &s)
{
    if (s.prec < 0) s.prec = h->prec == AER_PREC_FP32 ? 1 : (h->prec == AER_PREC_SPLIT ? 2 : 0);
    if (s.prec == 1) return ensure_tables32(h);
    if (s.prec == 2) return ensure_tables_split(h);
    return AER_OK;
}

int run_jobs(aer_sht *h, JobSpec *specs, int njob, bool synth)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    if (njob > AER_MAXJOB) return fail(AER_EINVAL, "too many jobs in one launch");
    aer_sht *hs[AER_MAXJOB];
    for (int k = 0; k < njob; ++k) {
        hs[k] = h;
        int rc = resolve_prec(h, specs[k]);
        if (rc) return rc;
    }
    Table B;
    if (!build_table_tiled(hs, specs, njob, synth, B)) build_table(hs, specs, njob, synth, B);
    if (B.overflow) return fail(AER_EINVAL, "launch too large: its task count leaves 32 bits; split the fields");
    if (B.jobs.empty()) return AER_OK;
    int rc = h->four.ensure(B.four_bytes);
    if (rc) return rc;
    aer::LaunchArgs L{};
    L.njob = (int)B.jobs.size();
    L.four = h->four.p;
    L.djobs = nullptr;
    for (int k = 0; k < L.njob; ++k) {
        L.job[k] = B.jobs[k];
        L.lstart[k] = B.jobs[k].ltask0;
        L.fstart[k] = B.jobs[k].ftask0;
    }
    return launch_table(h, h->stream, L, B, synth);
}

// The (0,0) coefficient of the constant field 1, read back by an FP64 analysis
// rather than written down (reference src/dynamics/aeros_timestep.f90:573-588).
int read_c00(aer_sht *h)
{
    const size_t ng = (size_t)h->nlon * h->nlat;
    std::vector<double> one(ng, 1.0);
    int rc = upload(h->massg, one.data(), ng * sizeof(double));
    if (rc) return rc;
    if ((rc = h->stage_s.ensure(2 * sizeof(double) * h->nlm))) return rc;
    JobSpec s{aer::ANA_SCALAR, 1};
    s.gin0 = (const double *)h->massg.p;
    s.gstride = (long)ng;
    s.sout0 = (double *)h->stage_s.p;
    s.sstride = 2L * h->nlm;
    s.prec = 0;
    if ((rc = run_jobs(h, &s, 1, false))) return rc;
    double c[2];
    HIP_TRY(hipMemcpyAsync(c, h->stage_s.p, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->c00 = c[0];
    if (std::fabs(h->c00) < 1e-12) return fail(AER_EHIP, "aer_sht_set_grid: the (0,0) coefficient of 1 read back as ~0");
    return AER_OK;
}

long dense_s(const aer_sht *h) { return 2L * h->nlm; }
long dense_g(const aer_sht *h) { return (long)h->nlon * h->nlat; }

int check_batch(const aer_sht *h, int nfield, long qstride, long gstride)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    if (nfield < 0) return fail(AER_EINVAL, "nfield < 0");
    if (nfield > kMaxFields) return fail(AER_EINVAL, "nfield above 2^26 in one call");
    if (qstride < dense_s(h)) return fail(AER_EINVAL, "spectral stride smaller than 2*nlm");
    if (gstride < dense_g(h)) return fail(AER_EINVAL, "grid stride smaller than nlon*nlat");
    return AER_OK;
}

int stage(aer_sht *h, int ns, int ng)
{
    int rc = h->stage_s.ensure((size_t)ns * dense_s(h) * sizeof(double));
    if (rc) return rc;
    return h->stage_g.ensure((size_t)ng * dense_g(h) * sizeof(double));
}

}  // namespace

// =================================================================== C ABI

extern "C" {

const char *aer_strerror(int status)
{
    switch (status) {
        case AER_OK: return "success";
        case AER_EINVAL: return "invalid argument";
        case AER_ENOMEM: return "out of memory";
        case AER_EHIP: return "HIP runtime error";
        case AER_ENODEV: return "no gfx950 device";
        case AER_ESTATE: return "call out of order";
        default: return "unknown status";
    }
}

const char *aer_last_error(void) { return g_last_error.c_str(); }

int aer_sht_grid_size(int trunc, int *nlon, int *nlat)
{
    if (trunc < 0 || !nlon || !nlat) return fail(AER_EINVAL, "aer_sht_grid_size: bad argument");
    int n = 3 * trunc + 1;
    while (!smooth_even(n)) ++n;
    *nlon = n;
    *nlat = n / 2;
    return AER_OK;
}

int aer_sht_create(int lmax, int mmax, int mres, int norm, aer_sht_t *cfg)
{
    if (!cfg) return fail(AER_EINVAL, "aer_sht_create: null output");
    *cfg = nullptr;
    if (lmax < 1 || mmax != lmax || mres != 1)
        return fail(AER_EINVAL, "aer_sht_create: only triangular lmax = mmax >= 1, mres = 1 is supported");
    if ((norm & (AER_SHT_ORTHONORMAL | AER_SHT_NO_CS_PHASE)) != (AER_SHT_ORTHONORMAL | AER_SHT_NO_CS_PHASE))
        return fail(AER_EINVAL, "aer_sht_create: norm must be AER_SHT_ORTHONORMAL | AER_SHT_NO_CS_PHASE");
    aer_sht *h = new aer_sht();
    h->lmax = lmax;
    h->mmax = mmax;
    h->mres = mres;
    h->norm = norm;
    h->nlm = (lmax + 1) * (lmax + 2) / 2;
    for (int m = 0; m <= mmax; ++m)
        for (int l = m; l <= lmax; ++l) {
            h->li.push_back(l);
            h->mi.push_back(m);
        }
    *cfg = h;
    return AER_OK;
}

int aer_sht_set_grid(aer_sht_t h, int layout, double eps, int nlat, int nphi)
{
    if (!h) return fail(AER_EINVAL, "aer_sht_set_grid: null handle");
    if (!(layout & AER_SHT_PHI_CONTIGUOUS)) return fail(AER_EINVAL, "aer_sht_set_grid: only AER_SHT_PHI_CONTIGUOUS is supported");
    if (nlat < h->lmax + 1) return fail(AER_EINVAL, "aer_sht_set_grid: nlat < lmax+1 is not exact");
    if (nphi <= 2 * h->mmax) return fail(AER_EINVAL, "aer_sht_set_grid: nphi <= 2*mmax");
    {
        int n = nphi;
        while (n % 2 == 0) n /= 2;
        while (n % 3 == 0) n /= 3;
        while (n % 5 == 0) n /= 5;
        if (n != 1) return fail(AER_EINVAL, "aer_sht_set_grid: nphi must factor into 2, 3 and 5");
    }
    int rc = check_device();
    if (rc) return rc;
    h->layout = layout;
    h->eps = eps;
    h->nlat = nlat;
    h->nlon = nphi;
    h->nh = (nlat + 1) / 2;
    h->nhp = (h->nh + 15) / 16 * 16;
    h->npt = h->nhp / 16;
    // FFT block: fb grid fields x pb latitude pairs. Lengths with a compile-time
    // path use 8 fields x fft_fixed_rows/8 pairs; others at most 2048 complex per buffer.
    if (aer::fft_fixed(nphi)) {
        h->fb = 8;
        h->pb = aer::fft_fixed_rows(nphi) / 8;
    } else {
        h->fb = 8;
        h->pb = std::max(1, 128 / nphi);
        while (h->fb * h->pb * nphi > 2048 && h->fb > 1) h->fb /= 2;
    }
    gauss_legendre(nlat, h->ct, h->w);
    std::vector<double> P, D, Q;
    legendre_tables(h->lmax, h->nh, h->nhp, h->ct, P, D, Q);
    const std::vector<double> *tabs[3] = {&P, &D, &Q};
    h->kp = (h->lmax / 2 + 1 + 15) / 16 * 16;
    h->nq = ((h->lmax + 2) / 2 + 3) / 4;
    for (int k = 0; k < 3; ++k) {
        if ((rc = upload(h->tab64[k], tabs[k]->data(), tabs[k]->size() * sizeof(double)))) return rc;
        const std::vector<double> tt = transpose_table(*tabs[k], h->lmax, h->nhp, h->kp);
        if ((rc = upload(h->tabT64[k], tt.data(), tt.size() * sizeof(double)))) return rc;
        const std::vector<double> tz = tile_table(*tabs[k], h->nlm, h->nhp, (h->nh + 3) / 4);
        if ((rc = upload(h->tabZ64[k], tz.data(), tz.size() * sizeof(double)))) return rc;
        const std::vector<double> ty = quad_table(*tabs[k], h->lmax, h->nhp, (h->nh + 3) / 4 * 4, h->nq);
        if ((rc = upload(h->tabY64[k], ty.data(), ty.size() * sizeof(double)))) return rc;
    }
    for (auto &b : h->tab32) b.release();
    for (auto &b : h->tabT32) b.release();
    for (auto &b : h->tabZ32) b.release();
    for (auto &b : h->tabY32) b.release();
    // per-degree factors
    const int T = h->lmax;
    std::vector<double> lf((size_t)AER_NMODE * (T + 1), 0.0);
    for (int l = 0; l <= T; ++l) {
        const double ll1 = (double)(l * (l + 1));
        lf[aer::M_ONE * (T + 1) + l] = 1.0;
        lf[aer::M_S_FROM_DIV * (T + 1) + l] = l ? -(kREarth / ll1) : 0.0;
        lf[aer::M_T_FROM_VOR * (T + 1) + l] = l ? (kREarth / ll1) : 0.0;
        lf[aer::M_INV_LL1 * (T + 1) + l] = l ? 1.0 / ll1 : 0.0;
        lf[aer::M_DIV_FROM_S * (T + 1) + l] = l ? -1.0 / kREarth : 0.0;
        lf[aer::M_VOR_FROM_T * (T + 1) + l] = l ? 1.0 / kREarth : 0.0;
        lf[aer::M_NEG_LAP * (T + 1) + l] = ll1 / (kREarth * kREarth);
        lf[aer::M_ZERO * (T + 1) + l] = 0.0;
    }
    if ((rc = upload(h->lfac, lf.data(), lf.size() * sizeof(double)))) return rc;
    {
        std::vector<double> lfm((size_t)AER_NMODE * h->nlm);
        for (int md = 0; md < AER_NMODE; ++md)
            for (int m = 0, lm = 0; m <= T; ++m)
                for (int l = m; l <= T; ++l, ++lm) lfm[(size_t)md * h->nlm + lm] = lf[md * (T + 1) + l];
        if ((rc = upload(h->lfacm, lfm.data(), lfm.size() * sizeof(double)))) return rc;
    }
    std::vector<double> wp(h->nhp, 0.0);
    for (int j = 0; j < h->nh; ++j) {
        wp[j] = h->w[j] * 2.0 * kPi / nphi;
        if (nlat - 1 - j == j) wp[j] *= 0.5;
    }
    if ((rc = upload(h->wpair, wp.data(), wp.size() * sizeof(double)))) return rc;
    if ((rc = upload(h->sinlat, h->ct.data(), h->ct.size() * sizeof(double)))) return rc;
    if ((rc = upload(h->wfull, h->w.data(), h->w.size() * sizeof(double)))) return rc;
    if ((rc = h->massg.ensure((size_t)nlat * nphi * sizeof(double)))) return rc;
    // FFT plan: radices 4, 2, 3, 5 and the forward twiddles of each pass
    h->radix.clear();
    h->twoff.clear();
    {
        int n = nphi;
        while (n % 4 == 0) { h->radix.push_back(4); n /= 4; }
        while (n % 2 == 0) { h->radix.push_back(2); n /= 2; }
        while (n % 3 == 0) { h->radix.push_back(3); n /= 3; }
        while (n % 5 == 0) { h->radix.push_back(5); n /= 5; }
    }
    if ((int)h->radix.size() > AER_MAXPASS) return fail(AER_EINVAL, "aer_sht_set_grid: nphi needs too many FFT passes");
    std::vector<double> tw;
    int Ns = 1;
    for (int R : h->radix) {
        h->twoff.push_back((int)(tw.size() / 2));
        for (int k = 0; k < Ns; ++k)
            for (int r = 1; r < R; ++r) {
                const double a = -2.0 * kPi * r * k / ((double)Ns * R);
                tw.push_back(std::cos(a));
                tw.push_back(std::sin(a));
            }
        Ns *= R;
    }
    if (tw.empty()) tw.assign(2, 0.0);
    if ((rc = upload(h->tw64, tw.data(), tw.size() * sizeof(double)))) return rc;
    h->tw_host = tw;
    std::vector<double> twN(2 * (size_t)nphi);
    for (int k = 0; k < nphi; ++k) {
        const double a = -2.0 * kPi * k / nphi;
        twN[2 * k] = std::cos(a);
        twN[2 * k + 1] = std::sin(a);
    }
    if ((rc = upload(h->twN64, twN.data(), twN.size() * sizeof(double)))) return rc;
    h->twN_host = twN;
    h->tw32.release();
    h->twN32.release();
    h->grid_set = true;
    return read_c00(h);
}

int aer_sht_host_tables(int lmax, int nlat, double *ct, double *w, double *P, double *D, double *Q)
{
    if (lmax < 0 || nlat < 1 || !ct || !w || !P || !D || !Q) return fail(AER_EINVAL, "aer_sht_host_tables: bad argument");
    std::vector<double> x, ww, tP, tD, tQ;
    gauss_legendre(nlat, x, ww);
    const int nh = (nlat + 1) / 2, nlm = (lmax + 1) * (lmax + 2) / 2;
    legendre_tables(lmax, nh, nh, x, tP, tD, tQ);
    std::memcpy(ct, x.data(), sizeof(double) * nlat);
    std::memcpy(w, ww.data(), sizeof(double) * nlat);
    std::memcpy(P, tP.data(), sizeof(double) * (size_t)nlm * nh);
    std::memcpy(D, tD.data(), sizeof(double) * (size_t)nlm * nh);
    std::memcpy(Q, tQ.data(), sizeof(double) * (size_t)nlm * nh);
    return AER_OK;
}

int aer_sht_get_info(aer_sht_t h, aer_sht_info *info)
{
    if (!h || !info) return fail(AER_EINVAL, "aer_sht_get_info: null argument");
    info->lmax = h->lmax;
    info->mmax = h->mmax;
    info->mres = h->mres;
    info->nlm = h->nlm;
    info->nlat = h->nlat;
    info->nphi = h->nlon;
    info->nspat = h->nlat * h->nlon;
    info->nlat_2 = (h->nlat + 1) / 2;
    return AER_OK;
}

int aer_sht_get_lm(aer_sht_t h, int *li, int *mi)
{
    if (!h || !li || !mi) return fail(AER_EINVAL, "aer_sht_get_lm: null argument");
    std::memcpy(li, h->li.data(), sizeof(int) * h->nlm);
    std::memcpy(mi, h->mi.data(), sizeof(int) * h->nlm);
    return AER_OK;
}

int aer_sht_get_cos_theta(aer_sht_t h, double *ct)
{
    if (!h || !ct) return fail(AER_EINVAL, "aer_sht_get_cos_theta: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "aer_sht_get_cos_theta before set_grid");
    std::memcpy(ct, h->ct.data(), sizeof(double) * h->nlat);
    return AER_OK;
}

int aer_sht_gauss_wts(aer_sht_t h, double *wts_half)
{
    if (!h || !wts_half) return fail(AER_EINVAL, "aer_sht_gauss_wts: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "aer_sht_gauss_wts before set_grid");
    std::memcpy(wts_half, h->w.data(), sizeof(double) * ((h->nlat + 1) / 2));
    return AER_OK;
}

int aer_sht_timing(aer_sht_t h, int enable)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    for (auto &r : h->recs) {
        h->ev_pool.push_back(r.a);
        h->ev_pool.push_back(r.b);
    }
    h->recs.clear();
    h->timing = enable != 0;
    return AER_OK;
}

int aer_sht_timing_read(aer_sht_t h, double *total_ms, int *launches)
{
    if (!h || !total_ms || !launches) return fail(AER_EINVAL, "aer_sht_timing_read: null argument");
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int k = 0; k < AER_NKERNEL_KIND; ++k) {
        total_ms[k] = 0.0;
        launches[k] = 0;
    }
    for (auto &r : h->recs) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        total_ms[r.kind] += ms;
        launches[r.kind] += 1;
    }
    return AER_OK;
}

int aer_sht_destroy(aer_sht_t h)
{
    if (!h) return AER_OK;
    for (auto &r : h->recs) {
        h->ev_pool.push_back(r.a);
        h->ev_pool.push_back(r.b);
    }
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    for (auto &b : h->tab64) b.release();
    for (auto &b : h->tab32) b.release();
    for (auto &b : h->tabT64) b.release();
    for (auto &b : h->tabT32) b.release();
    for (auto &b : h->tabZ64) b.release();
    for (auto &b : h->tabZ32) b.release();
    for (auto &b : h->tabY64) b.release();
    for (auto &b : h->tabY32) b.release();
    h->lfac.release();
    h->lfacm.release();
    h->wpair.release();
    h->tw64.release();
    h->tw32.release();
    h->twN64.release();
    h->twN32.release();
    h->four.release();
    h->stage_s.release();
    h->stage_g.release();
    h->sinlat.release();
    for (int k = 0; k < 3; ++k) {
        h->tabS[k].release();
        h->tabTS[k].release();
        h->expS[k].release();
        h->expTS[k].release();
    }
    h->wfull.release();
    h->massg.release();
    delete h;
    return AER_OK;
}

int aer_sht_set_precision(aer_sht_t h, int prec)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    if (prec != AER_PREC_FP64 && prec != AER_PREC_FP32 && prec != AER_PREC_SPLIT)
        return fail(AER_EINVAL, "aer_sht_set_precision: unsupported precision");
    h->prec = prec;
    return AER_OK;
}

int aer_sht_set_path(aer_sht_t h, int path)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    if (path != AER_PATH_TILED && path != AER_PATH_STAGED) return fail(AER_EINVAL, "aer_sht_set_path: unknown path");
    h->path_syn = h->path_ana = path;
    return AER_OK;
}

int aer_sht_set_stream(aer_sht_t h, void *stream)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    h->stream = (hipStream_t)stream;
    return AER_OK;
}

}  // extern "C"

namespace {

// Public job descriptor -> internal job spec, with validation.
int job_to_spec(const aer_sht *h, const aer_job &j, JobSpec &s)
{
    if (j.kind < AER_JOB_SYNTH_SCALAR || j.kind > AER_JOB_ANAL_VECTOR)
        return fail(AER_EINVAL, "job: unknown kind");
    const bool synth = j.kind <= AER_JOB_SYNTH_GRAD;
    if (j.nfield < 0) return fail(AER_EINVAL, "job: nfield < 0");
    if (j.nfield > kMaxFields) return fail(AER_EINVAL, "job: nfield above 2^26");
    for (int md : {j.mode0, j.mode1, j.mode2})
        if (md < 0 || md >= AER_NMODE) return fail(AER_EINVAL, "job: unknown mode");
    const bool vec_in = j.kind == AER_JOB_SYNTH_VECTOR || j.kind == AER_JOB_ANAL_VECTOR;
    const bool two_out = j.kind == AER_JOB_SYNTH_VECTOR || j.kind == AER_JOB_SYNTH_GRAD || j.kind == AER_JOB_ANAL_VECTOR;
    if (j.nfield > 0 && (!j.in0 || !j.out0 || (vec_in && !j.in1) || (two_out && !j.out1)))
        return fail(AER_EINVAL, "job: missing array");
    const long ss = dense_s(h), gs = dense_g(h);
    s = JobSpec{j.kind, j.nfield};
    const long is = j.in_stride ? j.in_stride : (synth ? ss : gs);
    const long os = j.out_stride ? j.out_stride : (synth ? gs : ss);
    if (is < (synth ? ss : gs) || os < (synth ? gs : ss))
        return fail(AER_EINVAL, "job: stride smaller than a dense field");
    if (synth) {
        s.sin0 = j.in0; s.sin1 = j.in1; s.sstride = is;
        s.gout0 = j.out0; s.gout1 = j.out1; s.gstride = os;
    } else {
        s.gin0 = j.in0; s.gin1 = j.in1; s.gin2 = j.kind == AER_JOB_ANAL_VECTOR ? j.in2 : nullptr;
        s.gstride = is;
        s.sout0 = j.out0; s.sout1 = j.out1; s.sstride = os;
    }
    s.mode0 = j.mode0; s.mode1 = j.mode1; s.mode2 = j.mode2;
    s.gsc0 = j.scale0 != 0.0 ? j.scale0 : 1.0;
    s.gsc1 = j.scale1 != 0.0 ? j.scale1 : 1.0;
    s.accumulate = j.accumulate;
    switch (j.prec) {
        case AER_JOB_PREC_HANDLE: s.prec = -1; break;
        case AER_JOB_PREC_FP64: s.prec = 0; break;
        case AER_JOB_PREC_FP32: s.prec = 1; break;
        case AER_JOB_PREC_SPLIT: s.prec = 2; break;
        default: return fail(AER_EINVAL, "job: unknown precision");
    }
    return AER_OK;
}

}  // namespace

struct aer_plan {
    aer_sht *th = nullptr;           // timing hook owner (first job's handle)
    Table syn, ana;
    DevBuf djob_syn, djob_ana, four;
    DevBuf tj[4];                    // task -> job: synthesis Legendre, FFT; analysis FFT, Legendre
    std::vector<aer_sht *> hsy, han; // the jobs of each direction and their handles (for rebuilds)
    std::vector<JobSpec> ssy, san;
};

namespace {

int plan_build(aer_plan *p)
{
    std::vector<aer_sht *> &hsy = p->hsy, &han = p->han;
    std::vector<JobSpec> &ssy = p->ssy, &san = p->san;
    if (!build_table_tiled(hsy.data(), ssy.data(), (int)ssy.size(), true, p->syn))
        build_table(hsy.data(), ssy.data(), (int)ssy.size(), true, p->syn);
    if (!build_table_tiled(han.data(), san.data(), (int)san.size(), false, p->ana))
        build_table(han.data(), san.data(), (int)san.size(), false, p->ana);
    if (p->syn.overflow || p->ana.overflow)
        return fail(AER_EINVAL, "plan too large: a launch's task count leaves 32 bits; split the jobs");
    // fresh job tables on every (re)build: a graph that recorded this plan's launches
    // before an autotune keeps its own tables, consistent with its recorded grids
    p->djob_syn.renew();
    p->djob_ana.renew();
    for (auto &b : p->tj) b.renew();
    int rc;
    if (!p->syn.jobs.empty() && (rc = upload(p->djob_syn, p->syn.jobs.data(), p->syn.jobs.size() * sizeof(aer::Job))))
        return rc;
    if (!p->ana.jobs.empty() && (rc = upload(p->djob_ana, p->ana.jobs.data(), p->ana.jobs.size() * sizeof(aer::Job))))
        return rc;
    // task -> job index tables of the four launches
    const Table *tabs[2] = {&p->syn, &p->ana};
    for (int d = 0; d < 2; ++d) {
        const Table &B = *tabs[d];
        if (B.jobs.empty()) continue;
        std::vector<int> tl(std::max(B.ltask, 1), 0), tf(std::max(B.ftask, 1), 0);
        for (size_t k = 0; k < B.jobs.size(); ++k) {
            const int l1 = k + 1 < B.jobs.size() ? B.jobs[k + 1].ltask0 : B.ltask;
            const int f1 = k + 1 < B.jobs.size() ? B.jobs[k + 1].ftask0 : B.ftask;
            for (int t = B.jobs[k].ltask0; t < l1; ++t) tl[t] = (int)k;
            for (int t = B.jobs[k].ftask0; t < f1; ++t) tf[t] = (int)k;
        }
        DevBuf &bl = p->tj[d == 0 ? 0 : 3], &bf = p->tj[d == 0 ? 1 : 2];
        if ((rc = upload(bl, tl.data(), std::max<size_t>(tl.size(), 1) * sizeof(int)))) return rc;
        if ((rc = upload(bf, tf.data(), std::max<size_t>(tf.size(), 1) * sizeof(int)))) return rc;
    }
    return p->four.ensure(std::max<size_t>(std::max(p->syn.four_bytes, p->ana.four_bytes), 256));
}

int plan_launch(aer_plan *p, hipStream_t st)
{
    aer::LaunchArgs L{};
    L.four = p->four.p;
    int rc;
    if (!p->syn.jobs.empty()) {
        L.njob = (int)p->syn.jobs.size();
        L.djobs = (const aer::Job *)p->djob_syn.p;
        if ((rc = launch_table(p->th, st, L, p->syn, true, (const int *)p->tj[0].p, (const int *)p->tj[1].p)))
            return rc;
    }
    if (!p->ana.jobs.empty()) {
        L.njob = (int)p->ana.jobs.size();
        L.djobs = (const aer::Job *)p->djob_ana.p;
        if ((rc = launch_table(p->th, st, L, p->ana, false, (const int *)p->tj[3].p, (const int *)p->tj[2].p)))
            return rc;
    }
    return AER_OK;
}

}  // namespace

extern "C" {

int aer_plan_create(int njob, const aer_sht_t *cfgs, const aer_job *jobs, aer_plan_t *plan)
{
    if (!plan) return fail(AER_EINVAL, "aer_plan_create: null output");
    *plan = nullptr;
    if (njob <= 0 || !cfgs || !jobs) return fail(AER_EINVAL, "aer_plan_create: empty job list");
    std::vector<aer_sht *> hs(njob);
    std::vector<JobSpec> specs(njob);
    int rc;
    for (int k = 0; k < njob; ++k) {
        aer_sht *h = cfgs[k];
        if (!h) return fail(AER_EINVAL, "aer_plan_create: null handle");
        if (!h->grid_set) return fail(AER_ESTATE, "aer_plan_create: handle without a grid");
        if ((rc = job_to_spec(h, jobs[k], specs[k]))) return rc;
        if ((rc = resolve_prec(h, specs[k]))) return rc;
        hs[k] = h;
    }
    aer_plan *p = new aer_plan();
    p->th = hs[0];
    for (int k = 0; k < njob; ++k) {
        const bool synth = specs[k].kind <= aer::SYN_GRAD;
        (synth ? p->hsy : p->han).push_back(hs[k]);
        (synth ? p->ssy : p->san).push_back(specs[k]);
    }
    rc = plan_build(p);
    if (rc) {
        aer_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return AER_OK;
}

int aer_plan_run(aer_plan_t p, void *stream)
{
    if (!p) return fail(AER_EINVAL, "aer_plan_run: null plan");
    return plan_launch(p, (hipStream_t)stream);
}

int aer_plan_destroy(aer_plan_t p)
{
    if (!p) return AER_OK;
    p->djob_syn.release();
    p->djob_ana.release();
    p->four.release();
    for (auto &b : p->tj) b.release();
    delete p;
    return AER_OK;
}

}  // extern "C"

namespace {

// Autotuner candidates: the staged path, and the tiled path with m-band staging budgets
// of 16, 32 and 64 KiB (bands of whole orders; smaller bands leave more LDS for
// co-resident blocks, larger ones need fewer barriers).
struct TuneCand {
    int path, band_kb;
};
constexpr TuneCand kTune[AER_TUNE_NCAND] = {{AER_PATH_STAGED, 64}, {AER_PATH_TILED, 16}, {AER_PATH_TILED, 32},
                                            {AER_PATH_TILED, 64}};

void tune_apply(aer_sht *h, bool synth, const TuneCand &c)
{
    (synth ? h->path_syn : h->path_ana) = c.path;
    if (synth) h->band_kb = c.band_kb;
}

// Mean milliseconds of `run` over reps launches on stream st after one warm-up.
template <class F>
int tune_time(hipStream_t st, int reps, F &&run, double &ms)
{
    int rc = run();
    if (rc) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, st));
    for (int r = 0; r < reps && !rc; ++r) rc = run();
    HIP_TRY(hipEventRecord(b, st));
    HIP_TRY(hipEventSynchronize(b));
    float f = 0.0f;
    HIP_TRY(hipEventElapsedTime(&f, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ms = (double)f / std::max(reps, 1);
    return rc;
}

}  // namespace

extern "C" {

int aer_autotune_jobs(aer_sht_t h, int njob, const aer_job *jobs, int reps, double *times_ms, int *choice)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    if (njob <= 0 || njob > AER_MAX_JOBS || !jobs) return fail(AER_EINVAL, "aer_autotune_jobs: bad job count");
    if (reps <= 0) return fail(AER_EINVAL, "aer_autotune_jobs: reps must be positive");
    JobSpec spec[AER_MAX_JOBS];
    const bool synth = jobs[0].kind <= AER_JOB_SYNTH_GRAD;
    for (int k = 0; k < njob; ++k) {
        int rc = job_to_spec(h, jobs[k], spec[k]);
        if (rc) return rc;
        if ((jobs[k].kind <= AER_JOB_SYNTH_GRAD) != synth)
            return fail(AER_EINVAL, "aer_autotune_jobs: synthesis and analysis jobs cannot share a launch");
    }
    int best = 0;
    double tbest = 0.0;
    for (int c = 0; c < AER_TUNE_NCAND; ++c) {
        tune_apply(h, synth, kTune[c]);
        double ms = 0.0;
        int rc = tune_time(h->stream, reps, [&] { return run_jobs(h, spec, njob, synth); }, ms);
        if (rc) return rc;
        if (times_ms) times_ms[c] = ms;
        if (c == 0 || ms < tbest) { best = c; tbest = ms; }
    }
    tune_apply(h, synth, kTune[best]);
    if (choice) *choice = best;
    return AER_OK;
}

int aer_plan_autotune(aer_plan_t p, void *stream, int reps, double *times_ms, int *choice)
{
    if (!p) return fail(AER_EINVAL, "aer_plan_autotune: null plan");
    if (reps <= 0) return fail(AER_EINVAL, "aer_plan_autotune: reps must be positive");
    hipStream_t st = (hipStream_t)stream;
    for (int d = 0; d < 2; ++d) {
        const bool synth = d == 0;
        std::vector<aer_sht *> &hs = synth ? p->hsy : p->han;
        if (hs.empty()) {
            if (choice) choice[d] = 0;
            continue;
        }
        int best = 0;
        double tbest = 0.0;
        for (int c = 0; c < AER_TUNE_NCAND; ++c) {
            for (aer_sht *h : hs) tune_apply(h, synth, kTune[c]);
            int rc = plan_build(p);
            if (rc) return rc;
            aer::LaunchArgs L{};
            L.four = p->four.p;
            const Table &B = synth ? p->syn : p->ana;
            L.njob = (int)B.jobs.size();
            L.djobs = (const aer::Job *)(synth ? p->djob_syn.p : p->djob_ana.p);
            const int *tl = (const int *)p->tj[synth ? 0 : 3].p, *tf = (const int *)p->tj[synth ? 1 : 2].p;
            double ms = 0.0;
            rc = tune_time(st, reps, [&] { return launch_table(p->th, st, L, B, synth, tl, tf); }, ms);
            if (rc) return rc;
            if (times_ms) times_ms[d * AER_TUNE_NCAND + c] = ms;
            if (c == 0 || ms < tbest) { best = c; tbest = ms; }
        }
        for (aer_sht *h : hs) tune_apply(h, synth, kTune[best]);
        if (choice) choice[d] = best;
    }
    return plan_build(p);
}

int aer_run_jobs_d(aer_sht_t h, int njob, const aer_job *jobs)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    if (njob < 0 || njob > AER_MAX_JOBS || (njob > 0 && !jobs)) return fail(AER_EINVAL, "aer_run_jobs_d: bad job count");
    if (njob == 0) return AER_OK;
    static_assert(AER_MAX_JOBS <= AER_MAXJOB, "job table size");
    JobSpec spec[AER_MAX_JOBS];
    const bool synth = jobs[0].kind <= AER_JOB_SYNTH_GRAD;
    for (int k = 0; k < njob; ++k) {
        int rc = job_to_spec(h, jobs[k], spec[k]);
        if (rc) return rc;
        if ((jobs[k].kind <= AER_JOB_SYNTH_GRAD) != synth)
            return fail(AER_EINVAL, "aer_run_jobs_d: synthesis and analysis jobs cannot share a launch");
    }
    return run_jobs(h, spec, njob, synth);
}

int aer_dev_malloc(size_t bytes, void **ptr)
{
    if (!ptr) return fail(AER_EINVAL, "aer_dev_malloc: null output");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(AER_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return AER_OK;
}

int aer_dev_free(void *ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return AER_OK;
}

int aer_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes && (!dst || !src)) return fail(AER_EINVAL, "aer_memcpy_h2d: null pointer");
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return AER_OK;
}

int aer_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes && (!dst || !src)) return fail(AER_EINVAL, "aer_memcpy_d2h: null pointer");
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return AER_OK;
}

int aer_sht_sync(aer_sht_t h)
{
    if (!h) return fail(AER_EINVAL, "null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return AER_OK;
}

// ---- batched device transforms

int aer_synth_d(aer_sht_t h, int nfield, const double *qlm, long qstride, double *vr, long gstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::SYN_SCALAR, nfield};
    s.sin0 = qlm; s.sstride = qstride; s.gout0 = vr; s.gstride = gstride;
    return run_jobs(h, &s, 1, true);
}

int aer_anal_d(aer_sht_t h, int nfield, const double *vr, long gstride, double *qlm, long qstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::ANA_SCALAR, nfield};
    s.gin0 = vr; s.gstride = gstride; s.sout0 = qlm; s.sstride = qstride;
    return run_jobs(h, &s, 1, false);
}

int aer_synth_vec_d(aer_sht_t h, int nfield, const double *slm, const double *tlm, long qstride,
                    double *vt, double *vp, long gstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::SYN_VECTOR, nfield};
    s.sin0 = slm; s.sin1 = tlm; s.sstride = qstride;
    s.gout0 = vt; s.gout1 = vp; s.gstride = gstride;
    return run_jobs(h, &s, 1, true);
}

int aer_anal_vec_d(aer_sht_t h, int nfield, const double *vt, const double *vp, long gstride,
                   double *slm, double *tlm, long qstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::ANA_VECTOR, nfield};
    s.gin0 = vt; s.gin1 = vp; s.gstride = gstride;
    s.sout0 = slm; s.sout1 = tlm; s.sstride = qstride;
    s.mode0 = s.mode1 = aer::M_INV_LL1;
    return run_jobs(h, &s, 1, false);
}

int aer_synth_grad_d(aer_sht_t h, int nfield, const double *slm, long qstride,
                     double *gt, double *gp, long gstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::SYN_GRAD, nfield};
    s.sin0 = slm; s.sstride = qstride; s.gout0 = gt; s.gout1 = gp; s.gstride = gstride;
    return run_jobs(h, &s, 1, true);
}

// ---- adjoint (transpose) transforms

int aer_synth_adj_d(aer_sht_t h, int nfield, const double *vr, long gstride, double *qlm, long qstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::ANA_SCALAR, nfield};
    s.gin0 = vr; s.gstride = gstride; s.sout0 = qlm; s.sstride = qstride;
    s.adj = 1;
    return run_jobs(h, &s, 1, false);
}

int aer_anal_adj_d(aer_sht_t h, int nfield, const double *qlm, long qstride, double *vr, long gstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::SYN_SCALAR, nfield};
    s.sin0 = qlm; s.sstride = qstride; s.gout0 = vr; s.gstride = gstride;
    s.adj = 1;
    return run_jobs(h, &s, 1, true);
}

int aer_synth_vec_adj_d(aer_sht_t h, int nfield, const double *vt, const double *vp, long gstride,
                        double *slm, double *tlm, long qstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    if (!slm || !tlm) return fail(AER_EINVAL, "aer_synth_vec_adj_d: missing output");
    JobSpec s{aer::ANA_VECTOR, nfield};
    s.gin0 = vt; s.gin1 = vp; s.gstride = gstride;
    s.sout0 = slm; s.sout1 = tlm; s.sstride = qstride;
    s.adj = 1;
    return run_jobs(h, &s, 1, false);
}

int aer_anal_vec_adj_d(aer_sht_t h, int nfield, const double *slm, const double *tlm, long qstride,
                       double *vt, double *vp, long gstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    JobSpec s{aer::SYN_VECTOR, nfield};
    s.sin0 = slm; s.sin1 = tlm; s.sstride = qstride;
    s.gout0 = vt; s.gout1 = vp; s.gstride = gstride;
    s.mode0 = s.mode1 = aer::M_INV_LL1;
    s.adj = 1;
    return run_jobs(h, &s, 1, true);
}

int aer_synth_grad_adj_d(aer_sht_t h, int nfield, const double *gt, const double *gp, long gstride,
                         double *slm, long qstride)
{
    int rc = check_batch(h, nfield, qstride, gstride);
    if (rc) return rc;
    if (!slm) return fail(AER_EINVAL, "aer_synth_grad_adj_d: missing output");
    JobSpec s{aer::ANA_VECTOR, nfield};
    s.gin0 = gt; s.gin1 = gp; s.gstride = gstride;
    s.sout0 = slm; s.sout1 = nullptr; s.sstride = qstride;
    s.adj = 1;
    return run_jobs(h, &s, 1, false);
}

// ---- aeros_vordiv

int aer_uv_from_vordiv_d(aer_sht_t h, int nlev, const double *vor, const double *div, double *u, double *v)
{
    int rc = check_batch(h, nlev, h ? dense_s(h) : 0, h ? dense_g(h) : 0);
    if (rc) return rc;
    JobSpec s{aer::SYN_VECTOR, nlev};
    s.sin0 = div; s.mode0 = aer::M_S_FROM_DIV;
    s.sin1 = vor; s.mode1 = aer::M_T_FROM_VOR;
    s.sstride = dense_s(h);
    s.gout0 = v; s.gsc0 = -1.0;  // v = -V_theta
    s.gout1 = u; s.gsc1 = 1.0;   // u = V_phi
    s.gstride = dense_g(h);
    return run_jobs(h, &s, 1, true);
}

int aer_vordiv_from_uv_d(aer_sht_t h, int nlev, const double *u, const double *v, double *vor, double *div)
{
    int rc = check_batch(h, nlev, h ? dense_s(h) : 0, h ? dense_g(h) : 0);
    if (rc) return rc;
    JobSpec s{aer::ANA_VECTOR, nlev};
    s.gin0 = v; s.gsc0 = -1.0;   // V_theta = -v
    s.gin1 = u; s.gsc1 = 1.0;    // V_phi = u
    s.gstride = dense_g(h);
    s.sout0 = div; s.mode0 = aer::M_DIV_FROM_S;
    s.sout1 = vor; s.mode1 = aer::M_VOR_FROM_T;
    s.sstride = dense_s(h);
    return run_jobs(h, &s, 1, false);
}

int aer_gradient_d(aer_sht_t h, int nlev, const double *f, double *dfdx, double *dfdy)
{
    int rc = check_batch(h, nlev, h ? dense_s(h) : 0, h ? dense_g(h) : 0);
    if (rc) return rc;
    JobSpec s{aer::SYN_GRAD, nlev};
    s.sin0 = f; s.sstride = dense_s(h);
    s.gout0 = dfdy; s.gsc0 = -1.0 / kREarth;  // north = -g_theta/a
    s.gout1 = dfdx; s.gsc1 = 1.0 / kREarth;   // east  =  g_phi/a
    s.gstride = dense_g(h);
    return run_jobs(h, &s, 1, true);
}

// ---- tendency transform halves

int aer_tendency_grid_transforms_d(aer_sht_t h, int nlev,
        const double *vor, const double *div, const double *temp, const double *lnps,
        double *u, double *v, double *vor_g, double *div_g, double *t_g,
        double *dtdx, double *dtdy, double *lnps_g, double *dlnpsdx, double *dlnpsdy)
{
    return aer_tendency_grid_transforms_batch_d(h, 1, nlev, vor, div, temp, lnps, u, v, vor_g, div_g, t_g,
                                                dtdx, dtdy, lnps_g, dlnpsdx, dlnpsdy);
}

int aer_tendency_grid_transforms_batch_d(aer_sht_t h, int nmem, int nlev,
        const double *vor, const double *div, const double *temp, const double *lnps,
        double *u, double *v, double *vor_g, double *div_g, double *t_g,
        double *dtdx, double *dtdy, double *lnps_g, double *dlnpsdx, double *dlnpsdy)
{
    if (nmem < 1 || nlev < 1 || (long)nmem * nlev > (1L << 30))
        return fail(AER_EINVAL, "aer_tendency_grid_transforms_batch_d: bad member or level count");
    int rc = check_batch(h, nmem * nlev, h ? dense_s(h) : 0, h ? dense_g(h) : 0);
    if (rc) return rc;
    const long ss = dense_s(h), gs = dense_g(h);
    const int nf = nmem * nlev;
    JobSpec j[7] = {{aer::SYN_SCALAR, nmem}, {aer::SYN_GRAD, nmem}, {aer::SYN_VECTOR, nf},
                    {aer::SYN_SCALAR, nf}, {aer::SYN_SCALAR, nf}, {aer::SYN_SCALAR, nf},
                    {aer::SYN_GRAD, nf}};
    for (auto &s : j) { s.sstride = ss; s.gstride = gs; }
    j[0].sin0 = lnps; j[0].gout0 = lnps_g;
    j[1].sin0 = lnps; j[1].gout0 = dlnpsdy; j[1].gsc0 = -1.0 / kREarth; j[1].gout1 = dlnpsdx; j[1].gsc1 = 1.0 / kREarth;
    j[2].sin0 = div; j[2].mode0 = aer::M_S_FROM_DIV; j[2].sin1 = vor; j[2].mode1 = aer::M_T_FROM_VOR;
    j[2].gout0 = v; j[2].gsc0 = -1.0; j[2].gout1 = u;
    j[3].sin0 = vor; j[3].gout0 = vor_g;
    j[4].sin0 = div; j[4].gout0 = div_g;
    j[5].sin0 = temp; j[5].gout0 = t_g;
    j[6].sin0 = temp; j[6].gout0 = dtdy; j[6].gsc0 = -1.0 / kREarth; j[6].gout1 = dtdx; j[6].gsc1 = 1.0 / kREarth;
    return run_jobs(h, j, 7, true);
}

int aer_tendency_spectral_d(aer_sht_t h, int nlev,
        const double *ae, const double *an, const double *ek, const double *dtdt,
        const double *dlnpsdt, double *tvor, double *tdiv, double *ttemp, double *tlnps)
{
    return aer_tendency_spectral_batch_d(h, 1, nlev, ae, an, ek, dtdt, dlnpsdt, tvor, tdiv, ttemp, tlnps);
}

int aer_tendency_spectral_batch_d(aer_sht_t h, int nmem, int nlev,
        const double *ae, const double *an, const double *ek, const double *dtdt,
        const double *dlnpsdt, double *tvor, double *tdiv, double *ttemp, double *tlnps)
{
    if (nmem < 1 || nlev < 1 || (long)nmem * nlev > (1L << 30))
        return fail(AER_EINVAL, "aer_tendency_spectral_batch_d: bad member or level count");
    int rc = check_batch(h, nmem * nlev, h ? dense_s(h) : 0, h ? dense_g(h) : 0);
    if (rc) return rc;
    const long ss = dense_s(h), gs = dense_g(h);
    const int nf = nmem * nlev;
    JobSpec j[3] = {{aer::ANA_VECTOR, nf}, {aer::ANA_SCALAR, nf}, {aer::ANA_SCALAR, nmem}};
    for (auto &s : j) { s.sstride = ss; s.gstride = gs; }
    j[0].gin0 = an; j[0].gsc0 = -1.0; j[0].gin1 = ae; j[0].gin2 = ek;
    j[0].sout0 = tdiv; j[0].mode0 = aer::M_DIV_FROM_S;
    j[0].sout1 = tvor; j[0].mode1 = aer::M_VOR_FROM_T;
    j[0].mode2 = aer::M_NEG_LAP;
    j[1].gin0 = dtdt; j[1].sout0 = ttemp;
    j[2].gin0 = dlnpsdt; j[2].sout0 = tlnps;
    return run_jobs(h, j, 3, false);
}

// ---- per-call host API (one field, synchronous)

int aer_SH_to_spat(aer_sht_t h, const double *qlm, double *vr)
{
    if (!h || !qlm || !vr) return fail(AER_EINVAL, "aer_SH_to_spat: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    int rc = stage(h, 1, 1);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(h->stage_s.p, qlm, dense_s(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = aer_synth_d(h, 1, (const double *)h->stage_s.p, dense_s(h), (double *)h->stage_g.p, dense_g(h)))) return rc;
    HIP_TRY(hipMemcpyAsync(vr, h->stage_g.p, dense_g(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return AER_OK;
}

int aer_spat_to_SH(aer_sht_t h, double *vr, double *qlm)
{
    if (!h || !qlm || !vr) return fail(AER_EINVAL, "aer_spat_to_SH: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    int rc = stage(h, 1, 1);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(h->stage_g.p, vr, dense_g(h) * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = aer_anal_d(h, 1, (const double *)h->stage_g.p, dense_g(h), (double *)h->stage_s.p, dense_s(h)))) return rc;
    HIP_TRY(hipMemcpyAsync(qlm, h->stage_s.p, dense_s(h) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return AER_OK;
}

int aer_SHsphtor_to_spat(aer_sht_t h, const double *slm, const double *tlm, double *vt, double *vp)
{
    if (!h || !slm || !tlm || !vt || !vp) return fail(AER_EINVAL, "aer_SHsphtor_to_spat: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    int rc = stage(h, 2, 2);
    if (rc) return rc;
    double *ds = (double *)h->stage_s.p, *dg = (double *)h->stage_g.p;
    const long ss = dense_s(h), gs = dense_g(h);
    HIP_TRY(hipMemcpyAsync(ds, slm, ss * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(ds + ss, tlm, ss * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = aer_synth_vec_d(h, 1, ds, ds + ss, ss, dg, dg + gs, gs))) return rc;
    HIP_TRY(hipMemcpyAsync(vt, dg, gs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(vp, dg + gs, gs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return AER_OK;
}

int aer_spat_to_SHsphtor(aer_sht_t h, double *vt, double *vp, double *slm, double *tlm)
{
    if (!h || !slm || !tlm || !vt || !vp) return fail(AER_EINVAL, "aer_spat_to_SHsphtor: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    int rc = stage(h, 2, 2);
    if (rc) return rc;
    double *ds = (double *)h->stage_s.p, *dg = (double *)h->stage_g.p;
    const long ss = dense_s(h), gs = dense_g(h);
    HIP_TRY(hipMemcpyAsync(dg, vt, gs * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(dg + gs, vp, gs * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = aer_anal_vec_d(h, 1, dg, dg + gs, gs, ds, ds + ss, ss))) return rc;
    HIP_TRY(hipMemcpyAsync(slm, ds, ss * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(tlm, ds + ss, ss * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return AER_OK;
}

int aer_SHsph_to_spat(aer_sht_t h, const double *slm, double *vt, double *vp)
{
    if (!h || !slm || !vt || !vp) return fail(AER_EINVAL, "aer_SHsph_to_spat: null argument");
    if (!h->grid_set) return fail(AER_ESTATE, "transform before aer_sht_set_grid");
    int rc = stage(h, 1, 2);
    if (rc) return rc;
    double *ds = (double *)h->stage_s.p, *dg = (double *)h->stage_g.p;
    const long ss = dense_s(h), gs = dense_g(h);
    HIP_TRY(hipMemcpyAsync(ds, slm, ss * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if ((rc = aer_synth_grad_d(h, 1, ds, ss, dg, dg + gs, gs))) return rc;
    HIP_TRY(hipMemcpyAsync(vt, dg, gs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(vp, dg + gs, gs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return AER_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ column work

struct aer_vgrid {
    int nlev = 0;
    bool ztop = false;
    std::vector<double> A, B;   // host copies
    DevBuf ab;   // A then B, nlev+1 each
};

extern "C" {

int aer_vgrid_create(int nlev, const double *a_half, const double *b_half, aer_vgrid_t *vg)
{
    if (!vg) return fail(AER_EINVAL, "aer_vgrid_create: null output");
    *vg = nullptr;
    if (nlev < 2 || nlev > AER_COL_MAXLEV || !a_half || !b_half)
        return fail(AER_EINVAL, "aer_vgrid_create: need 2 <= nlev <= 64 and both coefficient arrays");
    for (int j = 0; j < nlev; ++j)
        if (!(a_half[j + 1] + b_half[j + 1] * 1.0e5 > a_half[j] + b_half[j] * 1.0e5))
            return fail(AER_EINVAL, "aer_vgrid_create: half-level pressures must increase downward");
    std::vector<double> ab(2 * (nlev + 1));
    std::copy(a_half, a_half + nlev + 1, ab.begin());
    std::copy(b_half, b_half + nlev + 1, ab.begin() + nlev + 1);
    aer_vgrid *g = new aer_vgrid();
    g->nlev = nlev;
    g->ztop = a_half[0] <= 0.0 && b_half[0] <= 0.0;   // src/dynamics/aeros_tendency.f90:499
    g->A.assign(a_half, a_half + nlev + 1);
    g->B.assign(b_half, b_half + nlev + 1);
    int rc = upload(g->ab, ab.data(), ab.size() * sizeof(double));
    if (rc) {
        delete g;
        return rc;
    }
    *vg = g;
    return AER_OK;
}

int aer_vgrid_destroy(aer_vgrid_t vg)
{
    if (!vg) return AER_OK;
    vg->ab.release();
    delete vg;
    return AER_OK;
}

}  // extern "C"

namespace {

int column_args(aer_sht *h, aer_vgrid *vg, int held_suarez, const double *u, const double *v, const double *vor_g,
                const double *div_g, const double *t_g, const double *dtdx, const double *dtdy, const double *lnps_g,
                const double *dlnpsdx, const double *dlnpsdy, const double *phis, const double *dteq,
                aer::ColumnArgs &C, const char *who)
{
    if (!h || !vg) return fail(AER_EINVAL, std::string(who) + ": null handle");
    if (!h->grid_set) return fail(AER_ESTATE, std::string(who) + ": handle without a grid");
    if (!u || !v || !vor_g || !div_g || !t_g || !dtdx || !dtdy || !lnps_g || !dlnpsdx || !dlnpsdy)
        return fail(AER_EINVAL, std::string(who) + ": missing array");
    C = aer::ColumnArgs{};
    C.nlon = h->nlon;
    C.ncol = h->nlon * h->nlat;
    C.nlev = vg->nlev;
    C.hs = held_suarez ? 1 : 0;
    C.ztop = vg->ztop ? 1 : 0;
    C.A = (const double *)vg->ab.p;
    C.B = C.A + vg->nlev + 1;
    C.sinlat = (const double *)h->sinlat.p;
    C.u = u; C.v = v; C.vor = vor_g; C.div = div_g; C.t = t_g; C.dtdx = dtdx; C.dtdy = dtdy;
    C.lnps = lnps_g; C.dlnpsdx = dlnpsdx; C.dlnpsdy = dlnpsdy; C.phis = phis;
    C.dteq = dteq;
    return AER_OK;
}

// Member strides of ensemble arrays: [nmem][nlev][ncol], [nmem][ncol], dteq [nmem][nlev][nlat]
// (or shared when dteq_mstride is 0).
int column_members(aer::ColumnArgs &C, int nmem, long dteq_mstride, const char *who)
{
    if (nmem < 1 || nmem > 65535) return fail(AER_EINVAL, std::string(who) + ": member count outside 1..65535");
    if (dteq_mstride < 0) return fail(AER_EINVAL, std::string(who) + ": negative dteq member stride");
    C.ms3 = (long)C.nlev * C.ncol;
    C.ms2 = C.ncol;
    C.msq = dteq_mstride;
    return AER_OK;
}

int column_launch(aer_sht *h, aer::ColumnArgs &C, int nmem, double *ae, double *an, double *ek, double *dtdt,
                  double *dlnpsdt, const char *who)
{
    if (!ae || !an || !ek || !dtdt || !dlnpsdt) return fail(AER_EINVAL, std::string(who) + ": missing array");
    C.ae = ae; C.an = an; C.ek = ek; C.dtdt = dtdt; C.dlnpsdt = dlnpsdt;
    const int nb = (C.ncol + AER_COL_CB - 1) / AER_COL_CB;
    const size_t lds = (size_t)(2 * (C.nlev + 1) + (7 * C.nlev + 1) * AER_COL_CB) * sizeof(double);
    // levels per thread: the smallest of 2, 3, 4, 8 that covers nlev over the block's level lanes
    const dim3 grid(nb, nmem), block(AER_COL_CB * AER_COL_KL);
    if (C.nlev <= 2 * AER_COL_KL)
        hipLaunchKernelGGL(aer::k_column_terms<2>, grid, block, lds, h->stream, C);
    else if (C.nlev <= 3 * AER_COL_KL)
        hipLaunchKernelGGL(aer::k_column_terms<3>, grid, block, lds, h->stream, C);
    else if (C.nlev <= 4 * AER_COL_KL)
        hipLaunchKernelGGL(aer::k_column_terms<4>, grid, block, lds, h->stream, C);
    else
        hipLaunchKernelGGL(aer::k_column_terms<AER_COL_MAXLEV / AER_COL_KL>, grid, block, lds, h->stream, C);
    HIP_TRY(hipGetLastError());
    return AER_OK;
}

}  // namespace

extern "C" {

int aer_column_terms_d(aer_sht_t h, aer_vgrid_t vg, int held_suarez,
                       const double *u, const double *v, const double *vor_g, const double *div_g,
                       const double *t_g, const double *dtdx, const double *dtdy,
                       const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                       const double *phis,
                       double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt)
{
    return aer_column_terms_batch_d(h, vg, 1, held_suarez, nullptr, 0, u, v, vor_g, div_g, t_g, dtdx, dtdy,
                                    lnps_g, dlnpsdx, dlnpsdy, phis, ae, an, ek, dtdt, dlnpsdt);
}

int aer_column_terms_batch_d(aer_sht_t h, aer_vgrid_t vg, int nmem, int held_suarez,
                             const double *dteq, long dteq_mstride,
                             const double *u, const double *v, const double *vor_g, const double *div_g,
                             const double *t_g, const double *dtdx, const double *dtdy,
                             const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                             const double *phis,
                             double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt)
{
    aer::ColumnArgs C;
    if (dteq && !held_suarez) return fail(AER_EINVAL, "aer_column_terms_batch_d: dteq without the Held-Suarez forcing");
    int rc = column_args(h, vg, held_suarez, u, v, vor_g, div_g, t_g, dtdx, dtdy, lnps_g, dlnpsdx, dlnpsdy, phis,
                         dteq, C, "aer_column_terms_batch_d");
    if (rc) return rc;
    if ((rc = column_members(C, nmem, dteq_mstride, "aer_column_terms_batch_d"))) return rc;
    return column_launch(h, C, nmem, ae, an, ek, dtdt, dlnpsdt, "aer_column_terms_batch_d");
}

int aer_column_terms_dteq_d(aer_sht_t h, aer_vgrid_t vg, const double *dteq,
                            const double *u, const double *v, const double *vor_g, const double *div_g,
                            const double *t_g, const double *dtdx, const double *dtdy,
                            const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                            const double *phis,
                            double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt)
{
    return aer_column_terms_batch_d(h, vg, 1, 1, dteq, 0, u, v, vor_g, div_g, t_g, dtdx, dtdy,
                                    lnps_g, dlnpsdx, dlnpsdy, phis, ae, an, ek, dtdt, dlnpsdt);
}

int aer_column_terms_adj_d(aer_sht_t h, aer_vgrid_t vg, int held_suarez, const double *dteq,
                           const double *u, const double *v, const double *vor_g, const double *div_g,
                           const double *t_g, const double *dtdx, const double *dtdy,
                           const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                           const double *phis,
                           const double *b_ae, const double *b_an, const double *b_ek, const double *b_dtdt,
                           const double *b_dlnpsdt,
                           double *g_u, double *g_v, double *g_vor, double *g_div, double *g_t,
                           double *g_dtdx, double *g_dtdy, double *g_lnps, double *g_dlnpsdx, double *g_dlnpsdy,
                           double *g_teq)
{
    return aer_column_terms_adj_batch_d(h, vg, 1, held_suarez, dteq, 0, u, v, vor_g, div_g, t_g, dtdx, dtdy,
                                        lnps_g, dlnpsdx, dlnpsdy, phis, b_ae, b_an, b_ek, b_dtdt, b_dlnpsdt,
                                        g_u, g_v, g_vor, g_div, g_t, g_dtdx, g_dtdy, g_lnps, g_dlnpsdx, g_dlnpsdy,
                                        g_teq);
}

int aer_column_terms_adj_batch_d(aer_sht_t h, aer_vgrid_t vg, int nmem, int held_suarez,
                                 const double *dteq, long dteq_mstride,
                                 const double *u, const double *v, const double *vor_g, const double *div_g,
                                 const double *t_g, const double *dtdx, const double *dtdy,
                                 const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                                 const double *phis,
                                 const double *b_ae, const double *b_an, const double *b_ek, const double *b_dtdt,
                                 const double *b_dlnpsdt,
                                 double *g_u, double *g_v, double *g_vor, double *g_div, double *g_t,
                                 double *g_dtdx, double *g_dtdy, double *g_lnps, double *g_dlnpsdx,
                                 double *g_dlnpsdy, double *g_teq)
{
    aer::ColumnAdjArgs X{};
    if (dteq && !held_suarez) return fail(AER_EINVAL, "aer_column_terms_adj_batch_d: dteq without the Held-Suarez forcing");
    int rc = column_args(h, vg, held_suarez, u, v, vor_g, div_g, t_g, dtdx, dtdy, lnps_g, dlnpsdx, dlnpsdy, phis,
                         dteq, X.f, "aer_column_terms_adj_batch_d");
    if (rc) return rc;
    if ((rc = column_members(X.f, nmem, dteq_mstride, "aer_column_terms_adj_batch_d"))) return rc;
    if (vg->nlev > AER_ADJ_MAXLEV) return fail(AER_EINVAL, "aer_column_terms_adj_batch_d: nlev above 64");
    if (!b_ae || !b_an || !b_ek || !b_dtdt || !b_dlnpsdt || !g_u || !g_v || !g_vor || !g_div || !g_t || !g_dtdx ||
        !g_dtdy || !g_lnps || !g_dlnpsdx || !g_dlnpsdy)
        return fail(AER_EINVAL, "aer_column_terms_adj_batch_d: missing array");
    X.dteq = dteq;
    X.bae = b_ae; X.ban = b_an; X.bek = b_ek; X.bdtdt = b_dtdt; X.bdlnpsdt = b_dlnpsdt;
    X.gu = g_u; X.gv = g_v; X.gvor = g_vor; X.gdiv = g_div; X.gt = g_t; X.gdtdx = g_dtdx; X.gdtdy = g_dtdy;
    X.glnps = g_lnps; X.gdlnpsdx = g_dlnpsdx; X.gdlnpsdy = g_dlnpsdy; X.gteq = g_teq;
    const int nb = (X.f.ncol + 63) / 64;
    hipLaunchKernelGGL(aer::k_column_terms_adj, dim3(nb, nmem), dim3(64), 0, h->stream, X);
    HIP_TRY(hipGetLastError());
    return AER_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ step completion

struct aer_semiimp {
    aer_sht *h = nullptr;
    int nlev = 0, lmax = 0, nblk = 0;
    double alpha = 0.5, dt = 0.0;
    std::vector<double> G, tau, W, nu, pcoef;        // host copies, row-major
    bool diff = false;
    double tau_diff = 0.0, tau_div = 0.0;
    int ndiff = 0;
    DevBuf dG, dtau, dW, dnu, dpcoef, dMinv, dfvt, dfdiv, dtl;   // dtl: degree of each lm
};

namespace {

// Doolittle LU with partial pivoting in place, row-major (the factorization of
// reference src/dynamics/aeros_semiimp.f90:658-713).
bool lu_factor(std::vector<double> &a, std::vector<int> &piv, int n)
{
    for (int k = 0; k < n; ++k) {
        int ip = k;
        double amax = std::fabs(a[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a[i * n + k]) > amax) {
                amax = std::fabs(a[i * n + k]);
                ip = i;
            }
        if (amax <= 0.0) return false;
        piv[k] = ip;
        if (ip != k)
            for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[ip * n + j]);
        for (int i = k + 1; i < n; ++i) {
            a[i * n + k] = a[i * n + k] / a[k * n + k];
            for (int j = k + 1; j < n; ++j) a[i * n + j] = a[i * n + j] - a[i * n + k] * a[k * n + j];
        }
    }
    return true;
}

void lu_solve(const std::vector<double> &a, const std::vector<int> &piv, double *b, int n)
{
    for (int k = 0; k < n; ++k) {
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int j = k + 1; j < n; ++j) b[j] = b[j] - a[j * n + k] * b[k];
    }
    for (int k = n - 1; k >= 0; --k) {
        for (int j = k + 1; j < n; ++j) b[k] = b[k] - a[k * n + j] * b[j];
        b[k] = b[k] / a[k * n + k];
    }
}

// Factors of the degree-l Helmholtz matrices for the current step, inverted column
// by column through the LU factors, and the diffusion factors for that step.
int si_build_step(aer_semiimp *s)
{
    const int n = s->nlev;
    const double hh = s->alpha * s->alpha * s->dt * s->dt, a2 = kREarth * kREarth;
    std::vector<double> minv((size_t)(s->lmax + 1) * n * n), a((size_t)n * n), col(n);
    std::vector<int> piv(n);
    for (int l = 0; l <= s->lmax; ++l) {
        const double lam = (double)(l * (l + 1)) / a2;
        for (int j = 0; j < n; ++j) {
            for (int k = 0; k < n; ++k) a[k * n + j] = hh * lam * s->W[k * n + j];
            a[j * n + j] = a[j * n + j] + 1.0;
        }
        if (!lu_factor(a, piv, n)) return fail(AER_EINVAL, "aer_semiimp: singular Helmholtz matrix");
        double *mi = minv.data() + (size_t)l * n * n;
        for (int j = 0; j < n; ++j) {
            std::fill(col.begin(), col.end(), 0.0);
            col[j] = 1.0;
            lu_solve(a, piv, col.data(), n);
            for (int k = 0; k < n; ++k) mi[k * n + j] = col[k];
        }
    }
    int rc = upload(s->dMinv, minv.data(), minv.size() * sizeof(double));
    if (rc || !s->diff) return rc;
    // implicit del^ndiff factors (reference src/dynamics/aeros_timestep.f90:560-563, 1673-1726)
    const double lref = (double)(s->lmax * (s->lmax + 1));
    const double rate = s->dt / s->tau_diff, rate_div = s->dt / (s->tau_div > 0.0 ? s->tau_div : s->tau_diff);
    std::vector<double> fvt((size_t)n * (s->lmax + 1)), fdiv(fvt.size());
    for (int l = 0; l <= s->lmax; ++l) {
        const double x = (double)(l * (l + 1)) / lref;
        double r = s->ndiff / 2 > 0 ? x : 1.0;
        for (int i = 1; i < s->ndiff / 2; ++i) r = r * x;
        for (int k = 0; k < n; ++k) {
            fvt[(size_t)k * (s->lmax + 1) + l] = 1.0 / (1.0 + 1.0 * rate * r);
            fdiv[(size_t)k * (s->lmax + 1) + l] = 1.0 / (1.0 + 1.0 * rate_div * r);
        }
    }
    if ((rc = upload(s->dfvt, fvt.data(), fvt.size() * sizeof(double)))) return rc;
    return upload(s->dfdiv, fdiv.data(), fdiv.size() * sizeof(double));
}

int launch_step(aer_semiimp *s, int flags, double eps, double raw_alpha,
                const double *ov, const double *od, const double *ot, const double *op,
                double *cv, double *cd, double *ct, double *cp,
                const double *tv, const double *td, const double *tt, const double *tp,
                double *nv, double *nd, double *nt, double *np, int nmem = 1)
{
    if (nmem < 1 || nmem > 65535) return fail(AER_EINVAL, "step: member count outside 1..65535");
    aer::StepArgs A{};
    A.ms3 = 2L * s->h->nlm * s->nlev;
    A.ms2 = 2L * s->h->nlm;
    A.nlev = s->nlev;
    A.lmax = s->lmax;
    A.nlm = s->h->nlm;
    A.flags = flags;
    A.h = s->dt;
    A.a1 = s->alpha;
    A.astep = s->alpha * s->dt;
    A.hh_old = s->alpha * (1.0 - s->alpha) * s->dt * s->dt;
    A.half_nu = 0.5 * eps;
    A.raw_a = raw_alpha;
    A.raw_b = 1.0 - raw_alpha;
    A.G = (const double *)s->dG.p;
    A.tau = (const double *)s->dtau.p;
    A.W = (const double *)s->dW.p;
    A.nu = (const double *)s->dnu.p;
    A.pcoef = (const double *)s->dpcoef.p;
    A.Minv = (const double *)s->dMinv.p;
    A.fvt = (const double *)s->dfvt.p;
    A.fdiv = (const double *)s->dfdiv.p;
    A.l_of_lm = (const int *)s->dtl.p;
    A.ovor = ov; A.odiv = od; A.otemp = ot; A.olnps = op;
    A.cvor = cv; A.cdiv = cd; A.ctemp = ct; A.clnps = cp;
    A.tvor = tv; A.tdiv = td; A.ttemp = tt; A.tlnps = tp;
    A.nvor = nv; A.ndiv = nd; A.ntemp = nt; A.nlnps = np;
    const size_t lds = (size_t)(11 * s->nlev + 5) * AER_SI_CB * sizeof(double);
    // stage G, tau, W, nu, pcoef in LDS when the block still fits in 80 KB (two blocks per CU)
    const size_t ops = (size_t)(3 * s->nlev * s->nlev + 2 * s->nlev) * sizeof(double);
    const dim3 grid(s->nblk, nmem), block(AER_SI_CB * AER_SI_KL);
    if (lds + ops <= 80 * 1024)
        hipLaunchKernelGGL(aer::k_step<true>, grid, block, lds + ops, s->h->stream, A);
    else
        hipLaunchKernelGGL(aer::k_step<false>, grid, block, lds, s->h->stream, A);
    HIP_TRY(hipGetLastError());
    return AER_OK;
}

}  // namespace

extern "C" {

int aer_semiimp_create(aer_sht_t h, aer_vgrid_t vg, double t_ref, double ps_ref, double dt_step,
                       double alpha, aer_semiimp_t *out)
{
    if (!out) return fail(AER_EINVAL, "aer_semiimp_create: null output");
    *out = nullptr;
    if (!h || !vg) return fail(AER_EINVAL, "aer_semiimp_create: null handle");
    if (!h->grid_set) return fail(AER_ESTATE, "aer_semiimp_create: handle without a grid");
    if (alpha < 0.5 || alpha > 1.0) return fail(AER_EINVAL, "aer_semiimp_create: alpha must be in [0.5, 1]");
    if (dt_step <= 0.0 || t_ref <= 0.0 || ps_ref <= 0.0) return fail(AER_EINVAL, "aer_semiimp_create: step, t_ref and ps_ref must be > 0");
    const int n = vg->nlev;
    aer_semiimp *s = new aer_semiimp();
    s->h = h;
    s->nlev = n;
    s->lmax = h->lmax;
    s->alpha = alpha;
    s->dt = dt_step;
    // reference column at ps_ref (src/dynamics/aeros_semiimp.f90:222-275)
    std::vector<double> ph(n + 1), dp(n), al(n), dl(n);
    for (int j = 0; j <= n; ++j) ph[j] = vg->A[j] + vg->B[j] * ps_ref;
    for (int k = 0; k < n; ++k) {
        dp[k] = ph[k + 1] - ph[k];
        if (ph[k] <= 0.0) {
            al[k] = std::log(2.0);
            dl[k] = aer::kMV;
        } else {
            const double lr = std::log(ph[k + 1] / ph[k]);
            al[k] = 1.0 - (ph[k] / dp[k]) * lr;
            dl[k] = lr;
        }
    }
    if (vg->ztop) dl[0] = 0.0;
    s->G.assign((size_t)n * n, 0.0);
    s->tau.assign((size_t)n * n, 0.0);
    s->W.assign((size_t)n * n, 0.0);
    s->nu.assign(n, 0.0);
    s->pcoef.assign(n, 0.0);
    for (int k = 0; k < n; ++k) {
        s->G[k * n + k] = aer::kRd * al[k];
        for (int j = k + 1; j < n; ++j) s->G[k * n + j] = aer::kRd * dl[j];
        s->tau[k * n + k] = aer::kKappa * t_ref * al[k];
        for (int j = 0; j < k; ++j) s->tau[k * n + j] = aer::kKappa * t_ref * dl[k] * dp[j] / dp[k];
        s->nu[k] = dp[k] / ps_ref;
        s->pcoef[k] = aer::kRd * t_ref;
    }
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            double w = s->pcoef[k] * s->nu[j];
            for (int m = 0; m < n; ++m) w = w + s->G[k * n + m] * s->tau[m * n + j];
            s->W[k * n + j] = w;
        }
    // degree of each m-major spectral index; a block takes 16 consecutive real entries
    std::vector<int> tl;
    for (int m = 0; m <= s->lmax; ++m)
        for (int l = m; l <= s->lmax; ++l) tl.push_back(l);
    if ((int)tl.size() != s->h->nlm) {
        aer_semiimp_destroy(s);
        return fail(AER_EINVAL, "aer_semiimp_create: the step needs a triangular truncation (mmax = lmax, mres = 1)");
    }
    s->nblk = (int)((2L * s->h->nlm + AER_SI_CB - 1) / AER_SI_CB);
    int rc = 0;
    if (!rc) rc = upload(s->dG, s->G.data(), s->G.size() * sizeof(double));
    if (!rc) rc = upload(s->dtau, s->tau.data(), s->tau.size() * sizeof(double));
    if (!rc) rc = upload(s->dW, s->W.data(), s->W.size() * sizeof(double));
    if (!rc) rc = upload(s->dnu, s->nu.data(), s->nu.size() * sizeof(double));
    if (!rc) rc = upload(s->dpcoef, s->pcoef.data(), s->pcoef.size() * sizeof(double));
    if (!rc) rc = upload(s->dtl, tl.data(), tl.size() * sizeof(int));
    if (!rc) rc = si_build_step(s);
    if (rc) {
        aer_semiimp_destroy(s);
        return rc;
    }
    *out = s;
    return AER_OK;
}

int aer_semiimp_set_step(aer_semiimp_t s, double dt_step)
{
    if (!s) return fail(AER_EINVAL, "aer_semiimp_set_step: null object");
    if (dt_step <= 0.0) return fail(AER_EINVAL, "aer_semiimp_set_step: step must be > 0");
    s->dt = dt_step;
    return si_build_step(s);
}

int aer_semiimp_set_diffusion(aer_semiimp_t s, double tau_diff, double tau_diff_div, int ndiff)
{
    if (!s) return fail(AER_EINVAL, "aer_semiimp_set_diffusion: null object");
    if (tau_diff <= 0.0 || ndiff < 2 || ndiff % 2) return fail(AER_EINVAL, "aer_semiimp_set_diffusion: need tau_diff > 0 and an even ndiff >= 2");
    s->diff = true;
    s->tau_diff = tau_diff;
    s->tau_div = tau_diff_div;
    s->ndiff = ndiff;
    return si_build_step(s);
}

int aer_semiimp_gwspeed(aer_semiimp_t s, double *cmax)
{
    if (!s || !cmax) return fail(AER_EINVAL, "aer_semiimp_gwspeed: null argument");
    const int n = s->nlev;
    std::vector<double> x(n, 1.0 / std::sqrt((double)n)), y(n);
    double lam = 0.0;
    for (int it = 0; it < 200; ++it) {
        for (int k = 0; k < n; ++k) {
            y[k] = 0.0;
            for (int j = 0; j < n; ++j) y[k] = y[k] + s->W[k * n + j] * x[j];
        }
        double nrm = 0.0;
        for (int k = 0; k < n; ++k) nrm += y[k] * y[k];
        nrm = std::sqrt(nrm);
        if (nrm <= 0.0) break;
        for (int k = 0; k < n; ++k) x[k] = y[k] / nrm;
        lam = 0.0;
        for (int k = 0; k < n; ++k) lam += x[k] * y[k];
    }
    *cmax = std::sqrt(std::max(lam, 0.0));
    return AER_OK;
}

int aer_semiimp_destroy(aer_semiimp_t s)
{
    if (!s) return AER_OK;
    for (DevBuf *b : {&s->dG, &s->dtau, &s->dW, &s->dnu, &s->dpcoef, &s->dMinv, &s->dfvt, &s->dfdiv, &s->dtl})
        b->release();
    delete s;
    return AER_OK;
}

int aer_semiimp_linear_d(aer_semiimp_t s, const double *div, const double *temp, const double *lnps,
                         double *ldiv, double *ltemp, double *lps)
{
    if (!s) return fail(AER_EINVAL, "aer_semiimp_linear_d: null object");
    if (!div || !temp || !lnps || !ldiv || !ltemp || !lps) return fail(AER_EINVAL, "aer_semiimp_linear_d: missing array");
    return launch_step(s, aer::ST_LINEAR, 0.0, 1.0, nullptr, nullptr, nullptr, nullptr,
                       nullptr, (double *)div, (double *)temp, (double *)lnps,
                       nullptr, nullptr, nullptr, nullptr, nullptr, ldiv, ltemp, lps);
}

int aer_semiimp_step_d(aer_semiimp_t s,
                       const double *old_div, const double *old_temp, const double *old_lnps,
                       const double *now_div, const double *now_temp, const double *now_lnps,
                       const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                       double *new_div, double *new_temp, double *new_lnps)
{
    if (!s) return fail(AER_EINVAL, "aer_semiimp_step_d: null object");
    if (!old_div || !old_temp || !old_lnps || !now_div || !now_temp || !now_lnps || !tnd_div || !tnd_temp ||
        !tnd_lnps || !new_div || !new_temp || !new_lnps)
        return fail(AER_EINVAL, "aer_semiimp_step_d: missing array");
    return launch_step(s, aer::ST_SEMI, 0.0, 1.0, nullptr, old_div, old_temp, old_lnps,
                       nullptr, (double *)now_div, (double *)now_temp, (double *)now_lnps,
                       nullptr, tnd_div, tnd_temp, tnd_lnps, nullptr, new_div, new_temp, new_lnps);
}

int aer_leapfrog_d(aer_semiimp_t s, int flags, double eps_filter, double raw_alpha,
                   const double *old_vor, const double *old_div, const double *old_temp, const double *old_lnps,
                   double *now_vor, double *now_div, double *now_temp, double *now_lnps,
                   const double *tnd_vor, const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                   double *new_vor, double *new_div, double *new_temp, double *new_lnps)
{
    return aer_leapfrog_batch_d(s, 1, flags, eps_filter, raw_alpha, old_vor, old_div, old_temp, old_lnps,
                                now_vor, now_div, now_temp, now_lnps, tnd_vor, tnd_div, tnd_temp, tnd_lnps,
                                new_vor, new_div, new_temp, new_lnps);
}

int aer_leapfrog_batch_d(aer_semiimp_t s, int nmem, int flags, double eps_filter, double raw_alpha,
                         const double *old_vor, const double *old_div, const double *old_temp, const double *old_lnps,
                         double *now_vor, double *now_div, double *now_temp, double *now_lnps,
                         const double *tnd_vor, const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                         double *new_vor, double *new_div, double *new_temp, double *new_lnps)
{
    if (!s) return fail(AER_EINVAL, "aer_leapfrog_d: null object");
    if (flags & ~(AER_STEP_SEMI_IMPLICIT | AER_STEP_DIFFUSE | AER_STEP_FILTER))
        return fail(AER_EINVAL, "aer_leapfrog_d: unknown flag");
    if ((flags & AER_STEP_DIFFUSE) && !s->diff)
        return fail(AER_ESTATE, "aer_leapfrog_d: diffusion requested before aer_semiimp_set_diffusion");
    if (!old_vor || !old_div || !old_temp || !old_lnps || !now_vor || !now_div || !now_temp || !now_lnps ||
        !tnd_vor || !tnd_div || !tnd_temp || !tnd_lnps || !new_vor || !new_div || !new_temp || !new_lnps)
        return fail(AER_EINVAL, "aer_leapfrog_d: missing array");
    int f = aer::ST_VOR;
    if (flags & AER_STEP_SEMI_IMPLICIT) f |= aer::ST_SEMI;
    if (flags & AER_STEP_DIFFUSE) f |= aer::ST_DIFF;
    if ((flags & AER_STEP_FILTER) && eps_filter > 0.0) f |= aer::ST_RAW;
    return launch_step(s, f, eps_filter, raw_alpha, old_vor, old_div, old_temp, old_lnps,
                       now_vor, now_div, now_temp, now_lnps, tnd_vor, tnd_div, tnd_temp, tnd_lnps,
                       new_vor, new_div, new_temp, new_lnps, nmem);
}

int aer_leapfrog_adj_d(aer_semiimp_t s, int flags, double eps_filter, double raw_alpha,
                       const double *bn_vor, const double *bn_div, const double *bn_temp, const double *bn_lnps,
                       const double *bc_vor, const double *bc_div, const double *bc_temp, const double *bc_lnps,
                       double *go_vor, double *go_div, double *go_temp, double *go_lnps,
                       double *gc_vor, double *gc_div, double *gc_temp, double *gc_lnps,
                       double *gt_vor, double *gt_div, double *gt_temp, double *gt_lnps)
{
    return aer_leapfrog_adj_batch_d(s, 1, flags, eps_filter, raw_alpha, bn_vor, bn_div, bn_temp, bn_lnps,
                                    bc_vor, bc_div, bc_temp, bc_lnps, go_vor, go_div, go_temp, go_lnps,
                                    gc_vor, gc_div, gc_temp, gc_lnps, gt_vor, gt_div, gt_temp, gt_lnps);
}

int aer_leapfrog_adj_batch_d(aer_semiimp_t s, int nmem, int flags, double eps_filter, double raw_alpha,
                             const double *bn_vor, const double *bn_div, const double *bn_temp, const double *bn_lnps,
                             const double *bc_vor, const double *bc_div, const double *bc_temp, const double *bc_lnps,
                             double *go_vor, double *go_div, double *go_temp, double *go_lnps,
                             double *gc_vor, double *gc_div, double *gc_temp, double *gc_lnps,
                             double *gt_vor, double *gt_div, double *gt_temp, double *gt_lnps)
{
    if (!s) return fail(AER_EINVAL, "aer_leapfrog_adj_d: null object");
    if (nmem < 1 || nmem > 65535) return fail(AER_EINVAL, "aer_leapfrog_adj_d: member count outside 1..65535");
    if (flags & ~(AER_STEP_SEMI_IMPLICIT | AER_STEP_DIFFUSE | AER_STEP_FILTER))
        return fail(AER_EINVAL, "aer_leapfrog_adj_d: unknown flag");
    if ((flags & AER_STEP_DIFFUSE) && !s->diff)
        return fail(AER_ESTATE, "aer_leapfrog_adj_d: diffusion requested before aer_semiimp_set_diffusion");
    const void *ptrs[] = {bn_vor, bn_div, bn_temp, bn_lnps, bc_vor, bc_div, bc_temp, bc_lnps, go_vor, go_div, go_temp,
                          go_lnps, gc_vor, gc_div, gc_temp, gc_lnps, gt_vor, gt_div, gt_temp, gt_lnps};
    for (const void *p : ptrs)
        if (!p) return fail(AER_EINVAL, "aer_leapfrog_adj_d: missing array");
    aer::StepAdjArgs A{};
    A.nlev = s->nlev;
    A.lmax = s->lmax;
    A.nlm = s->h->nlm;
    A.flags = 0;
    if (flags & AER_STEP_SEMI_IMPLICIT) A.flags |= aer::ST_SEMI;
    if (flags & AER_STEP_DIFFUSE) A.flags |= aer::ST_DIFF;
    if ((flags & AER_STEP_FILTER) && eps_filter > 0.0) A.flags |= aer::ST_RAW;
    A.h = s->dt;
    A.a1 = s->alpha;
    A.astep = s->alpha * s->dt;
    A.hh_old = s->alpha * (1.0 - s->alpha) * s->dt * s->dt;
    A.half_nu = 0.5 * eps_filter;
    A.raw_a = raw_alpha;
    A.raw_b = 1.0 - raw_alpha;
    A.G = (const double *)s->dG.p;
    A.tau = (const double *)s->dtau.p;
    A.W = (const double *)s->dW.p;
    A.nu = (const double *)s->dnu.p;
    A.pcoef = (const double *)s->dpcoef.p;
    A.Minv = (const double *)s->dMinv.p;
    A.fvt = (const double *)s->dfvt.p;
    A.fdiv = (const double *)s->dfdiv.p;
    A.l_of_lm = (const int *)s->dtl.p;
    A.bnvor = bn_vor; A.bndiv = bn_div; A.bntemp = bn_temp; A.bnlnps = bn_lnps;
    A.bcvor = bc_vor; A.bcdiv = bc_div; A.bctemp = bc_temp; A.bclnps = bc_lnps;
    A.govor = go_vor; A.godiv = go_div; A.gotemp = go_temp; A.golnps = go_lnps;
    A.gcvor = gc_vor; A.gcdiv = gc_div; A.gctemp = gc_temp; A.gclnps = gc_lnps;
    A.gtvor = gt_vor; A.gtdiv = gt_div; A.gttemp = gt_temp; A.gtlnps = gt_lnps;
    A.ms3 = 2L * s->h->nlm * s->nlev;
    A.ms2 = 2L * s->h->nlm;
    const size_t lds = (size_t)(5 * s->nlev + 3) * AER_SI_CB * sizeof(double);
    hipLaunchKernelGGL(aer::k_step_adj, dim3(s->nblk, nmem), dim3(AER_SI_CB * AER_SI_KL), lds, s->h->stream, A);
    HIP_TRY(hipGetLastError());
    return AER_OK;
}

}  // extern "C"

namespace {

int mass_impl(aer_sht *h, int nmem, double *lnps, int fix, double target, const double *targets,
              double *mass_out, double *lnr_out, const char *who)
{
    if (!h) return fail(AER_EINVAL, std::string(who) + ": null handle");
    if (!h->grid_set) return fail(AER_ESTATE, std::string(who) + ": handle without a grid");
    if (!lnps) return fail(AER_EINVAL, std::string(who) + ": missing array");
    if (nmem < 1 || nmem > 65535) return fail(AER_EINVAL, std::string(who) + ": member count outside 1..65535");
    const long ng = (long)h->nlon * h->nlat;
    // grows once per member count; a caller that captures graphs makes its first call outside capture
    int rc = h->massg.ensure((size_t)nmem * ng * sizeof(double));
    if (rc) return rc;
    JobSpec sp{aer::SYN_SCALAR, nmem};
    sp.sin0 = lnps;
    sp.sstride = 2L * h->nlm;
    sp.gout0 = (double *)h->massg.p;
    sp.gstride = ng;
    sp.prec = 0;   // global mass sums stay FP64
    if ((rc = run_jobs(h, &sp, 1, true))) return rc;
    hipLaunchKernelGGL(aer::k_mass, dim3(nmem), dim3(256), (size_t)5 * h->nlat * sizeof(double), h->stream,
                       (const double *)h->massg.p, h->nlon, h->nlat, (const double *)h->wfull.p, 9.80665,
                       fix ? 1 : 0, target, targets, h->c00, lnps, mass_out, lnr_out, ng, 2L * h->nlm);
    HIP_TRY(hipGetLastError());
    return AER_OK;
}

}  // namespace

extern "C" {

int aer_mass_d(aer_sht_t h, double *lnps, int fix, double mass_target, double *mass_out, double *lnr_out)
{
    if (fix && mass_target <= 0.0) return fail(AER_EINVAL, "aer_mass_d: mass target must be > 0");
    return mass_impl(h, 1, lnps, fix, mass_target, nullptr, mass_out, lnr_out, "aer_mass_d");
}

int aer_mass_batch_d(aer_sht_t h, int nmem, double *lnps, int fix, const double *mass_targets,
                     double *mass_out, double *lnr_out)
{
    if (fix && !mass_targets) return fail(AER_EINVAL, "aer_mass_batch_d: fix needs the device array of targets");
    return mass_impl(h, nmem, lnps, fix, 0.0, mass_targets, mass_out, lnr_out, "aer_mass_batch_d");
}

}  // extern "C"

// ------------------------------------------------------------------ zonal-mean statistics

struct aer_diag {
    aer_sht *h = nullptr;
    int nlev = 0;
    int nsample = 0;
    DevBuf sum;
};

namespace {

long diag_len(const aer_diag *d) { return ((long)aer::DG_NFIELD * d->nlev + 2) * d->h->nlat; }

int diag_zero(aer_diag *d)
{
    const long n = diag_len(d);
    hipLaunchKernelGGL(aer::k_diag_zero, dim3((unsigned)std::min<long>((n + 255) / 256, 1024)), dim3(256), 0,
                       d->h->stream, (double *)d->sum.p, n);
    HIP_TRY(hipGetLastError());
    d->nsample = 0;
    return AER_OK;
}

}  // namespace

extern "C" {

int aer_diag_create(aer_sht_t h, int nlev, aer_diag_t *diag)
{
    if (!diag) return fail(AER_EINVAL, "aer_diag_create: null output");
    *diag = nullptr;
    if (!h || !h->grid_set) return fail(AER_ESTATE, "aer_diag_create: handle without a grid");
    if (nlev < 1) return fail(AER_EINVAL, "aer_diag_create: nlev < 1");
    if (h->nlon > 64 * aer::DG_MAXPL) return fail(AER_EINVAL, "aer_diag_create: nlon above 512");
    aer_diag *d = new aer_diag();
    d->h = h;
    d->nlev = nlev;
    int rc = d->sum.ensure((size_t)diag_len(d) * sizeof(double));
    if (!rc) rc = diag_zero(d);
    if (rc) {
        delete d;
        return rc;
    }
    *diag = d;
    return AER_OK;
}

int aer_diag_accumulate_d(aer_diag_t d, const double *u, const double *v, const double *temp, const double *lnps)
{
    if (!d) return fail(AER_EINVAL, "aer_diag_accumulate_d: null object");
    if (!u || !v || !temp || !lnps) return fail(AER_EINVAL, "aer_diag_accumulate_d: null field");
    aer::DiagArgs A{};
    A.nlon = d->h->nlon;
    A.nlat = d->h->nlat;
    A.nlev = d->nlev;
    A.u = u;
    A.v = v;
    A.temp = temp;
    A.lnps = lnps;
    A.sum = (double *)d->sum.p;
    hipLaunchKernelGGL(aer::k_diag_accumulate, dim3(A.nlat, A.nlev + 1), dim3(64), 0, d->h->stream, A);
    HIP_TRY(hipGetLastError());
    d->nsample += 1;
    return AER_OK;
}

int aer_diag_read(aer_diag_t d, int *nsample, double *mean)
{
    if (!d || !nsample || !mean) return fail(AER_EINVAL, "aer_diag_read: null argument");
    const long n = diag_len(d);
    HIP_TRY(hipMemcpyAsync(mean, d->sum.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, d->h->stream));
    HIP_TRY(hipStreamSynchronize(d->h->stream));
    *nsample = d->nsample;
    if (d->nsample > 0) {
        const double rn = 1.0 / (double)d->nsample;
        for (long i = 0; i < n; ++i) mean[i] *= rn;
    }
    return AER_OK;
}

int aer_diag_reset(aer_diag_t d)
{
    if (!d) return fail(AER_EINVAL, "aer_diag_reset: null object");
    return diag_zero(d);
}

int aer_diag_destroy(aer_diag_t d)
{
    if (!d) return AER_OK;
    d->sum.release();
    delete d;
    return AER_OK;
}

}  // extern "C"
