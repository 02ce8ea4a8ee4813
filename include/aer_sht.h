/*
 * aer_sht.h -- C ABI of the MI355X (gfx950) spherical-harmonic transform
 * library that replaces SHTns on the aeros transform path.
 *
 * The reference reaches SHTns from Fortran through SHTns' own iso_c_binding
 * header (`include 'shtns.f03'`, reference src/aeros_spectral.f90:41-45). Each
 * entry point below names the SHTns call it replaces and the reference line
 * that makes that call. The Fortran module aeros_amd/fortran/aeros_spectral.f90
 * binds these with iso_c_binding and keeps the reference's module API.
 *
 * Conventions (reference src/aeros_spectral.f90:17-26): orthonormal harmonics,
 * no Condon-Shortley phase; grids are (nlon, nlat) Fortran order, longitude
 * contiguous, latitude north -> south, no padding; spectral arrays are nlm
 * complex coefficients, m-major with l fastest, stored as interleaved
 * (re, im) doubles (Fortran complex(c_double_complex)); triangular
 * truncation lmax = mmax, mres = 1.
 *
 * Every function returns an int status: 0 on success, an AER_E* code
 * otherwise. aer_strerror() gives a message; aer_last_error() gives the detail
 * of the most recent failure on the calling thread. The Fortran shim turns a
 * non-zero status into a message on error_unit plus `error stop`, as the
 * reference does for its own fatal conditions (src/aeros_spectral.f90:236-244).
 *
 * Threading: a handle is not safe for concurrent calls (the reference's SHTns
 * configs are not either, src/aeros_spectral.f90:50-54); create one handle
 * per host thread. Results are deterministic: no atomics, fixed reduction
 * order, so repeated calls are bit-identical.
 */
#ifndef AER_SHT_H
#define AER_SHT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define AER_OK      0
#define AER_EINVAL  1  /* invalid argument, truncation, grid or flag combination */
#define AER_ENOMEM  2  /* device or host allocation failed */
#define AER_EHIP    3  /* HIP runtime error (detail in aer_last_error) */
#define AER_ENODEV  4  /* no gfx950 device visible: there is no CPU fallback */
#define AER_ESTATE  5  /* call out of order, e.g. a transform before set_grid */

/* ---- normalization / layout flags: the subset aeros requests
 * (SHT_ORTHONORMAL + SHT_NO_CS_PHASE, SHT_PHI_CONTIGUOUS, SHT_GAUSS or
 * SHT_QUICK_INIT, optional SHT_LOAD_SAVE_CFG; src/aeros_spectral.f90:200-221). */
#define AER_SHT_ORTHONORMAL    0x1
#define AER_SHT_NO_CS_PHASE    0x2
#define AER_SHT_PHI_CONTIGUOUS 0x10
#define AER_SHT_GAUSS          0x20
#define AER_SHT_QUICK_INIT     0x40
#define AER_SHT_LOAD_SAVE_CFG  0x80

/* ---- arithmetic of the Legendre contraction and the longitude FFT */
#define AER_PREC_FP64  0  /* FP64 MFMA Legendre GEMM, FP64 FFT (the plain form) */
#define AER_PREC_FP32  1  /* FP32 MFMA Legendre GEMM, FP32 FFT */
#define AER_PREC_SPLIT 2  /* split-FP32 MFMA Legendre GEMM (near-FP64), FP64 FFT: operands as
                             five 10-bit integer slices under power-of-two scales, each
                             MFMA step summed exactly in FP32 and promoted to FP64 */

typedef struct aer_sht *aer_sht_t;

typedef struct aer_sht_info {
    int lmax, mmax, mres;
    int nlm;      /* (lmax+1)(lmax+2)/2 */
    int nlat, nphi, nspat, nlat_2;  /* nspat = nphi*nlat (never padded); nlat_2 = (nlat+1)/2 */
} aer_sht_info;

const char *aer_strerror(int status);
const char *aer_last_error(void);

/* Smallest quadratically unaliased Gaussian grid for truncation T: nlon >= 3T+1,
 * 2/3/5-smooth and even, nlat = nlon/2. Replaces aeros_sht_grid_size
 * (src/aeros_spectral.f90:399-438). */
int aer_sht_grid_size(int trunc, int *nlon, int *nlat);

/* Replaces shtns_create (src/aeros_spectral.f90:225). Host-only; norm must be
 * AER_SHT_ORTHONORMAL | AER_SHT_NO_CS_PHASE, mres must be 1, mmax == lmax. */
int aer_sht_create(int lmax, int mmax, int mres, int norm, aer_sht_t *cfg);

/* Replaces shtns_set_grid (src/aeros_spectral.f90:227; NB SHTns' argument order
 * nlat, nphi). Builds the Gauss grid and the FP64 Legendre tables and uploads
 * them to the current HIP device. layout must contain AER_SHT_PHI_CONTIGUOUS;
 * SHT_GAUSS / QUICK_INIT / LOAD_SAVE_CFG are accepted and change nothing (there
 * is no auto-tuning step). eps is recorded only: the contraction is always the
 * full dense sum (SHTns' polar skip below eps is not applied). */
int aer_sht_set_grid(aer_sht_t cfg, int layout, double eps, int nlat, int nphi);

/* Read-back of the realized configuration (the shtns_info fields read at
 * src/aeros_spectral.f90:230-260). */
int aer_sht_get_info(aer_sht_t cfg, aer_sht_info *info);
/* li[lm], mi[lm] for lm = 0..nlm-1 (SHTns info%li / info%mi, src/aeros_spectral.f90:246-252). */
int aer_sht_get_lm(aer_sht_t cfg, int *li, int *mi);
/* cos(colatitude) of the Gauss nodes, north -> south (info%ct, src/aeros_spectral.f90:255-258). */
int aer_sht_get_cos_theta(aer_sht_t cfg, double *ct);
/* Gauss weights of the northern half, nlat_2 values (shtns_gauss_wts, src/aeros_spectral.f90:271). */
int aer_sht_gauss_wts(aer_sht_t cfg, double *wts_half);
/* Host-only geometry and FP64 Legendre tables, exactly as set_grid uploads them:
 * ct[nlat] (cos colatitude, N->S), w[nlat] (Gauss weights), and P, dP/dtheta,
 * m P/sin(theta) at the northern nodes, each [nlm][(nlat+1)/2]. Needs no GPU; no
 * reference counterpart (SHTns keeps its tables internal, src/aeros_spectral.f90:246-275
 * reads back only ct and the weights). */
int aer_sht_host_tables(int lmax, int nlat, double *ct, double *w, double *P, double *D, double *Q);
/* Replaces shtns_destroy (src/aeros_spectral.f90:288). Accepts NULL. */
int aer_sht_destroy(aer_sht_t cfg);

/* Precision of subsequent transforms on this handle (default AER_PREC_FP64). */
int aer_sht_set_precision(aer_sht_t cfg, int prec);
/* HIP stream (hipStream_t) for subsequent device-pointer calls; NULL = default stream. */
int aer_sht_set_stream(aer_sht_t cfg, void *stream);
/* Kernel path of subsequent transforms on this handle, both directions (no reference
 * counterpart; SHTns picks its own kernels). AER_PATH_STAGED (default): Legendre and FFT
 * as separate launches with an HBM intermediate, every precision and grid length.
 * AER_PATH_TILED: one fused Legendre+FFT synthesis launch (4 grid fields x 4 latitude
 * pairs per block, 4x4 MFMA) and a two-launch analysis, for FP64 and FP32 jobs on the
 * compile-time grid lengths 64, 96, 128, 160, 192, 256 and 320; other jobs fall back to
 * the staged path. Both give the same results to rounding. */
#define AER_PATH_TILED  0
#define AER_PATH_STAGED 1
int aer_sht_set_path(aer_sht_t cfg, int path);

/* ---- per-call transforms on HOST arrays, one field; synchronous.
 * Drop-in replacements for the five SHTns calls aeros makes. Analyses leave
 * their grid inputs unchanged (the reference allows them to be destroyed,
 * src/aeros_spectral.f90:125-126). */
int aer_spat_to_SH(aer_sht_t cfg, double *vr, double *qlm);                    /* spat_to_SH,       src/aeros_spectral.f90:496 */
int aer_SH_to_spat(aer_sht_t cfg, const double *qlm, double *vr);              /* SH_to_spat,       src/aeros_spectral.f90:511 */
int aer_spat_to_SHsphtor(aer_sht_t cfg, double *vt, double *vp,
                         double *slm, double *tlm);                             /* spat_to_SHsphtor, src/aeros_spectral.f90:539 */
int aer_SHsphtor_to_spat(aer_sht_t cfg, const double *slm, const double *tlm,
                         double *vt, double *vp);                               /* SHsphtor_to_spat, src/aeros_spectral.f90:554 */
int aer_SHsph_to_spat(aer_sht_t cfg, const double *slm, double *vt, double *vp);/* SHsph_to_spat,    src/aeros_spectral.f90:580 */

/* ---- batched transforms on DEVICE arrays (asynchronous on the handle's stream).
 * nfield fields; field k of a spectral array starts at q + k*qstride doubles
 * (qstride >= 2*nlm), of a grid array at g + k*gstride doubles
 * (gstride >= nlon*nlat). These are the level loops of the reference
 * (src/dynamics/aeros_tendency.f90:409-419, 449-465) as one launch. */
int aer_synth_d(aer_sht_t cfg, int nfield, const double *qlm, long qstride, double *vr, long gstride);
int aer_anal_d(aer_sht_t cfg, int nfield, const double *vr, long gstride, double *qlm, long qstride);
int aer_synth_vec_d(aer_sht_t cfg, int nfield, const double *slm, const double *tlm, long qstride,
                    double *vt, double *vp, long gstride);
int aer_anal_vec_d(aer_sht_t cfg, int nfield, const double *vt, const double *vp, long gstride,
                   double *slm, double *tlm, long qstride);
int aer_synth_grad_d(aer_sht_t cfg, int nfield, const double *slm, long qstride,
                     double *gt, double *gp, long gstride);

/* ---- aeros_vordiv on DEVICE arrays, nlev levels, dense (nlm | nlon*nlat per level).
 * uv_from_vordiv:  src/dynamics/aeros_vordiv.f90:82-122
 * vordiv_from_uv:  src/dynamics/aeros_vordiv.f90:124-168 (inputs preserved)
 * gradient:        src/dynamics/aeros_vordiv.f90:170-205 (east, north components) */
int aer_uv_from_vordiv_d(aer_sht_t cfg, int nlev, const double *vor, const double *div,
                         double *u, double *v);
int aer_vordiv_from_uv_d(aer_sht_t cfg, int nlev, const double *u, const double *v,
                         double *vor, double *div);
int aer_gradient_d(aer_sht_t cfg, int nlev, const double *f, double *dfdx, double *dfdy);

/* ---- the transform halves of the tendency, DEVICE arrays, all levels fused.
 * 3-D spectral arrays are (nlm, nlev), 3-D grid arrays (nlon, nlat, nlev), Fortran order.
 * grid_transforms: the spectral->grid half of aeros_tendency_grid
 *                  (src/dynamics/aeros_tendency.f90:403-419)
 * spectral:        aeros_tendency_spectral, the spectral tendency assembly
 *                  (src/dynamics/aeros_tendency.f90:428-473): (zeta_t, D_t) from the
 *                  vector analysis of (ae, an), D_t -= laplacian(analysis(ek)),
 *                  T_t = analysis(dtdt), lnps_t = analysis(dlnpsdt). Inputs preserved. */
int aer_tendency_grid_transforms_d(aer_sht_t cfg, int nlev,
        const double *vor, const double *div, const double *temp, const double *lnps,
        double *u, double *v, double *vor_g, double *div_g, double *t_g,
        double *dtdx, double *dtdy, double *lnps_g, double *dlnpsdx, double *dlnpsdy);
int aer_tendency_spectral_d(aer_sht_t cfg, int nlev,
        const double *ae, const double *an, const double *ek, const double *dtdt,
        const double *dlnpsdt,
        double *tvor, double *tdiv, double *ttemp, double *tlnps);
/* Ensemble forms of the two halves: nmem members of this size, member b's 3-D arrays
 * at b*nlev fields and its single-level arrays at b fields from the first pointer
 * ([nmem][nlev][...] and [nmem][...] contiguous). The reference steps one model per
 * call (src/dynamics/aeros_tendency.f90:362-473); members batch into the same fused
 * launches as levels do. */
int aer_tendency_grid_transforms_batch_d(aer_sht_t cfg, int nmem, int nlev,
        const double *vor, const double *div, const double *temp, const double *lnps,
        double *u, double *v, double *vor_g, double *div_g, double *t_g,
        double *dtdx, double *dtdy, double *lnps_g, double *dlnpsdx, double *dlnpsdy);
int aer_tendency_spectral_batch_d(aer_sht_t cfg, int nmem, int nlev,
        const double *ae, const double *an, const double *ek, const double *dtdt,
        const double *dlnpsdt,
        double *tvor, double *tdiv, double *ttemp, double *tlnps);

/* ---- fused launches: a list of batched transforms of one direction run as ONE
 * launch per stage (the Legendre GEMM kernel and the FFT kernel). Every entry
 * point above is such a list; this is the general form, used to put all the
 * transforms of a model step (for example the tendency halves together with
 * passive-scalar transforms) into one launch pair, and to record that launch
 * pair into a HIP graph (no allocation happens inside after the first call of a
 * given shape). A later, larger call on the same handle grows the handle's Fourier
 * intermediate without freeing the previous one (kept until aer_sht_destroy), so a
 * graph captured earlier stays valid. No reference counterpart: the reference loops
 * over levels and fields with one SHTns call each (src/dynamics/aeros_tendency.f90:409-419,
 * 449-465). */
#define AER_JOB_SYNTH_SCALAR 0   /* in0 spectral -> out0 grid */
#define AER_JOB_SYNTH_VECTOR 1   /* (in0 = S, in1 = T) -> (out0 = V_theta, out1 = V_phi) */
#define AER_JOB_SYNTH_GRAD   2   /* in0 -> (out0 = d/dtheta, out1 = (1/sin) d/dphi) */
#define AER_JOB_ANAL_SCALAR  3   /* in0 grid -> out0 spectral */
#define AER_JOB_ANAL_VECTOR  4   /* (in0 = V_theta, in1 = V_phi[, in2 = X]) -> (out0 = S, out1 = T) */

/* Per-degree factors on spectral inputs (synthesis) or outputs (analysis). */
#define AER_MODE_ONE        0   /* 1 */
#define AER_MODE_S_FROM_DIV 1   /* -a/(l(l+1)), 0 at l = 0 */
#define AER_MODE_T_FROM_VOR 2   /* +a/(l(l+1)), 0 at l = 0 */
#define AER_MODE_INV_LL1    3   /* 1/(l(l+1)), 0 at l = 0 */
#define AER_MODE_DIV_FROM_S 4   /* -1/a after the 1/(l(l+1)) of the vector analysis, 0 at l = 0 */
#define AER_MODE_VOR_FROM_T 5   /* +1/a likewise */
#define AER_MODE_NEG_LAP    6   /* +l(l+1)/a^2 (minus the Laplacian) */
#define AER_MODE_ZERO       7   /* 0 */

typedef struct aer_job {
    int kind;              /* AER_JOB_* */
    int nfield;            /* scalar fields, or vector pairs */
    const double *in0, *in1, *in2;
    double *out0, *out1;
    long in_stride, out_stride;   /* doubles between consecutive fields; 0 = dense */
    int mode0, mode1, mode2;      /* AER_MODE_*: spectral factor of in0/in1 (synthesis),
                                     of out0/out1 (analysis); mode2 applies to the scalar
                                     analysis of in2, which is added into out0 */
    double scale0, scale1;        /* grid-side factor of component 0 / 1 (0 means 1) */
    int accumulate;               /* analysis: add into out0/out1 instead of overwriting */
    int prec;                     /* AER_JOB_PREC_*: compute type of this job */
} aer_job;

/* Per-job compute type. Jobs of different types may share one launch (the kernels
 * branch per job), e.g. an FP32 transform set whose cancellation-prone analysis
 * (D_t = div(A) - lap(Phi+K)) runs in FP64. */
#define AER_JOB_PREC_HANDLE 0     /* the handle's aer_sht_set_precision setting */
#define AER_JOB_PREC_FP64   1
#define AER_JOB_PREC_FP32   2
#define AER_JOB_PREC_SPLIT  3

/* All jobs must be of one direction (all AER_JOB_SYNTH_* or all AER_JOB_ANAL_*);
 * at most AER_MAX_JOBS jobs per call. Asynchronous on the handle's stream. */
#define AER_MAX_JOBS 8
int aer_run_jobs_d(aer_sht_t cfg, int njob, const aer_job *jobs);

/* ---- compiled plans: a fixed list of jobs, each on its own handle (so members of
 * an ensemble may differ in truncation and level count), built once into
 * device-resident job tables. Running a plan is at most four launches (synthesis
 * Legendre + FFT, then analysis FFT + Legendre) on the given stream, with no host
 * work beyond the launches and no allocation, so it can be captured in a HIP graph.
 * All handles must have the same precision (taken at creation); arrays named in
 * the jobs must stay valid while the plan is used. No reference counterpart: the
 * reference has no ensemble or batching layer (its parallel unit is one transform
 * per OpenMP thread, src/aeros_spectral.f90:87-113). */
typedef struct aer_plan *aer_plan_t;
int aer_plan_create(int njob, const aer_sht_t *cfgs, const aer_job *jobs, aer_plan_t *plan);
int aer_plan_run(aer_plan_t plan, void *stream);
int aer_plan_destroy(aer_plan_t plan);

/* ---- adjoint (transpose) transforms, for tangent-linear / adjoint models (no
 * reference counterpart). Each is the exact transpose of the named transform in
 * the real inner products sum(re*re + im*im) over the nlm coefficients and
 * sum over the nlon*nlat grid points (unweighted):
 *   aer_synth_adj_d      spectral = synthesis^T (grid)
 *   aer_anal_adj_d       grid     = analysis^T (spectral)
 *   aer_synth_vec_adj_d  (S, T)   = vector synthesis^T (V_theta, V_phi)
 *   aer_anal_vec_adj_d   (V_theta, V_phi) = vector analysis^T (S, T)
 *   aer_synth_grad_adj_d S        = gradient synthesis^T (g_theta, g_phi)
 * They run the forward kernels with unit row weights and the m > 0 factor 2 in the
 * analysis direction, and with the m > 0 coefficients halved and the rows weighted
 * by w_j 2 pi/nlon in the synthesis direction. Strides as for the batched calls. */
int aer_synth_adj_d(aer_sht_t cfg, int nfield, const double *vr, long gstride, double *qlm, long qstride);
int aer_anal_adj_d(aer_sht_t cfg, int nfield, const double *qlm, long qstride, double *vr, long gstride);
int aer_synth_vec_adj_d(aer_sht_t cfg, int nfield, const double *vt, const double *vp, long gstride,
                        double *slm, double *tlm, long qstride);
int aer_anal_vec_adj_d(aer_sht_t cfg, int nfield, const double *slm, const double *tlm, long qstride,
                       double *vt, double *vp, long gstride);
int aer_synth_grad_adj_d(aer_sht_t cfg, int nfield, const double *gt, const double *gp, long gstride,
                         double *slm, long qstride);

/* ---- grid-space column work between the transform halves of a step.
 *
 * Vertical coordinate: the hybrid coefficients of the reference's
 * aeros_vgrid_class (src/dynamics/aeros_vertical.f90:46-55), A (Pa) and B (-) at
 * the nlev+1 half levels, top first, copied to the device once. 2 <= nlev <= 64. */
typedef struct aer_vgrid *aer_vgrid_t;
int aer_vgrid_create(int nlev, const double *a_half, const double *b_half, aer_vgrid_t *vg);
int aer_vgrid_destroy(aer_vgrid_t vg);

/* Grid right-hand sides of every column (column_terms + vert_adv,
 * src/dynamics/aeros_tendency.f90:475-630): from the synthesized fields of
 * aer_tendency_grid_transforms_d (3-D arrays [nlev][nlat][nlon], 2-D [nlat][nlon],
 * device pointers) to ae, an, ek = Phi + K, dtdt and dlnpsdt. phis (surface
 * geopotential) may be NULL for a flat surface. held_suarez != 0 adds the
 * Held-Suarez forcing in the same pass (aeros_hs_apply,
 * src/physics/aeros_held_suarez.f90:177-223). Asynchronous on the handle's stream. */
int aer_column_terms_d(aer_sht_t cfg, aer_vgrid_t vg, int held_suarez,
                       const double *u, const double *v, const double *vor_g, const double *div_g,
                       const double *t_g, const double *dtdx, const double *dtdy,
                       const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                       const double *phis,
                       double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt);
/* The same with the Held-Suarez forcing on and an additive equilibrium-temperature
 * field dteq [nlev][nlat] (T_eq = formula + dteq): the control variable of the adjoint.
 * No reference counterpart (the reference's T_eq is the formula alone). */
int aer_column_terms_dteq_d(aer_sht_t cfg, aer_vgrid_t vg, const double *dteq,
                            const double *u, const double *v, const double *vor_g, const double *div_g,
                            const double *t_g, const double *dtdx, const double *dtdy,
                            const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                            const double *phis,
                            double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt);
/* Adjoint (transpose of the linearization at the given inputs) of aer_column_terms_d /
 * aer_column_terms_dteq_d: from the adjoints of (ae, an, ek, dtdt, dlnpsdt) to those of the
 * ten grid inputs (overwritten) and, if g_teq is not null, the per-column sensitivity to
 * dteq [nlev][nlat*nlon] (sum it over longitudes for the field's gradient). nlev <= 64.
 * No reference counterpart. */
int aer_column_terms_adj_d(aer_sht_t cfg, aer_vgrid_t vg, int held_suarez, const double *dteq,
                           const double *u, const double *v, const double *vor_g, const double *div_g,
                           const double *t_g, const double *dtdx, const double *dtdy,
                           const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                           const double *phis,
                           const double *b_ae, const double *b_an, const double *b_ek, const double *b_dtdt,
                           const double *b_dlnpsdt,
                           double *g_u, double *g_v, double *g_vor, double *g_div, double *g_t,
                           double *g_dtdx, double *g_dtdy, double *g_lnps, double *g_dlnpsdx, double *g_dlnpsdy,
                           double *g_teq);
/* Ensemble forms of the column terms and their adjoint: nmem members of this size,
 * grid arrays [nmem][nlev][nlat][nlon] and [nmem][nlat][nlon], one launch (the member
 * is the grid's second dimension, at most 65535). dteq (may be NULL) is member b's at
 * b*dteq_mstride doubles (0: one field shared by all members); phis is shared. With
 * nmem = 1 and dteq_mstride = 0 they are aer_column_terms_d / _dteq_d / _adj_d.
 * No reference counterpart (the reference steps one model per call). */
int aer_column_terms_batch_d(aer_sht_t cfg, aer_vgrid_t vg, int nmem, int held_suarez,
                             const double *dteq, long dteq_mstride,
                             const double *u, const double *v, const double *vor_g, const double *div_g,
                             const double *t_g, const double *dtdx, const double *dtdy,
                             const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                             const double *phis,
                             double *ae, double *an, double *ek, double *dtdt, double *dlnpsdt);
int aer_column_terms_adj_batch_d(aer_sht_t cfg, aer_vgrid_t vg, int nmem, int held_suarez,
                                 const double *dteq, long dteq_mstride,
                                 const double *u, const double *v, const double *vor_g, const double *div_g,
                                 const double *t_g, const double *dtdx, const double *dtdy,
                                 const double *lnps_g, const double *dlnpsdx, const double *dlnpsdy,
                                 const double *phis,
                                 const double *b_ae, const double *b_an, const double *b_ek, const double *b_dtdt,
                                 const double *b_dlnpsdt,
                                 double *g_u, double *g_v, double *g_vor, double *g_div, double *g_t,
                                 double *g_dtdx, double *g_dtdy, double *g_lnps, double *g_dlnpsdx,
                                 double *g_dlnpsdy, double *g_teq);

/* ---- spectral step completion: the semi-implicit gravity-wave solve
 * (src/dynamics/aeros_semiimp.f90) and the leapfrog pieces of
 * aeros_timestep_step (src/dynamics/aeros_timestep.f90:1043-1100, 1225-1276,
 * 1673-1815). Spectral 3-D fields are complex [nlev][nlm] device arrays, 2-D
 * (ln p_s) [nlm]; all arithmetic FP64. New-state arrays must not alias the
 * old, now or tendency arrays. */
typedef struct aer_semiimp *aer_semiimp_t;

/* aeros_semiimp_init (src/dynamics/aeros_semiimp.f90:171-276): the linear
 * operators about an isothermal t_ref at ps_ref on the handle's truncation, and
 * the degree-by-degree Helmholtz factors for a step dt_step (h: 2 dt for a
 * leapfrog step, dt for the start-up step). alpha in [0.5, 1] is the decentering
 * weight on X^(n+1). nlev <= 64. */
int aer_semiimp_create(aer_sht_t cfg, aer_vgrid_t vg, double t_ref, double ps_ref, double dt_step,
                       double alpha, aer_semiimp_t *si);
/* aeros_semiimp_set_step (:301-343): refactor for a new step length. The device
 * tables are updated in place while the leapfrog takes the step length as a launch
 * argument: a HIP graph that recorded aer_leapfrog_d replays the new tables with the
 * step it recorded, so capture after the last set_step or set_diffusion. */
int aer_semiimp_set_step(aer_semiimp_t si, double dt_step);
/* Implicit del^ndiff horizontal diffusion (src/dynamics/aeros_timestep.f90:1673-1726)
 * with e-folding time tau_diff [s] at l = lmax for vorticity and temperature and
 * tau_diff_div for divergence (<= 0: tau_diff); rebuilt with every set_step. */
int aer_semiimp_set_diffusion(aer_semiimp_t si, double tau_diff, double tau_diff_div, int ndiff);
/* aeros_semiimp_gwspeed (:573-620): fastest linear gravity-wave speed [m/s]. Host only. */
int aer_semiimp_gwspeed(aer_semiimp_t si, double *cmax);
int aer_semiimp_destroy(aer_semiimp_t si);
/* aeros_semiimp_linear (:345-399): L applied to (div, temp, lnps). */
int aer_semiimp_linear_d(aer_semiimp_t si, const double *div, const double *temp, const double *lnps,
                         double *ldiv, double *ltemp, double *lps);
/* aeros_semiimp_step (:462-571): D, T, ln p_s at n+1 from n-1, n and R(X^n). */
int aer_semiimp_step_d(aer_semiimp_t si,
                       const double *old_div, const double *old_temp, const double *old_lnps,
                       const double *now_div, const double *now_temp, const double *now_lnps,
                       const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                       double *new_div, double *new_temp, double *new_lnps);

/* Sections 2-4 of aeros_timestep_step in one launch: the vorticity leapfrog, the
 * semi-implicit (AER_STEP_SEMI_IMPLICIT) or explicit update of D, T, ln p_s, the
 * implicit diffusion (AER_STEP_DIFFUSE; after aer_semiimp_set_diffusion) and the
 * Robert-Asselin-Williams filter (AER_STEP_FILTER, eps_filter > 0, raw_alpha = 1:
 * classical Robert-Asselin), which updates the now arrays in place. The step
 * length is the semi-implicit object's. */
#define AER_STEP_SEMI_IMPLICIT 1
#define AER_STEP_DIFFUSE       2
#define AER_STEP_FILTER        4
int aer_leapfrog_d(aer_semiimp_t si, int flags, double eps_filter, double raw_alpha,
                   const double *old_vor, const double *old_div, const double *old_temp, const double *old_lnps,
                   double *now_vor, double *now_div, double *now_temp, double *now_lnps,
                   const double *tnd_vor, const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                   double *new_vor, double *new_div, double *new_temp, double *new_lnps);

/* Adjoint of aer_leapfrog_d with the same flags, filter and step settings: the transpose
 * of the linear map (X^(n-1), X^n, R) -> (X^(n+1), filtered X^n). Inputs are the adjoints
 * of X^(n+1) (bn_*) and of the filtered X^n (bc_*); outputs (overwritten) are the
 * adjoints of X^(n-1) (go_*), of X^n as it entered the step (gc_*) and of the tendency
 * (gt_*). No reference counterpart. */
int aer_leapfrog_adj_d(aer_semiimp_t si, int flags, double eps_filter, double raw_alpha,
                       const double *bn_vor, const double *bn_div, const double *bn_temp, const double *bn_lnps,
                       const double *bc_vor, const double *bc_div, const double *bc_temp, const double *bc_lnps,
                       double *go_vor, double *go_div, double *go_temp, double *go_lnps,
                       double *gc_vor, double *gc_div, double *gc_temp, double *gc_lnps,
                       double *gt_vor, double *gt_div, double *gt_temp, double *gt_lnps);

/* Global dry-air mass of a spectral ln p_s (mass_of, src/dynamics/aeros_timestep.f90:1250-1276):
 * one FP64 synthesis and an FP64 quadrature, written to the device double
 * *mass_out (may be NULL). With fix != 0 also the mass fixer (mass_fix, :1225-1248):
 * ln p_s(0,0) += ln(mass_target/mass) c00, ln r written to *lnr_out (device, may be
 * NULL). c00 is read back from an analysis of the constant 1 at aer_sht_set_grid. */
int aer_mass_d(aer_sht_t cfg, double *lnps, int fix, double mass_target, double *mass_out, double *lnr_out);

/* Ensemble forms of the leapfrog, its adjoint and the mass fixer: nmem members of this
 * size, spectral arrays [nmem][nlev][nlm] and [nmem][nlm] (complex), one launch each.
 * aer_mass_batch_d reads member b's target from the DEVICE array mass_targets[b] and
 * writes mass_out[b], lnr_out[b] (device, may be NULL); its first call at a given
 * member count sizes the handle's grid buffer, so make that call outside graph capture.
 * No reference counterpart (the reference steps one model per call). */
int aer_leapfrog_batch_d(aer_semiimp_t si, int nmem, int flags, double eps_filter, double raw_alpha,
                         const double *old_vor, const double *old_div, const double *old_temp, const double *old_lnps,
                         double *now_vor, double *now_div, double *now_temp, double *now_lnps,
                         const double *tnd_vor, const double *tnd_div, const double *tnd_temp, const double *tnd_lnps,
                         double *new_vor, double *new_div, double *new_temp, double *new_lnps);
int aer_leapfrog_adj_batch_d(aer_semiimp_t si, int nmem, int flags, double eps_filter, double raw_alpha,
                             const double *bn_vor, const double *bn_div, const double *bn_temp, const double *bn_lnps,
                             const double *bc_vor, const double *bc_div, const double *bc_temp, const double *bc_lnps,
                             double *go_vor, double *go_div, double *go_temp, double *go_lnps,
                             double *gc_vor, double *gc_div, double *gc_temp, double *gc_lnps,
                             double *gt_vor, double *gt_div, double *gt_temp, double *gt_lnps);
int aer_mass_batch_d(aer_sht_t cfg, int nmem, double *lnps, int fix, const double *mass_targets,
                     double *mass_out, double *lnr_out);

/* ---- zonal-mean statistics: replaces aeros_diag_init / aeros_diag_accumulate /
 * aeros_diag_finalize's sums (src/aeros_diagnostics.f90:114-276) with FP64 running sums on
 * the device. One sample (asynchronous on the handle's stream) takes the grid fields
 * u, v, T [nlev][nlat][nlon] and ln p_s [nlat][nlon] and adds, per (level, latitude), the
 * instantaneous zonal means [u], [v], [T] and the covariances about them [u'v'], [v'T'],
 * [(u'^2+v'^2)/2], [T'^2], plus [p_s] (p_s = exp(ln p_s)) and the lowest level's [u].
 * aer_diag_read synchronizes and returns the time means (sums / samples) as
 * [7][nlev][nlat] (u, v, T, u'v', v'T', eke, T'^2; level 0 at the top, latitude 0 at the
 * north) followed by [2][nlat] (p_s, lowest-level u), and the sample count. */
typedef struct aer_diag *aer_diag_t;
int aer_diag_create(aer_sht_t cfg, int nlev, aer_diag_t *diag);
int aer_diag_accumulate_d(aer_diag_t diag, const double *u, const double *v, const double *temp,
                          const double *lnps_grid);
int aer_diag_read(aer_diag_t diag, int *nsample, double *mean);
int aer_diag_reset(aer_diag_t diag);
int aer_diag_destroy(aer_diag_t diag);

/* ---- device memory for hosts without a HIP binding of their own (Fortran): thin
 * wrappers of hipMalloc / hipFree / synchronous hipMemcpy. No reference counterpart. */
int aer_dev_malloc(size_t bytes, void **ptr);
int aer_dev_free(void *ptr);
int aer_memcpy_h2d(void *dst, const void *src, size_t bytes);
int aer_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* Block until all work queued on the handle's stream has finished. */
int aer_sht_sync(aer_sht_t cfg);

/* ---- autotuner (no reference counterpart). Times a job list (one direction, as for
 * aer_run_jobs_d) or a compiled plan under every kernel configuration and keeps the
 * fastest for that direction on the handles involved. Candidates: 0 the staged path;
 * 1, 2, 3 the tiled path with m-band staging budgets of 16, 32 and 64 KiB. Each candidate
 * runs once untimed and reps times timed; synchronous; the jobs' outputs are overwritten.
 * times_ms (optional) receives the mean milliseconds per candidate (a plan: synthesis
 * candidates, then analysis candidates), choice (optional) the kept index (a plan: two).
 * aer_plan_autotune rebuilds the plan into fresh device job tables: a HIP graph that
 * recorded aer_plan_run of this plan before keeps replaying the configuration it
 * recorded (its tables stay allocated until aer_plan_destroy); capture again to replay
 * the tuned one. */
#define AER_TUNE_NCAND 4
int aer_autotune_jobs(aer_sht_t cfg, int njob, const aer_job *jobs, int reps, double *times_ms, int *choice);
int aer_plan_autotune(aer_plan_t plan, void *stream, int reps, double *times_ms, int *choice);

/* ---- observability: per-kernel timing with HIP events recorded on the handle's
 * stream around every kernel launch (off by default; no reference counterpart --
 * the reference times whole transforms with system_clock, drivers/bench_sht.f90:214-227).
 * Kernel kinds: 0 Legendre synthesis, 1 FFT synthesis, 2 FFT analysis, 3 Legendre analysis,
 * 4 fused Legendre+FFT synthesis (the tiled path).
 * aer_sht_timing(cfg, 1) starts and clears; aer_sht_timing_read() synchronizes the stream
 * and returns the summed milliseconds and launch counts per kind since the start. */
#define AER_NKERNEL_KIND 5
int aer_sht_timing(aer_sht_t cfg, int enable);
int aer_sht_timing_read(aer_sht_t cfg, double *total_ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif
