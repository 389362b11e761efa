/*
 * slfmm.h -- C ABI of the MI355X (gfx950) streaming LFMM2 library.
 *
 * Drop-in boundary for schnelLFMM's Rust library entry points (reference
 * crate `schnellfmm` v0.3.2).  Every entry point returns an int status
 * (SLFMM_OK = 0), never throws across the ABI, and records a message that
 * slfmm_last_error() returns.  Matrices are row-major float64 unless noted.
 * Caller memory is never freed by the library.  Output pointers are host
 * memory unless SLFMM_OUT_DEVICE is passed in `flags`, in which case they
 * are device pointers on the context's device (inputs stay host memory).
 *
 * Threading: one context owns one HIP stream at default priority; calls on
 * one context must not overlap.  Several contexts may coexist.
 */
#ifndef SLFMM_H
#define SLFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slfmm_ctx slfmm_ctx;
typedef struct slfmm_geno slfmm_geno;

/* Status codes.  Validation codes mirror the reference's `bail!` sites. */
enum {
  SLFMM_OK = 0,
  SLFMM_E_ARG = 1,        /* bad argument / shape                              */
  SLFMM_E_HIP = 2,        /* HIP runtime error                                 */
  SLFMM_E_IO = 3,         /* file open / read (src/bed.rs:86-97)               */
  SLFMM_E_BED_MAGIC = 4,  /* bad magic bytes (src/bed.rs:105-110)              */
  SLFMM_E_BED_SIZE = 5,   /* size mismatch (src/bed.rs:113-123)                */
  SLFMM_E_PARSE = 6,      /* .bim/.fam field count or number (src/bed.rs:456-503) */
  SLFMM_E_LAMBDA = 7,     /* lambda <= 0 (src/precompute.rs:31-38)             */
  SLFMM_E_DF = 8,         /* n <= 1+d+K (src/testing.rs:272-278)               */
  SLFMM_E_SINGULAR = 9,   /* safe_inv failed even after epsilon (src/testing.rs:1014-1021) */
  SLFMM_E_NOMEM = 10,     /* device allocation failed                          */
  SLFMM_E_NUMERIC = 11    /* non-finite intermediate                           */
};

enum { SLFMM_NORM_CENTER_ONLY = 0, SLFMM_NORM_EIGENSTRAT = 1 };   /* src/bed.rs:19-26 */
enum { SLFMM_IMPUTE_MEAN = 0, SLFMM_IMPUTE_NMF = 1 };              /* src/lib.rs:39-44 */

/* Arithmetic of the genotype-stream GEMMs.  Small dense algebra (QR of the
 * sketch, small SVD, ridge / K x K solves, GIF calibration) is FP64 on the
 * GPU in every mode. */
enum {
  SLFMM_PREC_FP64 = 0,    /* FP64 throughout                                   */
  SLFMM_PREC_I8 = 1       /* int8 MFMA on exact genotypes, multi-term int8
                             operands, exact int32 accumulation, FP64 epilogue.
                             The int8 kernels carry at most 32 columns: a sketch
                             with K + oversampling > 32, or an association with
                             K + d > 32, runs that stage's FP64 kernels, and
                             each such stage is reported once per context on
                             stderr. */
};

/* SLFMM_OUT_DEVICE: outputs stay in device memory (the pointers in the
 * output structs are device pointers).  SLFMM_EAGER: queue every kernel of
 * a fit directly; without it, fits with device-resident panels and device
 * outputs replay a HIP graph captured by an earlier call with the same
 * shapes, configuration and pointers (no counterpart in the reference).
 * SLFMM_NMF_INT8 (slfmm_nmf_estimate): the forward pass (the W update's
 * numerator) and the H update of every iteration run on the int8 matrix
 * cores for ranks k <= 8 (see slfmm_nmf_estimate). */
enum { SLFMM_OUT_DEVICE = 1u, SLFMM_EAGER = 2u, SLFMM_NMF_INT8 = 4u };

/* Lfmm2Config (src/lib.rs:47-98) plus the GPU precision switch. */
typedef struct {
  uint64_t k;             /* latent factors K            (default 5)       */
  double lambda;          /* ridge penalty               (default 1e-5)    */
  uint64_t chunk_size;    /* SNPs per chunk; only seeds CV masks (10000)   */
  uint64_t oversampling;  /* RSVD oversampling           (default 10)      */
  uint64_t n_power_iter;  /* RSVD power iterations       (default 2)       */
  uint64_t seed;          /* RNG seed                    (default 42)      */
  int32_t norm;           /* SLFMM_NORM_*                (CENTER_ONLY)     */
  int32_t scale_cov;      /* divide covariates by sd     (0)               */
  int32_t impute;         /* SLFMM_IMPUTE_*              (MEAN)            */
  int32_t precision;      /* SLFMM_PREC_*                (FP64)            */
} slfmm_config;

/* NmfConfig (src/nmf.rs:15-50). */
typedef struct {
  uint64_t k;             /* components                  (default 5)       */
  uint64_t n_iter;        /* multiplicative iterations   (default 10)      */
  double eps;             /* clamp / update threshold    (default 1e-16)   */
  double cv_rate;         /* CV hold-out rate            (default 5e-4)    */
  uint64_t chunk_size;    /* CV mask seeding granularity (default 10000)   */
  uint64_t seed;          /* base seed                   (default 42)      */
} slfmm_nmf_config;

/* TestResults (src/testing.rs:76-93) outputs of the association pass. */
typedef struct {
  double* beta;           /* p x d ridge effect sizes (src/testing.rs:359-362)     */
  double* tstat;          /* p x d t statistics       (src/testing.rs:387-390)     */
  double* r2;             /* p x 3 r2_cov, r2_latent, r2_resid (:392-404)          */
  double* pvalue;         /* p x d calibrated p values (src/testing.rs:685-713)    */
  uint64_t* hist;         /* d x 100000 t^2 histogram, may be NULL (:114-149)      */
  double* gif_per_trait;  /* d, clamped >= 1 (src/testing.rs:159-177)              */
  double* gif;            /* 1, average of unclamped estimates                     */
} slfmm_assoc_out;

void slfmm_config_default(slfmm_config* cfg);
void slfmm_nmf_config_default(slfmm_nmf_config* cfg);
const char* slfmm_version(void);

int slfmm_create(int device, slfmm_ctx** out);
/* Destroys a context.  Panels (slfmm_geno) created on it may be freed before
 * or after: a context with live panels is released by its last
 * slfmm_geno_free.  No other call may use the context after this. */
void slfmm_destroy(slfmm_ctx* ctx);
const char* slfmm_last_error(const slfmm_ctx* ctx);
int slfmm_synchronize(slfmm_ctx* ctx);

/* ---- L1: PLINK .bed (src/bed.rs) --------------------------------------- */

/* BedFile::open validation (src/bed.rs:78-135): parses <stem>.fam/.bim,
 * checks magic and exact size; reports n_samples and n_snps. */
int slfmm_bed_open(slfmm_ctx* ctx, const char* bed_path, uint64_t* n_samples, uint64_t* n_snps);

/* BedFile::open + mmap upload; sample_keep (nullable) mirrors
 * BedFile::subset_samples (src/bed.rs:155-159): physical indices, new order. */
int slfmm_geno_load_bed(slfmm_ctx* ctx, const char* bed_path, const int64_t* sample_keep,
                        uint64_t n_keep, slfmm_geno** out);

/* Upload packed SNP-major bytes (the .bed payload after the 3 magic bytes,
 * ceil(n_phys/4) bytes per SNP). */
int slfmm_geno_upload(slfmm_ctx* ctx, const uint8_t* packed, uint64_t n_phys, uint64_t n_snps,
                      const int64_t* sample_keep, uint64_t n_keep, slfmm_geno** out);

/* Wrap packed bytes already resident on the device (row stride `ld` bytes,
 * ld % 16 == 0, >= ceil(n/4); bytes past ceil(n/4) and padding codes must be
 * zero, 128 readable bytes must follow the last row).  Not owned. */
int slfmm_geno_wrap_device(slfmm_ctx* ctx, void* d_packed, uint64_t n, uint64_t n_snps,
                           uint64_t ld, slfmm_geno** out);

/* Host-resident panels (out-of-core): the rows stay in host memory and are
 * streamed to the device in chunks of `chunk_snps` rows (0: about 1 GiB of
 * rows per chunk) through two pinned staging buffers, the copies overlapping
 * the association of the previous chunk on a second stream.  Counterpart of
 * the reference's mmap + IO-thread chunk pipeline (BedFile::open,
 * src/bed.rs:78-135; parallel_stream, src/parallel.rs:156-329) for panels
 * larger than device memory.  Every entry point that takes a panel accepts
 * one, except that the test panel of slfmm_fit_lfmm2_nmf (NMF imputation in
 * the association pass) must be device-resident.  Results are identical to
 * those of the same rows held on the device.
 *
 * open_bed_host maps <bed_path> read-only (same validation as
 * slfmm_bed_open); the panel owns the mapping.  wrap_host borrows `packed`
 * (n_snps rows of ceil(n/4) bytes, the .bed payload after the magic bytes),
 * which must stay valid and unchanged until slfmm_geno_free; page-locked
 * memory (hipHostMalloc) is read by DMA directly, other memory is staged. */
int slfmm_geno_open_bed_host(slfmm_ctx* ctx, const char* bed_path, uint64_t chunk_snps, slfmm_geno** out);
int slfmm_geno_wrap_host(slfmm_ctx* ctx, const uint8_t* packed, uint64_t n, uint64_t n_snps,
                         uint64_t chunk_snps, slfmm_geno** out);

/* Copies rows [j0, j0 + rows) of a panel to host memory `out` in .bed payload
 * layout (ceil(n/4) bytes per SNP, padding codes zero): the payload that
 * write_bed_file (src/bed.rs:398-429) writes after the magic bytes, e.g. to
 * save a panel generated on the device. */
int slfmm_geno_read_packed(slfmm_ctx* ctx, const slfmm_geno* g, uint64_t j0, uint64_t rows, uint8_t* out);

void slfmm_geno_free(slfmm_geno* g);
int slfmm_geno_info(const slfmm_geno* g, uint64_t* n_samples, uint64_t* n_snps, uint64_t* ld);

/* Per-SNP integer genotype counts: missing calls, sum of dosages, count of
 * hom-alt calls (exact; mean = sum_g / (n - n_miss)). */
int slfmm_decode_stats(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t n_idx,
                       int32_t* n_miss, int32_t* sum_g, int32_t* n_hom_alt);

/* decode_bed_chunk_into (src/bed.rs:211-235, :300-379): n x c row-major,
 * optional n x c fill for missing entries (NULL = mean imputation). */
int slfmm_decode(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t c,
                 int32_t norm, const double* fill, double* out);

/* decode_raw_bed_chunk_into (src/bed.rs:246-281): {0,1,2,NaN}, n x c. */
int slfmm_decode_raw(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t c,
                     double* out);

/* ---- L3 algorithms ------------------------------------------------------- */

/* center_covariates (src/lib.rs:102-117).  Host arithmetic (n x d). */
int slfmm_center_covariates(const double* x, uint64_t n, uint64_t d, int32_t scale, double* out);

/* safe_inv (src/testing.rs:996-1024) on the GPU in FP64; *regularized = 1
 * when the epsilon*I fallback was taken. */
int slfmm_safe_inv(slfmm_ctx* ctx, const double* a, uint64_t m, double* out, int32_t* regularized);

/* precompute (src/precompute.rs:30-88) in thin form: sigma (d), d_lambda
 * (first d entries; the rest are 1), Q_d (n x d, left singular vectors) and
 * ridge_inv (d x d).  No n x n matrix is formed. */
int slfmm_precompute(slfmm_ctx* ctx, const double* xs, uint64_t n, uint64_t d, double lambda,
                     double* sigma, double* d_lambda, double* q_thin, double* ridge_inv);

/* estimate_factors_streaming (src/rsvd.rs:26-177).  xs is the centred
 * covariate matrix.  b0 = M^T Omega (n x l, l = k + oversampling) may be
 * injected; NULL draws Omega on the device from cfg->seed.  u_hat: n x k. */
int slfmm_estimate_factors(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* subset_idx,
                           uint64_t n_idx, const double* xs, uint64_t d, const slfmm_config* cfg,
                           const double* b0, double* u_hat, uint32_t flags);

/* test_associations_fused (src/testing.rs:252-528) with mean imputation,
 * GIF and calibration; outputs per slfmm_assoc_out. */
int slfmm_test_associations(slfmm_ctx* ctx, const slfmm_geno* g, const double* xs, uint64_t d,
                            const double* u_hat, const slfmm_config* cfg, slfmm_assoc_out* out,
                            uint32_t flags);

/* Allocation-free variant of slfmm_test_associations (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and outputs; every scratch buffer comes from a static device
 * array of 96 MiB per GPU (held until the process ends), so the call makes no
 * device allocation.  Requires SLFMM_OUT_DEVICE, a device-resident panel and
 * SLFMM_PREC_I8 with K + d <= 32; shapes whose scratch would exceed the array
 * (about n (4 (K + d) + ...) doubles plus the histogram) are rejected with
 * SLFMM_E_ARG before any work is queued.  Calls that overlap on one GPU share
 * the array and race. */
int slfmm_test_associations_noalloc(slfmm_ctx* ctx, const slfmm_geno* g, const double* xs, uint64_t d,
                                    const double* u_hat, const slfmm_config* cfg, slfmm_assoc_out* out,
                                    uint32_t flags);

/* fit_lfmm2 (src/lib.rs:162-263), mean imputation: centres x, estimates
 * U_hat on est (subset_idx of est, NULL = all), tests every SNP of g. */
int slfmm_fit_lfmm2(slfmm_ctx* ctx, const slfmm_geno* est, const int64_t* subset_idx,
                    uint64_t n_idx, const slfmm_geno* g, const double* x, uint64_t d,
                    const slfmm_config* cfg, const double* b0, double* u_hat,
                    slfmm_assoc_out* out, uint32_t flags);

/* ---- L3: NMF imputation (src/nmf.rs) -------------------------------------- */

/* nmf_impute_estimation (src/nmf.rs:80-249) in FP64 on the estimation SNPs
 * subset_idx of g (NULL = all SNPs): w (n x k), h (k x p_est; the
 * reference's h_full), cv_mae (n_iter: held-out MAE of W H per iteration),
 * cv_mae_mean (1: MAE of column-mean imputation in the last iteration),
 * mask_count (1: held-out entries summed over the iterations).  omega
 * (n x (k + 10)) may be injected; NULL draws it from cfg->seed.  The CV
 * hold-out draw of entry (sample i, column c of chunk q) in iteration t is
 * element c * n + i of the library's uniform stream seeded cfg->seed + t + q
 * (the reference draws from ChaCha8).  w and h follow `flags`; the CV
 * outputs are host memory.  With SLFMM_NMF_INT8 and k <= 8 both passes of
 * each iteration sum the filled genotypes on the int8 matrix cores, with
 * exact int32 sums recombined in FP64:
 *   forward (num_W, src/nmf.rs:131-147), on a sample-major copy of the rows:
 *   H in 3 int8 terms against the dosages and the products H_r H_s in 2
 *   against the missing-call indicator (the fill W_k . H_j enters as
 *   (sum over missing j of H_j H_j^T) W_k);
 *   backward (the H update, src/nmf.rs:171-221): W in 3 terms against the
 *   dosages, W_r W_s in 2 against the indicator (the fill enters as
 *   (sum over missing k of W_k W_k^T) H_j).
 * W and H then agree with the FP64 path to about 1e-6 of their largest
 * entries; everything else (initialisation, updates, CV) stays FP64. */
int slfmm_nmf_estimate(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* subset_idx, uint64_t n_idx,
                       const slfmm_nmf_config* cfg, const double* omega, double* w, double* h,
                       double* cv_mae, double* cv_mae_mean, uint64_t* mask_count, uint32_t flags);

/* One problem of a batched fit: the arguments of slfmm_fit_lfmm2. */
typedef struct slfmm_problem {
  const slfmm_geno* est;
  const int64_t* subset_idx;
  uint64_t n_idx;
  const slfmm_geno* g;
  const double* x;
  uint64_t d;
  slfmm_config cfg;
  const double* b0;
  double* u_hat;
  slfmm_assoc_out out;
} slfmm_problem;

/* Many independent fit_lfmm2 problems (each src/lib.rs:162-263) in one call.
 * No counterpart in the reference, which fits one problem per process; this
 * is the batching pattern of the hot path (many small LFMM2 problems, uneven
 * sizes).  The problems are spread over n_streams (1..16) internal streams
 * of the context, longest estimated first, each to the stream with the least
 * queued work, so small problems overlap on the device; the host synchronises
 * once per stream at the end.  With SLFMM_OUT_DEVICE each problem's work is
 * replayed from a captured graph (see SLFMM_EAGER).  Each problem's results are those of
 * slfmm_fit_lfmm2 on the same inputs; outputs follow `flags` as there.  An
 * error names the problem it came from ("problem i: ..."). */
int slfmm_fit_lfmm2_batch(slfmm_ctx* ctx, slfmm_problem* problems, uint64_t count, int n_streams,
                          uint32_t flags);

/* fit_lfmm2 with NMF imputation (src/lib.rs:162-263, ImputeConfig::Nmf*),
 * FP64: NMF estimation on est's subset, W^+ = safe_inv(W^T W) W^T, the RSVD
 * on the NMF-filled estimation rows (NmfInRam) and the association on every
 * SNP of g with NMF on-the-fly fill (NmfOnTheFly).  nmf_omega
 * (n x (nmf k + 10)) may be injected; nmf_cv_mae (n_iter + 1, host) receives
 * the estimation CV error per iteration, then the mean-imputation error on the
 * final iteration's mask (results.nmf_cv_mean, src/lib.rs:259); gwas_cv (3, host: NMF MAE, mean-fill
 * MAE, held-out count) enables the association-phase CV with rate
 * nmf->cv_rate and seed cfg->seed (NULL disables it).  With SLFMM_NMF_INT8
 * in flags the NMF estimation runs its update passes on the int8 matrix
 * cores (nmf k <= 8, as slfmm_nmf_estimate).  With cfg->precision
 * SLFMM_PREC_I8 (nmf k <= 8) the passes over the NMF-filled rows run there
 * too: the RSVD sketch/backward and forward passes and the association
 * functionals split each filled sum into the dosages against the operand (3
 * int8 terms) and the missing-call indicator against the operand's products
 * with W (or H) (2 terms), recombined in FP64 with the SNP's own H_j (the
 * sample's own W_k); the fill moments and the association's H_j run there
 * as well; counts, the NMF initialisation, the CV and the epilogue stay
 * FP64.  Layouts past the int8 tile plans run those passes in FP64. */
int slfmm_fit_lfmm2_nmf(slfmm_ctx* ctx, const slfmm_geno* est, const int64_t* subset_idx, uint64_t n_idx,
                        const slfmm_geno* g, const double* x, uint64_t d, const slfmm_config* cfg,
                        const slfmm_nmf_config* nmf, const double* b0, const double* nmf_omega, double* u_hat,
                        slfmm_assoc_out* out, double* nmf_cv_mae, double* gwas_cv, uint32_t flags);

/* ---- Launch shapes (no reference counterpart) ----
 * The tile and launch shapes that the kernels pick per problem, as a table
 * of thresholds and targets.  The defaults are the measured MI355X choices;
 * an offline tuning pass (python -m schnellfmm_amd.autotune) times the
 * candidates on the device and writes a table that the host layer applies
 * to every context it creates.  The tile choice leaves the results bit for
 * bit unchanged; the workgroup targets change only how FP64 partial sums are
 * grouped before they are added (rounding level). */
typedef struct {
  /* int8 SNP-side kernels (association pass, RSVD sketch) with at most 3 row
   * tiles: 2 tiles of 32 SNPs per wave (256 SNPs per workgroup) when that
   * grid holds at least this many workgroup blocks (workgroups x 256-sample
   * blocks per row), else 1 tile (128 SNPs per workgroup: twice the
   * workgroups for short or narrow panels).  0: always 2. */
  uint32_t snp_two_tiles_min_blocks;
  /* int8 RSVD forward pass: target number of workgroups (sample blocks x
   * SNP chunks); fewer chunks mean fewer partial sums to reduce. */
  uint32_t fwd_target_wg;
  /* tall-skinny Gram: maximum number of row blocks (partial Grams). */
  uint32_t gram_max_blocks;
  /* batched fits: the schedule's estimate of a problem's device time,
   * batch_fit_ms + (genotype bytes streamed) / (batch_gb_per_ms 1e9). */
  double batch_fit_ms;
  double batch_gb_per_ms;
} slfmm_tuning;

/* Replace the context's launch shapes (its batch streams take them too). */
int slfmm_set_tuning(slfmm_ctx* ctx, const slfmm_tuning* t);
/* The context's launch shapes (the defaults until slfmm_set_tuning). */
int slfmm_get_tuning(slfmm_ctx* ctx, slfmm_tuning* t);

/* ---- Kernel timing (no reference counterpart; used by the benchmark) ---- */

/* Enable HIP-event timing of the genotype-stream kernels on the context's
 * stream (names: "assoc_stream", "rsvd_project", "rsvd_forward"). */
int slfmm_profile_enable(slfmm_ctx* ctx, int32_t on);
/* Accumulated device time (ms) and launch count of one timed kernel; reset=1
 * clears it after reading. */
int slfmm_profile_read(slfmm_ctx* ctx, const char* name, double* total_ms, uint64_t* launches,
                       int32_t reset);
/* Fits queued as HIP graphs on the context and its batch streams since it
 * was created: graphs captured, and replays of a cached graph (no reference
 * counterpart). */
int slfmm_graph_stats(slfmm_ctx* ctx, uint64_t* captured, uint64_t* replayed);

/* Diagnostic (no reference counterpart): one int8 MFMA 16x16x64 with the
 * fragment mapping of the mixed-precision kernels; D = A (16x64) B (64x16),
 * all row-major.  A GPU test pins the lane layout with it. */
int slfmm_debug_mfma_i8(slfmm_ctx* ctx, const int8_t* a, const int8_t* b, int32_t* d);

/* Diagnostic (no reference counterpart): sustained dense MFMA rate on the
 * whole chip, in tera-operations per second (2 per multiply-add), of kind 0
 * v_mfma_i32_32x32x32_i8, 1 v_mfma_scale_f32_32x32x64_f8f6f4 with FP8 e4m3
 * operands, 2 v_mfma_f32_32x32x16_bf16: iters MFMAs on each of 4
 * accumulators per wave, 2 waves per SIMD.  The benchmark's precision
 * comparison prices int8, FP8 and BF16 passes with it. */
int slfmm_debug_mfma_rate(slfmm_ctx* ctx, int32_t kind, uint64_t iters, double* tops);

/* Diagnostic (no reference counterpart, host only): the packed-operand row
 * layout of the int8 SNP-side kernels for m functionals, columns >= k3 with 3
 * dosage-side terms and columns >= km with 2 miss-side terms.
 * dims[5] = {NQ, NT0, NT1, g1, ones_g}; rows (cap >= 32 (NT0 + NT1) ints,
 * may be NULL) = (column, term) per packed row, column -1 padding, -2 the
 * all-ones row.  A CPU test pins the layout with it. */
int slfmm_debug_i8_layout(int32_t m, int32_t k3, int32_t km, int32_t* dims, int32_t* rows, uint64_t cap);

/* ---- Synthetic data (test / benchmark inputs, no reference counterpart) --- */

/* Balding-Nichols panel generated on the device into a packed SNP-major
 * buffer owned by a new geno handle: n samples split evenly into n_pops
 * populations, F_k ~ U(fst_lo, fst_hi), p_anc ~ U(maf_lo, maf_hi),
 * missing calls with probability miss_rate. */
int slfmm_synth_bn(slfmm_ctx* ctx, uint64_t n, uint64_t n_snps, uint32_t n_pops, double fst_lo,
                   double fst_hi, double maf_lo, double maf_hi, double miss_rate, uint64_t seed,
                   slfmm_geno** out);

/* Generalised synthetic panel (no reference counterpart): admixed samples,
 * inbred lines, Beta-distributed ancestral frequencies and structured
 * missingness.  Per SNP: p_anc ~ U(maf_a, maf_b) or Beta(maf_a, maf_b)
 * clamped to [maf_min, 1 - maf_min]; per population k: F_k ~ U(fst_lo, fst_hi),
 * p_k ~ Beta(p_anc (1-F)/F, (1-p_anc)(1-F)/F).  Sample i's frequency is
 * sum_k ancestry[i][k] p_k (or p of its group when ancestry is NULL: n_pops
 * equal contiguous groups).  Genotype: Bin(2, p) (het_rate < 0) or an inbred
 * line, a het call with probability het_rate and else 2 Bernoulli(p).
 * Missing with probability min(1, miss_scale (1 - (1 - miss_rate)(1 - a_i)
 * (1 - b_j))), a_i = sample_miss[i] (or 0), b_j ~ Beta(snp_miss_a,
 * snp_miss_b) (0 when snp_miss_a <= 0); miss_scale <= 0 means 1.  Host
 * arrays; n_pops <= 64. */
typedef struct {
  uint32_t n_pops;
  double fst_lo, fst_hi;
  int32_t maf_beta;
  double maf_a, maf_b, maf_min;
  const float* ancestry;
  double het_rate;
  double miss_rate;
  const float* sample_miss;
  double snp_miss_a, snp_miss_b;
  double miss_scale;
  uint64_t seed;
} slfmm_synth_spec;
int slfmm_synth_bn_ex(slfmm_ctx* ctx, uint64_t n, uint64_t n_snps, const slfmm_synth_spec* spec, slfmm_geno** out);

/* ---- Output stage (src/testing.rs:740-801) ------------------------------ */

/* Writes the calibrated results table of coalesce_output_tsv: header
 * "chr pos snp_id {p_<c> beta_<c> t_<c>} r2_cov r2_latent r2_resid", tab
 * separated, one row per SNP in SNP order, every number in Rust's `{:.6e}`
 * form.  SNP metadata comes from the .bim at `bim_path` (validated as
 * parse_bim, src/bed.rs:480-503; it must hold p records) or, when bim_path is
 * NULL, from the arrays chrom / pos / snp_id.  cov_names may be NULL
 * ("cov_0", "cov_1", ...).  pvalue, beta, tstat are p x d and r2 p x 3
 * (row-major, as slfmm_fit_lfmm2 returns them); with SLFMM_OUT_DEVICE they are
 * device pointers (ctx required).  ctx may be NULL for host arrays.
 * n_threads <= 0: up to 16 formatting threads. */
int slfmm_write_tsv(slfmm_ctx* ctx, const char* path, const char* bim_path, const char* const* chrom,
                    const uint64_t* pos, const char* const* snp_id, uint64_t p, uint64_t d,
                    const char* const* cov_names, const double* pvalue, const double* beta,
                    const double* tstat, const double* r2, uint32_t flags, int n_threads);

/* The number formatter of slfmm_write_tsv on its own: x[i] in Rust's `{:.6e}`
 * form, NUL-terminated, at out + i * stride (stride >= 16).  No counterpart in
 * the reference (Rust's core::fmt does this there); exported for callers that
 * write their own tables. */
int slfmm_format_e6(const double* x, uint64_t count, char* out, uint64_t stride);

#ifdef __cplusplus
}
#endif
#endif /* SLFMM_H */
