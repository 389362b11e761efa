// Internal declarations shared by the slfmm HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <new>
#include <stdexcept>
#include <set>
#include <string>
#include <vector>

#include "../../include/slfmm.h"

namespace slfmm {

// Error carried from deep inside a call back to the C ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Thrown by scratch() when a buffer would have to be allocated while the
// context's stream is being captured into a graph; the capture is dropped
// and the work is queued directly.
struct CaptureAbort : std::runtime_error {
  CaptureAbort() : std::runtime_error("graph capture needs a new device buffer") {}
};

#define SLFMM_HIP(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::slfmm::Error(SLFMM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define SLFMM_CHECK(cond, code, msg)                 \
  do {                                               \
    if (!(cond)) throw ::slfmm::Error((code), (msg)); \
  } while (0)

constexpr int kBins = 100000;        // TsqHistogram bins (src/testing.rs:125)
constexpr double kBinWidth = 0.001;  // (src/testing.rs:126)
constexpr int kMaxRows = 64;         // max functionals per SNP handled by one launch row-block
constexpr int kRowBlk = 32;          // rows per FP64 functional row-block
// Zero bytes allocated after the last packed row: kernels load whole
// 16-byte segments of a 256-sample block and may run past ld by < 128 bytes.
constexpr uint64_t kSlack = 128;
constexpr int kMaxM = 64;   // functionals per SNP (K + d, or K + oversampling)
constexpr int kNmfMaxK = 16; // NMF components
constexpr int kNmfMaxK8 = 8;  // NMF rank of the int8 matrix-core backward pass
// every NMF rank 1 .. kNmfMaxK gets its own kernel instantiation
#define SLFMM_NMF_RANKS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
static_assert(kNmfMaxK == 16, "rank instantiations cover 1 .. kNmfMaxK");
constexpr int kMaxD = 16;   // traits

__host__ __device__ inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
__host__ __device__ inline uint64_t rup(uint64_t a, uint64_t b) { return cdiv(a, b) * b; }

}  // namespace slfmm

// Opaque handle definitions.
struct slfmm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Launch shapes (slfmm_set_tuning); the defaults are the measured MI355X choices.
  slfmm_tuning tuning{0u, 512u, 1024u, 1.4, 2.3};
  // Panels bound to this context that are still alive.  slfmm_destroy on a
  // context with live panels only marks it closed; the last slfmm_geno_free
  // deletes it, so the two may be called in either order.
  int live_genos = 0;
  bool closed = false;
  std::string err;
  // Named, grow-only device scratch buffers reused across calls.
  std::map<std::string, std::pair<void*, size_t>> pool;
  // Allocation-free calls: while `arena` is set, scratch() hands out 256-byte
  // aligned slices of that device buffer (bump allocation, name -> slice)
  // instead of allocating; a request past arena_cap throws SLFMM_E_NOMEM.
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_top = 0;
  std::map<std::string, std::pair<size_t, size_t>> arena_map;
  void* scratch(const std::string& name, size_t bytes);
  template <class T>
  T* buf(const std::string& name, size_t count) {
    return static_cast<T*>(scratch(name, count * sizeof(T)));
  }
  // workgroups of a kernel the device holds at once (CUs x occupancy), by kernel
  std::map<const void*, uint64_t> resident_wgs;
  // HIP-event timing of named kernels (enabled by slfmm_profile_enable)
  bool profiling = false;
  struct Timer { double ms = 0.0; uint64_t launches = 0; std::vector<std::pair<hipEvent_t, hipEvent_t>> pending; };
  std::map<std::string, Timer> timers;
  std::pair<hipEvent_t, hipEvent_t> tic(const std::string& name);
  void toc(const std::string& name, std::pair<hipEvent_t, hipEvent_t> ev);
  void collect();
  // Device status words of the current API call.  Kernels that can meet a
  // singular matrix write 0 (ok), 1 (regularised / shifted) or 2 (failed)
  // into their slot; the host reads every slot once, after the call's work
  // is queued (one synchronisation instead of one per small solve).
  struct Check { std::string warn; int code; std::string err; };
  static constexpr int kMaxChecks = 4096;
  int* status = nullptr;
  std::vector<Check> checks;
  // warn is printed to stderr for value 1 (if non-empty); value 2 raises
  // Error(code, err) unless code is SLFMM_OK.
  int* check_slot(const std::string& warn, int code, const std::string& err);
  std::string check_prefix;   // prepended to the error text of new slots (batched fits)
  // Streaming of host-resident panels: a copy stream (default priority),
  // two pinned staging buffers and their events, created on first use.
  hipStream_t copy_stream = nullptr;
  void* pinned[2] = {nullptr, nullptr};
  size_t pinned_bytes[2] = {0, 0};
  hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_consumed[2] = {nullptr, nullptr};
  void* pinned_buf(int i, size_t bytes);   // grow-only, hipHostMalloc
  void ensure_copy_stream();
  // A second compute stream (default priority) for independent work that
  // overlaps the main stream's chain of small kernels: forked from and
  // joined back into `stream` with the two events (capturable).
  // Stages that SLFMM_PREC_I8 asked for but whose shape runs the FP64
  // kernels, each reported once per context on stderr (fp64_stage).
  std::set<std::string> fp64_noted;
  hipStream_t side_stream = nullptr;
  bool batch_worker = false;   // a stream of slfmm_fit_lfmm2_batch (problems already overlap)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void ensure_side_stream();
  // Uploads of host arrays: copied into a grow-only pinned buffer and queued
  // with hipMemcpyAsync from there, so the call returns without waiting for
  // the copy (a pageable source would be staged synchronously).  Slices are
  // bump-allocated within one API call; the next call reuses the buffer
  // after the event of the last queued upload.  In allocation-free calls the
  // copy is queued straight from the caller's memory.
  unsigned char* up_buf = nullptr;
  size_t up_cap = 0, up_top = 0;
  hipEvent_t ev_up = nullptr;
  void upload(void* dst, const void* src, size_t bytes);
  void upload_reset();
  // Graph replay of whole fits.  pool_gen counts pool (re)allocations: a
  // graph captured at one generation is replayed only at that generation.
  // While `capturing`, scratch() may not allocate (CaptureAbort), and
  // gather_subset takes the estimation indices that were uploaded before the
  // capture (staged_idx, the caller's host pointer) from its buffer.
  uint64_t pool_gen = 0;
  bool capturing = false;
  const int64_t* staged_idx = nullptr;
  struct FitGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<Check> checks;   // the status slots the graph writes, from slot `slot0` on
    uint64_t pool_gen = 0;
  };
  std::map<std::string, FitGraph> graphs;
  uint64_t graphs_captured = 0, graphs_replayed = 0;
  // Contexts of the extra streams of batched fits (same device, own scratch).
  std::vector<slfmm_ctx*> workers;
  // Reads the slots (synchronises the stream), prints warnings in order and
  // throws the first error; returns the values.  Clears the slot list.
  std::vector<int> finish_checks();
  ~slfmm_ctx();
};

// RAII timing scope for one kernel launch on ctx->stream.
struct slfmm_timed {
  slfmm_ctx* ctx; std::string name; std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  slfmm_timed(slfmm_ctx* c, const char* n) : ctx(c), name(n) { if (ctx->profiling) ev = ctx->tic(name); }
  ~slfmm_timed() { if (ctx->profiling && ev.first) ctx->toc(name, ev); }
};

struct slfmm_geno {
  slfmm_ctx* ctx = nullptr;
  uint8_t* d = nullptr;   // packed codes, SNP-major, row stride ld bytes
  uint64_t n = 0;         // samples represented (after sample_keep)
  uint64_t p = 0;         // SNPs
  uint64_t ld = 0;        // row stride in bytes (multiple of 16)
  bool owned = false;
  // Host-resident panel (streamed to the device chunk by chunk): rows of
  // bps = ceil(n/4) bytes at `host` (a mapped .bed or caller memory); d is null.
  const uint8_t* host = nullptr;
  uint64_t bps = 0;
  bool host_pinned = false;     // host rows are page-locked: copied without staging
  uint64_t chunk_snps = 0;      // rows per streamed chunk (0: automatic)
  void* map_base = nullptr;     // owned mapping of the .bed file
  size_t map_len = 0;
};

namespace slfmm {

// ---- host-resident panels (api.cpp) ---------------------------------------
// Copies rows rows[0..m) (or 0..m when rows is null) of a host panel to dst
// (device, row stride g->ld, padding bytes zero) through the pinned buffers,
// on ctx->stream; returns after the copies are queued and the pinned buffers
// are free again.
void host_rows_to_device(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* rows, uint64_t m, uint8_t* dst);
// Rows per streamed chunk of a host panel.
uint64_t host_chunk_rows(const slfmm_geno* g);
// Fills pinned staging buffer b with rows [j0, j1) of host panel g (several
// host threads) and queues its copy to dst (row stride g->ld) on the copy
// stream; records ev_copied[b].
void host_chunk_upload(slfmm_ctx* ctx, const slfmm_geno* g, uint64_t j0, uint64_t j1, int b, uint8_t* dst);

// Body of every C entry point: runs f with the context's device current,
// turns exceptions into status codes and the context's error text, and reads
// the device status words the call queued.
// Host-only entry points may run without a context; their error text is kept
// per thread (slfmm_last_error(NULL)).
std::string& no_ctx_error();

template <class F>
int guarded(slfmm_ctx* ctx, F&& f) {
  std::string& err = ctx ? ctx->err : no_ctx_error();
  try {
    if (ctx) {
      SLFMM_HIP(hipSetDevice(ctx->device));
      ctx->checks.clear();
      ctx->upload_reset();
    }
    f();
    if (ctx) ctx->finish_checks();
    err.clear();
    return SLFMM_OK;
  } catch (const Error& e) {
    err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    err = "host allocation failed";
    return SLFMM_E_NOMEM;
  } catch (const std::exception& e) {
    err = e.what();
    return SLFMM_E_ARG;
  }
}

// ---- geno.hip -------------------------------------------------------------
uint64_t geno_ld(uint64_t n);
// Zeroes the codes of the padding samples in the last byte of rows [0, p).
void geno_clear_tail(hipStream_t s, uint8_t* d, uint64_t ld, uint64_t p, uint64_t n);
void geno_alloc(slfmm_ctx* ctx, slfmm_geno* g);
void geno_upload(slfmm_ctx* ctx, const uint8_t* packed, uint64_t n_phys, uint64_t n_snps,
                 const int64_t* keep, uint64_t n_keep, slfmm_geno* g);
// Gather rows idx[0..m) of g into dst (m x g->ld), contiguous.
void geno_gather(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* d_idx, uint64_t m, uint8_t* dst);
// Integer stats per SNP (rows of a packed matrix).
void launch_stats(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                  int32_t* n_miss, int32_t* sum_g, int32_t* n_hom);
// mean and scale per SNP from the integer stats (compute_scale semantics).
void launch_mean_scale(hipStream_t s, const int32_t* n_miss, const int32_t* sum_g, uint64_t n,
                       uint64_t p, int norm, double* mu, double* sc);
void launch_decode_f64(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t n,
                       const int64_t* idx, uint64_t c, int norm, const double* fill, double* out,
                       int raw);
void launch_synth_bn(hipStream_t s, uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                     uint32_t n_pops, double fst_lo, double fst_hi, double maf_lo, double maf_hi,
                     double miss_rate, uint64_t seed);
void launch_synth_ex(hipStream_t s, uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p, const slfmm_synth_spec& spec);
// Transposed copy (sample-major packed: row = sample, 4 SNPs per byte).
void launch_transpose(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                      uint8_t* out, uint64_t ldt);

// ---- dense.hip: tall-skinny FP64 helpers and one-workgroup small solvers ---
// Kernels taking `const int* gate` do nothing unless gate is null or *gate == 1.
// out (a x b) = A^T B, A rows x a (lda), B rows x b (ldb); deterministic.
void ts_gram(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t a, const double* B,
             uint64_t ldb, uint64_t b, uint64_t rows, double* out, const int* gate = nullptr);
// out (rows x q, ldo) = alpha * A (rows x m, lda) * M' (m x q) + beta * out, with
// M' = M, or (dlam set) row t of M times -(1 - dlam[t]) (inverse: 1 / dlam[t] - 1)
void ts_gemm(hipStream_t s, const double* A, uint64_t lda, uint64_t m, const double* M,
             uint64_t ldm, uint64_t q, uint64_t rows, double alpha, double beta, double* out,
             uint64_t ldo, const int* gate = nullptr, const double* dlam = nullptr, bool inverse = false);
void launch_fill(hipStream_t s, double* p, double v, uint64_t count);
// Reports once per context that a stage requested in int8 runs its FP64
// kernels because of its shape (stage: what runs in FP64; why: the limit).
void fp64_stage(slfmm_ctx* ctx, const std::string& stage, const std::string& why);
// dst (rows x cols, ldd) <- src (rows x cols, lds)
void copy2d(hipStream_t s, const double* src, uint64_t lds, double* dst, uint64_t ldd, uint64_t rows,
            uint64_t cols, const int* gate = nullptr);
// Column reductions of a tall matrix (rows x m, lda), deterministic order.
void colmeans(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t rows, uint64_t m, double* means);
// out[c] = max_r |A[r][c] * rs[r] * rm[r]| / 127 (1 if zero); rs, rm optional.
// With rm2 given, also out[m + c] for the row multiplier rs * rm2.
void colabsmax_scale(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t rows, uint64_t m,
                     const double* rs, const double* rm2, double* out);
// In-place thin QR of A (rows x m, lda) by CholeskyQR2 (a shifted first
// pass, then two clean ones, if plain Cholesky breaks down); R (m x m upper)
// to d_R if non-null.  *d_st <- 0 ok, 1 ok after the shift, 2 not full rank.
void cholqr2(slfmm_ctx* ctx, double* A, uint64_t lda, uint64_t m, uint64_t rows, double* d_R, int* d_st,
             bool shift_all = false);
// One shifted-CholeskyQR pass, out = A R^{-1} (column span only; see dense.hip).
void cholqr1(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t m, uint64_t rows, double* out, uint64_t ldo,
             int* st);
// safe_inv on device (m <= 64): *d_st <- 0 ok, 1 regularised, 2 failed.
void safe_inv_dev(slfmm_ctx* ctx, const double* d_a, uint64_t m, double* d_out, int* d_st);
// out (m x m) <- In + lambda I
void add_diag(hipStream_t s, const double* In, uint64_t m, double lambda, double* out);
// Symmetric/one-sided Jacobi on an m x m matrix (m <= 64): columns of
// A V orthogonal; outputs V (m x m) and column norms, sorted descending.
void jacobi_svd_small(slfmm_ctx* ctx, const double* d_a, uint64_t m, double* d_v, double* d_sv);

// Missing calls of a streamed panel filled with W_k . H_j (NMF imputation,
// src/parallel.rs:233-321): W is n x K, Ht holds H_j for every streamed row
// j (rows x K); both device pointers.  W == nullptr means mean imputation.
struct FillSpec {
  const double* W = nullptr;
  const double* Ht = nullptr;
  int K = 0;
  uint64_t n = 0;
};

// ---- assoc.hip --------------------------------------------------------------
struct AssocPlan;   // defined in assoc.hip
// FP64 functionals F[j][c] = sum_k A[k][c] v_kj of the centred rows (v = g - mu,
// missing calls 0, or W_k . H_j - mu with a fill) for the columns c < m; A is
// npad x lda (lda >= rup(m, 32), zero past column m); F rows scaled by
// outscale.
void launch_func_f64(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t p, uint64_t npad,
                     const double* A, uint64_t lda, int m, const double* mu, double* F, uint64_t ldF,
                     const double* outscale, const FillSpec* fill = nullptr);
void pad_operand(hipStream_t s, const double* Q, uint64_t ldq, uint64_t n, int m, double* A, uint64_t lda,
                 uint64_t npad);
// Per-row moments of a filled panel: mu = (sum_g + S1) / n over all calls,
// sc = compute_scale(mu, norm), tss = sc^2 sum_k (y_k - mu)^2 (y filled).
// With Wp (n x K) also SWp[j] = sum over missing k of Wp[k] (rows x K).
// With i8 (rank <= 8) the fill sums come from fill_stats_i8.
void fill_moments(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t rows, const FillSpec& fs,
                  int norm, const double* Wp, double* mu, double* sc, double* tss, double* SWp, bool i8 = false);
// Association with NMF on-the-fly imputation (src/parallel.rs:252-282): the
// caller passes W, the pseudo-inverse rows Wp = W (W^T W)^-1 and the GWAS
// CV settings; cv_out (3) receives (NMF MAE, mean MAE, count) if non-null.
struct NmfAssoc {
  const double* W = nullptr;
  const double* Wp = nullptr;
  int K = 0;
  double cv_rate = 0.0;   // 0: no GWAS-phase CV
  uint64_t cv_seed = 0;
  uint64_t chunk = 10000;
  double* cv_out = nullptr;
};
void assoc_run(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_xs, uint64_t d,
               const double* d_u, uint64_t k, double lambda, const slfmm_config& cfg,
               double* d_beta, double* d_t, double* d_r2, double* d_p, uint64_t* d_hist,
               double* d_gif_pt, double* d_gif, const NmfAssoc* nmf = nullptr);

// ---- assoc_i8.hip: single-MFMA lane-layout probe (test-only entry point) ----
void probe_mfma_i8(slfmm_ctx* ctx, const int8_t* hA, const int8_t* hB, int32_t* hD);
void i8_layout(int m, int k3, int km, int32_t* dims, int32_t* rows, uint64_t cap);
// ---- mfma_rate.hip: sustained MFMA rate of int8 / FP8 / BF16 (diagnostic) ----
void mfma_rate(slfmm_ctx* ctx, int kind, uint64_t iters, double* tops);
// ---- noalloc.hip ----------------------------------------------------------
// Resolves the device address of the allocation-free entry point's static
// scratch array on the current device (once per device).
void noalloc_register(int device);

// ---- rsvd.hip ---------------------------------------------------------------
// fill (optional): NMF in-RAM imputation of the estimation rows, with
// fill.Ht indexed by estimation row (src/parallel.rs:284-299).
void rsvd_run(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* h_idx, uint64_t n_idx,
              const double* d_xs, uint64_t d, const slfmm_config& cfg, const double* d_b0,
              double* d_uhat, const FillSpec* fill = nullptr);
// N(0,1) matrix (n x l, row stride ld) from the library's counter-based stream.
void launch_normal(hipStream_t s, double* out, uint64_t n, uint64_t l, uint64_t ld, uint64_t seed);
// Rows h_idx of g as a contiguous block (g's buffer itself for the identity).
const uint8_t* gather_subset(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* h_idx, uint64_t n_idx,
                             const char* tag, uint64_t* pe);
// Thin precompute on the device: sigma(d), dlam(d), Qd (n x d), ridge_inv (d x d).
void precompute_dev(slfmm_ctx* ctx, const double* d_xs, uint64_t n, uint64_t d, double lambda,
                    double* d_sigma, double* d_dlam, double* d_qd, double* d_ridge);

// ---- nmf.hip ------------------------------------------------------------------
// NMF estimation (src/nmf.rs:80-249) on packed rows Y (pe SNPs): device
// outputs W (n x K), Ht (pe x K) and cv (3 per iteration: NMF MAE, mean MAE,
// held-out count); omega (n x (K + 10)) is drawn from cfg.seed when null.
// With i8 (and K <= kNmfMaxK8) the forward pass (num_W, on a sample-major
// copy of Y) and the backward pass (the H update) run on the int8 matrix
// cores (nmf_forward_i8, nmf_backward_i8); everything else is FP64.
void nmf_run(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const slfmm_nmf_config& cfg,
             const double* d_omega, double* W, double* Ht, double* cv, bool i8 = false);
// assoc_i8.hip: H^T (pe x K) <- the NMF H update (src/nmf.rs:171-221) from W
// (n x K) and W^T W on the rows Y, on the int8 matrix cores (K <= kNmfMaxK8).
void nmf_backward_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const double* W, int K,
                     const double* WtW, double eps, double* Ht);
// assoc_i8.hip: RSVD passes on the int8 matrix cores (l <= 32).  Sketch /
// backward: Z[j][c] = s_j (B^T y_j)_c for every SNP row (y normalised by
// norm, missing calls at the mean).  Forward: W = Y Q_z on the sample-major
// copy packedT.
void rsvd_project_i8(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                     const double* d_B, uint64_t l, int norm, double* Z, uint64_t ldZ);
void rsvd_accumulate_i8(slfmm_ctx* ctx, const uint8_t* packedT, uint64_t ldt, uint64_t n, uint64_t p,
                        const double* Qz, uint64_t ldq, uint64_t l, const double* mu, const double* sc,
                        double* W, uint64_t ldw);
// assoc_i8.hip: centred functionals of NMF-filled rows on the int8 matrix
// cores, F[j][c] = sum_k A[k][c] (y_kj - mu_j) for c < m with the missing
// calls of SNP j filled by W_k . H_j (W n x K, Ht p x K), times outscale[j]
// if given; A is n x m (row stride lda).  Supported when fill rank K <= 8 and
// the column groups fit the mode-3 tile plans (fill_functionals_i8_supported).
bool fill_functionals_i8_supported(int m, int K);
void fill_functionals_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t p, const double* A,
                         uint64_t lda, int m, const double* W, const double* Ht, int K, const double* mu,
                         const double* outscale, double* F, uint64_t ldF);
// assoc_i8.hip: S1 = sum of the fills, S2 = sum of their squares and (with
// Wp) SWp = sum of the Wp rows over each row's missing calls, on the int8
// matrix cores (rank K <= 8); the FP64 counterpart is fill_moments' kernel.
void fill_stats_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t p, const double* W,
                   const double* Ht, int K, const double* Wp, double* S1, double* S2, double* SWp);
// assoc_i8.hip: the RSVD forward pass over NMF-filled rows (before N is
// applied) on the int8 matrix cores: W[k][c] = sum_j (y_kj - mu_j) sc_j Qz[j][c] for c < l
// with the missing calls of sample k filled by W_k . H_j (Wf n x K, Ht pe x K),
// from the sample-major copy YT (n x ldt bytes).  Fill rank K <= 8.
bool fill_forward_i8_supported(int l, int K);
void fill_forward_i8(slfmm_ctx* ctx, const uint8_t* YT, uint64_t ldt, uint64_t n, uint64_t pe, const double* Qz,
                     uint64_t ldq, int l, const double* mu, const double* sc, const double* Wf, const double* Ht, int K,
                     double* W, uint64_t ldw);
// assoc_i8.hip: num_W (n x K) of the NMF forward pass (src/nmf.rs:131-147,
// missing calls filled by W_k . H_j) from the sample-major copy YT (n x ldt
// bytes) of the rows, Ht (pe x K) and W (K <= kNmfMaxK8).
void nmf_forward_i8(slfmm_ctx* ctx, const uint8_t* YT, uint64_t ldt, uint64_t n, uint64_t pe, const double* Ht,
                    const double* W, int K, double* num);
void nmf_transpose_h(hipStream_t s, const double* Ht, uint64_t pe, int K, double* H);
// Ht[j][r] = max(F[j][r], 0)
void relu_rows(hipStream_t s, const double* F, uint64_t ldf, uint64_t p, int K, double* Ht);
// GWAS-phase CV (src/testing.rs:423-487); out3 = (NMF MAE, mean MAE, count).
void gwas_cv(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p, const double* W,
             const double* Wp, int K, const double* Hraw, uint64_t ldh, const double* SWp, const int32_t* nmiss,
             const int32_t* sumg, uint64_t seed, uint64_t chunk, double cv_rate, double* out3);
// out[j][c] = sum_k Op[k][c] y_kj for the mean-filled raw rows (y = g, or
// the row mean on missing calls); ldout >= rup(m, 32).
void raw_functionals(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const double* mu,
                     const double* Op, uint64_t ldo, uint64_t m, double* out, uint64_t ldout);
// raw_functionals on the int8 matrix cores (m <= 32; no padding of Op).
void raw_functionals_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const double* mu,
                        const double* Op, uint64_t m, double* out, uint64_t ldout);

}  // namespace slfmm
