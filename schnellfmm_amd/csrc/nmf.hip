// NMF imputation of the estimation panel on the GPU (FP64, the reference's
// arithmetic): src/nmf.rs:80-455.
//
// Layout: W is n x K (row-major), H is kept transposed as Ht (p_est x K), so
// the K factor loadings of one SNP are contiguous.  Every pass reads the raw
// genotype codes of the estimation rows directly; missing calls are filled
// inside the kernels with the current reconstruction W_k . H_j (the
// reference's NmfRawInRam fill, src/parallel.rs:290-321), never stored.
//
//   forward   num_W[k] = sum_j y_kj H_j      (src/nmf.rs:131-147)
//   W update  W *= num_W / (W HH^T) where the denominator > eps, clamp,
//             normalise columns (src/nmf.rs:150-168, :259-271)
//   backward  per SNP: num_H = W^T y_j, H_j *= num_H / (W^T W H_j) where the
//             denominator > eps, clamp (src/nmf.rs:171-221), normalise
//   CV error  masked MAE against W H and against the column mean, one
//             uniform draw per entry (src/nmf.rs:379-455)
#include "internal.h"
#include "colreduce.h"
#include "rng.h"

namespace slfmm {

__device__ __forceinline__ uint32_t code_of(uint32_t word, int pos) { return (word >> (2 * pos)) & 3u; }
__device__ __forceinline__ double dosage_of(uint32_t code) { return code == 0u ? 0.0 : (code == 2u ? 1.0 : 2.0); }

// ---- forward pass: per-chunk partials of num_W (n x K) -----------------------
// One workgroup per (256-sample tile, SNP chunk); thread = sample.  The
// loadings H_j of the SNP in hand are the same for every thread, so they come
// in through scalar loads (one wide load per SNP: the rank K is a template
// parameter).
template <int K>
__global__ void __launch_bounds__(256) k_nmf_fwd(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t n,
                                                 uint64_t p, uint64_t snps_per_chunk, const double* __restrict__ W,
                                                 const double* __restrict__ Ht, double* __restrict__ part) {
  constexpr int KM = K;
  __shared__ uint32_t sw[32][16];
  const int t = threadIdx.x;
  const uint64_t k = (uint64_t)blockIdx.x * 256 + t;
  const uint64_t j0 = (uint64_t)blockIdx.y * snps_per_chunk;
  const uint64_t j1 = std::min(p, j0 + snps_per_chunk);
  double w[KM], acc[KM];
#pragma unroll
  for (int r = 0; r < KM; ++r) {
    w[r] = k < n ? W[k * K + r] : 0.0;
    acc[r] = 0.0;
  }
  for (uint64_t jb = j0; jb < j1; jb += 32) {
    const int nj = (int)std::min<uint64_t>(32, j1 - jb);
    __syncthreads();
    for (int e = t; e < 32 * 16; e += 256) {
      const int jj = e >> 4, wd = e & 15;
      sw[jj][wd] = jj < nj ? reinterpret_cast<const uint32_t*>(packed + (jb + jj) * ld + (uint64_t)blockIdx.x * 64)[wd]
                           : 0u;
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < nj; ++jj) {
      const double* hj = Ht + (jb + jj) * K;
      double h[KM];
#pragma unroll
      for (int r = 0; r < KM; ++r) h[r] = hj[r];
      const uint32_t code = code_of(sw[jj][t >> 4], t & 15);
      double f = 0.0;   // missing: current reconstruction
#pragma unroll
      for (int r = 0; r < KM; ++r) f += w[r] * h[r];
      const double v = code == 1u ? f : dosage_of(code);
#pragma unroll
      for (int r = 0; r < KM; ++r) acc[r] += v * h[r];
    }
  }
  if (k < n)
    for (int r = 0; r < K; ++r) part[((uint64_t)blockIdx.y * n + k) * K + r] = acc[r];
}

// num_W = sum of the chunk partials in chunk order
__global__ void k_nmf_sum_chunks(const double* __restrict__ part, uint64_t count, int nchunk, double* out) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(uint64_t)c * count + e];
  out[e] = s;
}

// W update, one thread per sample row (src/nmf.rs:150-167)
__global__ void k_nmf_wupd(double* __restrict__ W, const double* __restrict__ num, const double* __restrict__ hh,
                           uint64_t n, int K, double eps) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double w[kNmfMaxK], nw[kNmfMaxK];
  for (int r = 0; r < K; ++r) w[r] = W[k * K + r];
  for (int r = 0; r < K; ++r) {
    double d = 0.0;
    for (int s = 0; s < K; ++s) d += w[s] * hh[s * K + r];
    double v = w[r];
    if (d > eps) v *= num[k * K + r] / d;
    nw[r] = v > eps ? v : eps;
  }
  for (int r = 0; r < K; ++r) W[k * K + r] = nw[r];
}

// normalize_w (src/nmf.rs:259-271): W[:, r] *= 1 / s_r, H[r, :] *= s_r for
// column sums s_r > eps.
__global__ void k_nmf_scale(double* __restrict__ W, uint64_t n, double* __restrict__ Ht, uint64_t p, int K,
                            const double* __restrict__ colsum, double eps) {
  const uint64_t total = (n + p) * (uint64_t)K;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % K);
    const double s = colsum[r];
    if (!(s > eps)) continue;
    if (e < n * (uint64_t)K) W[e] *= 1.0 / s;
    else Ht[e - n * (uint64_t)K] *= s;
  }
}

// ---- backward pass with the H update fused (src/nmf.rs:171-221) --------------
// One thread per SNP; the W row of the sample in hand is the same for every
// thread (scalar loads; the rank K is a template parameter).
template <int K>
__global__ void __launch_bounds__(256) k_nmf_bwd(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t n,
                                                 uint64_t p, const double* __restrict__ W,
                                                 const double* __restrict__ WtW, double eps, double* Ht) {
  constexpr int KM = K;
  const int t = threadIdx.x;
  const uint64_t j = (uint64_t)blockIdx.x * 256 + t;
  const bool ok = j < p;
  double h[KM], acc[KM];
#pragma unroll
  for (int r = 0; r < KM; ++r) {
    h[r] = ok ? Ht[j * K + r] : 0.0;
    acc[r] = 0.0;
  }
  const uint4* row = reinterpret_cast<const uint4*>(packed + (ok ? j : 0) * ld);
  const uint64_t nkb = cdiv(n, 64);
  for (uint64_t kb = 0; kb < nkb; ++kb) {
    const uint4 q = ok ? row[kb] : make_uint4(0, 0, 0, 0);
    const uint32_t wd[4] = {q.x, q.y, q.z, q.w};
    const int lim = (int)std::min<uint64_t>(64, n - kb * 64);
#pragma unroll 4
    for (int kk = 0; kk < lim; ++kk) {
      const double* wk = W + (kb * 64 + kk) * K;
      double wr[KM];
#pragma unroll
      for (int r = 0; r < KM; ++r) wr[r] = wk[r];
      const uint32_t code = code_of(wd[kk >> 4], kk & 15);
      double f = 0.0;   // missing: fill with W_k . H_j (pre-update H)
#pragma unroll
      for (int r = 0; r < KM; ++r) f += wr[r] * h[r];
      const double v = code == 1u ? f : dosage_of(code);
#pragma unroll
      for (int r = 0; r < KM; ++r) acc[r] += v * wr[r];
    }
  }
  if (!ok) return;
  double nh[KM];
  for (int r = 0; r < K; ++r) {
    double d = 0.0;
    for (int s2 = 0; s2 < K; ++s2) d += WtW[r * K + s2] * h[s2];
    double v = h[r];
    if (d > eps) v *= acc[r] / d;
    nh[r] = v > eps ? v : eps;
  }
  for (int r = 0; r < K; ++r) Ht[j * K + r] = nh[r];
}

// ---- CV error (src/nmf.rs:379-455) -------------------------------------------
// Entry (sample k, SNP j) of chunk seq = j / chunk draws element
// (j % chunk) * n + k of the stream seeded mask_seed + seq (column, then
// row, every entry drawn); it is held out if the draw is < cv_rate and the
// call is observed.  Per-workgroup sums, combined in a fixed order.
// One thread per SNP.  The draws of a 16-code word are unrolled (a counter
// offset per code, a hit mask), and the rare drawn calls are then taken in
// sample order from the mask, so the hot loop has no branch and the sums
// keep the per-entry order.
// Iteration it draws chunk seq from the stream seeded seed + it + seq, so an
// iteration reuses the streams of the one before, shifted by a chunk: the
// nchunk + n_iter - 1 distinct streams are drawn once (k_cv_lists: the
// sample indices of the first kCvList draws below the rate of each (stream,
// row-in-chunk), in sample order, and their count), and each iteration
// walks its SNPs' lists; a SNP whose list overflowed draws again.
constexpr int kCvList = 16;

// Draw bits of 16 consecutive elements of a stream: bit c set if element
// base + c (counter ctr + c kStreamStep) falls below the rate.
__device__ __forceinline__ uint32_t cv_draw16(uint64_t ctr, uint64_t cv_thresh) {
  uint32_t hit = 0;
#pragma unroll
  for (int c = 0; c < 16; ++c) hit |= (mix_counter(ctr + (uint64_t)c * kStreamStep) >> 11) < cv_thresh ? 1u << c : 0u;
  return hit;
}

__global__ void __launch_bounds__(256) k_cv_lists(uint64_t n, uint64_t seed0, uint64_t ns, uint64_t jl_count,
                                                  uint64_t cv_thresh, uint32_t* __restrict__ lists,
                                                  uint8_t* __restrict__ counts) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ns * jl_count) return;
  const uint64_t srel = e / jl_count, jl = e % jl_count;
  const uint64_t ctr0 = stream_counter(seed0 + srel, jl * n);
  uint32_t cnt = 0;
  for (uint64_t k0 = 0; k0 < n; k0 += 16) {
    uint32_t hit = cv_draw16(ctr0 + k0 * kStreamStep, cv_thresh);
    if (n - k0 < 16) hit &= (1u << (n - k0)) - 1u;
    while (hit) {
      const int b = __builtin_ctz(hit);
      hit &= hit - 1;
      if (cnt < (uint32_t)kCvList) lists[e * kCvList + cnt] = (uint32_t)(k0 + b);
      cnt += cnt < 255u ? 1u : 0u;
    }
  }
  counts[e] = (uint8_t)cnt;
}

__global__ void __launch_bounds__(256) k_nmf_cv(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t n,
                                                uint64_t p, const double* __restrict__ W,
                                                const double* __restrict__ Ht, int K,
                                                const int32_t* __restrict__ nmiss, const int32_t* __restrict__ sumg,
                                                uint64_t mask_seed, uint64_t chunk, uint64_t cv_thresh,
                                                const uint32_t* __restrict__ lists, const uint8_t* __restrict__ counts,
                                                uint64_t jl_count, uint64_t it, double* __restrict__ part) {
  __shared__ double r0[256], r1[256];
  __shared__ unsigned long long r2[256];
  const int t = threadIdx.x;
  const uint64_t j = (uint64_t)blockIdx.x * 256 + t;
  double e_nmf = 0.0, e_mean = 0.0;
  unsigned long long cnt = 0;
  if (j < p && cv_thresh > 0) {
    const int64_t nobs = (int64_t)n - nmiss[j];
    const double mu = nobs > 0 ? (double)sumg[j] / (double)nobs : 0.0;
    auto held = [&](uint64_t k, uint32_t code) {   // a drawn call; held out if observed
      if (code == 1u) return;
      double rec = 0.0;
      for (int r = 0; r < K; ++r) rec += W[k * K + r] * Ht[j * K + r];
      const double g = dosage_of(code);
      e_nmf += fabs(g - rec);
      e_mean += fabs(g - mu);
      cnt += 1;
    };
    const uint64_t le = lists ? ((it + j / chunk) * jl_count + j % chunk) : 0;
    const uint32_t nl = lists ? counts[le] : 0u;
    if (lists && nl <= (uint32_t)kCvList) {
      const uint8_t* rowb = packed + j * ld;
      for (uint32_t i = 0; i < nl; ++i) {
        const uint32_t k = lists[le * kCvList + i];
        held(k, (uint32_t)(rowb[k >> 2] >> (2 * (k & 3))) & 3u);
      }
    } else {
      // draw base + k of the stream: u01 < cv_rate  <=>  (mix >> 11) < cv_thresh
      uint64_t ctr = stream_counter(mask_seed + j / chunk, (j % chunk) * n);
      const uint4* row = reinterpret_cast<const uint4*>(packed + j * ld);
      const uint64_t nkb = cdiv(n, 64);
      for (uint64_t kb = 0; kb < nkb; ++kb, ctr += 64 * kStreamStep) {
        const uint4 q = row[kb];
        const uint32_t wd[4] = {q.x, q.y, q.z, q.w};
        const int lim = (int)std::min<uint64_t>(64, n - kb * 64);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          uint32_t hit = cv_draw16(ctr + (uint64_t)(16 * w) * kStreamStep, cv_thresh);
          const int valid = lim - 16 * w;   // codes past the panel's last sample
          if (valid < 16) hit &= valid <= 0 ? 0u : (1u << valid) - 1u;
          while (hit) {
            const int b = __builtin_ctz(hit);
            hit &= hit - 1;
            held(kb * 64 + 16 * w + b, (wd[w] >> (2 * b)) & 3u);
          }
        }
      }
    }
  }
  r0[t] = e_nmf; r1[t] = e_mean; r2[t] = cnt;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { r0[t] += r0[t + s]; r1[t] += r1[t + s]; r2[t] += r2[t + s]; }
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x * 3 + 0] = r0[0];
    part[blockIdx.x * 3 + 1] = r1[0];
    part[blockIdx.x * 3 + 2] = (double)r2[0];
  }
}

// out[0..2] = (nmf MAE, mean-imputation MAE, held-out count) of one pass;
// one workgroup of 256 threads, partials combined in a fixed order.
__global__ void __launch_bounds__(256) k_nmf_cv_final(const double* __restrict__ part, int nblk, double* out) {
  __shared__ double s0[256], s1[256], s2[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = t; i < nblk; i += 256) { a += part[i * 3]; b += part[i * 3 + 1]; c += part[i * 3 + 2]; }
  s0[t] = a; s1[t] = b; s2[t] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) { s0[t] += s0[t + h]; s1[t] += s1[t + h]; s2[t] += s2[t + h]; }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = s2[0] > 0.0 ? s0[0] / s2[0] : 0.0;
    out[1] = s2[0] > 0.0 ? s1[0] / s2[0] : 0.0;
    out[2] = s2[0];
  }
}

// ---- initialisation helpers (src/nmf.rs:279-362) -------------------------------
// Z[j][c] += mu_j * colsum[c]: turns centred functionals into functionals of
// the mean-filled raw column (g on observed calls, mu on missing ones).
__global__ void k_add_mean_term(double* __restrict__ Z, uint64_t ldz, uint64_t p, int m, const double* __restrict__ mu,
                                const double* __restrict__ colsum) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p * (uint64_t)m) return;
  const uint64_t j = e / m;
  const int c = (int)(e % m);
  Z[j * ldz + c] += mu[j] * colsum[c];
}

__global__ void k_abs(double* __restrict__ a, uint64_t count) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) a[e] = fabs(a[e]);
}

// Ht[j][r] = max(F[j][r], lo)  (F with row stride ldf)
__global__ void k_relu_rows(const double* __restrict__ F, uint64_t ldf, uint64_t p, int K, double lo,
                            double* __restrict__ Ht) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p * (uint64_t)K) return;
  const double v = F[(e / K) * ldf + e % K];
  Ht[e] = v > lo ? v : lo;
}

__global__ void k_clamp_min(double* __restrict__ a, uint64_t count, double lo) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count && !(a[e] > lo)) a[e] = lo;
}

__global__ void k_transpose_kp(const double* __restrict__ Ht, uint64_t p, int K, double* __restrict__ H) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p * (uint64_t)K) return;
  const uint64_t j = e / K;
  const int r = (int)(e % K);
  H[(uint64_t)r * p + j] = Ht[e];
}


namespace {
struct FnColSum {
  const double* A; uint64_t lda;
  __device__ void operator()(uint64_t r, int c, double (&v)[1]) const { v[0] = A[r * lda + c]; }
};
struct GStore {
  double* out;
  __device__ void operator()(int c, const double (&v)[1]) const { out[c] = v[0]; }
};
}  // namespace

// Functionals of the mean-filled raw columns of Y (pe rows) against the
// n x m operand Op: out[j][c] = sum_k Op[k][c] y_kj, y = g or mu_j.
void raw_functionals(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const double* mu,
                            const double* Op, uint64_t ldo, uint64_t m, double* out, uint64_t ldout) {
  hipStream_t s = ctx->stream;
  const uint64_t npad = rup(n, 64);
  const int nrb = (int)cdiv(m, kRowBlk);
  const uint64_t lda = (uint64_t)nrb * kRowBlk;
  SLFMM_CHECK(ldout >= lda, SLFMM_E_ARG, "raw_functionals: output stride");
  double* A = ctx->buf<double>("nmf_A", npad * lda);
  pad_operand(s, Op, ldo, n, (int)m, A, lda, npad);
  launch_func_f64(s, Y, ld, pe, npad, A, lda, (int)m, mu, out, ldout, nullptr);
  double* cs = ctx->buf<double>("nmf_cs", m);
  colreduce<1, kColSum>(ctx, FnColSum{Op, ldo}, GStore{cs}, n, (int)m);
  k_add_mean_term<<<(unsigned)cdiv(pe * m, 256), 256, 0, s>>>(out, ldout, pe, (int)m, mu, cs);
  SLFMM_HIP(hipGetLastError());
}

// The same functionals on the int8 matrix cores (m <= 32): the centred
// sketch pass (rsvd_project_i8, centring only) plus mu_j * colsum.
void raw_functionals_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const double* mu,
                        const double* Op, uint64_t m, double* out, uint64_t ldout) {
  rsvd_project_i8(ctx, Y, ld, n, pe, Op, m, SLFMM_NORM_CENTER_ONLY, out, ldout);
  double* cs = ctx->buf<double>("nmf_cs", m);
  colreduce<1, kColSum>(ctx, FnColSum{Op, m}, GStore{cs}, n, (int)m);
  k_add_mean_term<<<(unsigned)cdiv(pe * m, 256), 256, 0, ctx->stream>>>(out, ldout, pe, (int)m, mu, cs);
  SLFMM_HIP(hipGetLastError());
}

static void normalize_w(slfmm_ctx* ctx, double* W, uint64_t n, double* Ht, uint64_t pe, int K, double eps) {
  double* cs = ctx->buf<double>("nmf_wsum", K);
  colreduce<1, kColSum>(ctx, FnColSum{W, (uint64_t)K}, GStore{cs}, n, K);
  k_nmf_scale<<<(unsigned)std::min<uint64_t>(cdiv((n + pe) * K, 256), 8192), 256, 0, ctx->stream>>>(W, n, Ht, pe, K,
                                                                                                    cs, eps);
  SLFMM_HIP(hipGetLastError());
}

// NMF estimation on the rows Y (pe SNPs, row stride ld, kSlack after the
// last row).  W (n x K), Ht (pe x K) and cv (3 x n_iter: nmf MAE, mean MAE,
// count per iteration) are device outputs; omega (n x (K + 10)) may be
// injected, else it is drawn from seed.
void nmf_run(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const slfmm_nmf_config& cfg,
             const double* d_omega, double* W, double* Ht, double* cv, bool i8) {
  hipStream_t s = ctx->stream;
  const int K = (int)cfg.k;
  const uint64_t l2 = cfg.k + 10;
  const double eps = cfg.eps;
  SLFMM_CHECK(K >= 1 && K <= kNmfMaxK, SLFMM_E_ARG, "NMF rank k must be 1..16");
  SLFMM_CHECK(pe >= 1 && n >= 1, SLFMM_E_ARG, "empty estimation panel");
  SLFMM_CHECK(cfg.chunk_size >= 1, SLFMM_E_ARG, "chunk_size must be positive");
  int32_t* nmiss = ctx->buf<int32_t>("nmf_nmiss", pe);
  int32_t* sumg = ctx->buf<int32_t>("nmf_sumg", pe);
  int32_t* nhom = ctx->buf<int32_t>("nmf_nhom", pe);
  double* mu = ctx->buf<double>("nmf_mu", pe);
  double* sc = ctx->buf<double>("nmf_sc", pe);
  launch_stats(s, Y, ld, n, pe, nmiss, sumg, nhom);
  launch_mean_scale(s, nmiss, sumg, n, pe, SLFMM_NORM_CENTER_ONLY, mu, sc);

  // -- init_from_random_probe: Z = Y_meanfill^T Omega', V = right singular
  //    vectors of Z, W = |Omega' V_k|, H = max(0, W^+ Y_meanfill)
  double* Om = ctx->buf<double>("nmf_omega", n * l2);
  if (d_omega) SLFMM_HIP(hipMemcpyAsync(Om, d_omega, n * l2 * 8, hipMemcpyDeviceToDevice, s));
  else launch_normal(s, Om, n, l2, l2, cfg.seed);
  const uint64_t ldz = rup(l2, kRowBlk);
  double* Z = ctx->buf<double>("nmf_Z", pe * ldz);
  raw_functionals(ctx, Y, ld, n, pe, mu, Om, l2, l2, Z, ldz);
  double* G = ctx->buf<double>("nmf_G", l2 * l2);
  ts_gram(ctx, Z, ldz, l2, Z, ldz, l2, pe, G);
  double* V = ctx->buf<double>("nmf_V", l2 * l2);
  double* sv = ctx->buf<double>("nmf_sv", l2);
  jacobi_svd_small(ctx, G, l2, V, sv);
  ts_gemm(s, Om, l2, l2, V, l2, K, n, 1.0, 0.0, W, K);
  k_abs<<<(unsigned)cdiv(n * K, 256), 256, 0, s>>>(W, n * K);
  double* WtW = ctx->buf<double>("nmf_WtW", K * K);
  double* Winv = ctx->buf<double>("nmf_Winv", K * K);
  ts_gram(ctx, W, K, K, W, K, K, n, WtW);
  safe_inv_dev(ctx, WtW, K, Winv,
               ctx->check_slot("Warning: W_init^T W_init is singular, adding epsilon diagonal regularization",
                               SLFMM_E_SINGULAR, "W_init^T W_init inversion failed even with epsilon regularization"));
  double* Wp = ctx->buf<double>("nmf_Wp", n * K);   // (W^+)^T = W (W^T W)^-1
  ts_gemm(s, W, K, K, Winv, K, K, n, 1.0, 0.0, Wp, K);
  const uint64_t ldf = rup(K, kRowBlk);
  double* F = ctx->buf<double>("nmf_F", pe * ldf);
  raw_functionals(ctx, Y, ld, n, pe, mu, Wp, K, K, F, ldf);
  k_relu_rows<<<(unsigned)cdiv(pe * K, 256), 256, 0, s>>>(F, ldf, pe, K, 0.0, Ht);
  k_clamp_min<<<(unsigned)cdiv(n * K, 256), 256, 0, s>>>(W, n * K, eps);
  k_clamp_min<<<(unsigned)cdiv(pe * K, 256), 256, 0, s>>>(Ht, pe * K, eps);
  SLFMM_HIP(hipGetLastError());
  normalize_w(ctx, W, n, Ht, pe, K, eps);

  // -- iterations.  With i8 (K <= kNmfMaxK8) both passes run on the int8
  // matrix cores; the initialisation above stays FP64: its functionals feed
  // an SVD, and int8 sums there moved W by 3e-5 of its largest entry after 5
  // iterations (the passes alone: about 1e-6).
  const bool mc = i8 && K <= kNmfMaxK8;
  const uint64_t ntile = cdiv(n, 256);
  uint64_t nchunk = std::max<uint64_t>(1, std::min<uint64_t>(cdiv(pe, 256), cdiv(2048, ntile)));
  const uint64_t spc = cdiv(pe, nchunk);
  nchunk = cdiv(pe, spc);
  double* part = mc ? nullptr : ctx->buf<double>("nmf_part", nchunk * n * K);
  double* num = ctx->buf<double>("nmf_num", n * K);
  double* hh = ctx->buf<double>("nmf_hh", K * K);
  const unsigned cvblk = (unsigned)cdiv(pe, 256);
  double* cvpart = ctx->buf<double>("nmf_cvpart", (uint64_t)cvblk * 3);
  // CV draws of the nchunk + n_iter - 1 distinct streams, once (when that is
  // fewer draws than n_iter passes and the lists stay within 2 GiB)
  const uint64_t cv_thresh = u01_threshold(cfg.cv_rate);
  const uint64_t cv_nchunk = cdiv(pe, cfg.chunk_size), cv_jl = std::min<uint64_t>(cfg.chunk_size, pe);
  const uint64_t cv_ns = cv_nchunk + cfg.n_iter - 1;
  uint32_t* cv_lists = nullptr;
  uint8_t* cv_counts = nullptr;
  if (cv_thresh > 0 && cfg.n_iter > 1 && cv_ns * cv_jl < cfg.n_iter * pe && cv_ns * cv_jl * (4 * kCvList + 1) < (2ull << 30)) {
    cv_lists = ctx->buf<uint32_t>("nmf_cvlists", cv_ns * cv_jl * kCvList);
    cv_counts = ctx->buf<uint8_t>("nmf_cvcounts", cv_ns * cv_jl);
    slfmm_timed tm(ctx, "nmf_cv_lists");
    k_cv_lists<<<(unsigned)cdiv(cv_ns * cv_jl, 256), 256, 0, s>>>(n, cfg.seed, cv_ns, cv_jl, cv_thresh, cv_lists,
                                                                 cv_counts);
    SLFMM_HIP(hipGetLastError());
  }
  uint8_t* YT = nullptr;
  uint64_t ldt = 0;
  if (mc) {   // sample-major copy for the forward passes
    ldt = rup(cdiv(pe, 4), 16);
    YT = ctx->buf<uint8_t>("nmf_YT", n * ldt + kSlack);
    launch_fill(s, reinterpret_cast<double*>(YT + n * ldt), 0.0, kSlack / 8);   // the transpose writes every row byte
    launch_transpose(s, Y, ld, n, pe, YT, ldt);
  }
  for (uint64_t it = 0; it < cfg.n_iter; ++it) {
    // forward: W update from the current reconstruction
    if (mc) {
      slfmm_timed tm(ctx, "nmf_forward");
      nmf_forward_i8(ctx, YT, ldt, n, pe, Ht, W, K, num);
    } else {
      slfmm_timed tm(ctx, "nmf_forward");
      const dim3 fgrid((unsigned)ntile, (unsigned)nchunk);
      switch (K) {
#define SLFMM_FWD(R) case R: k_nmf_fwd<R><<<fgrid, 256, 0, s>>>(Y, ld, n, pe, spc, W, Ht, part); break;
        SLFMM_NMF_RANKS(SLFMM_FWD)
#undef SLFMM_FWD
      }
      k_nmf_sum_chunks<<<(unsigned)cdiv(n * K, 256), 256, 0, s>>>(part, n * K, (int)nchunk, num);
    }
    ts_gram(ctx, Ht, K, K, Ht, K, K, pe, hh);
    k_nmf_wupd<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(W, num, hh, n, K, eps);
    SLFMM_HIP(hipGetLastError());
    normalize_w(ctx, W, n, Ht, pe, K, eps);
    // backward: H update in place, fill from the new W and the old H
    ts_gram(ctx, W, K, K, W, K, K, n, WtW);
    const unsigned bgrid = (unsigned)cdiv(pe, 256);
    if (mc) {
      slfmm_timed tm(ctx, "nmf_backward");
      nmf_backward_i8(ctx, Y, ld, n, pe, W, K, WtW, eps, Ht);
    } else {
      slfmm_timed tm(ctx, "nmf_backward");
      switch (K) {
#define SLFMM_BWD(R) case R: k_nmf_bwd<R><<<bgrid, 256, 0, s>>>(Y, ld, n, pe, W, WtW, eps, Ht); break;
        SLFMM_NMF_RANKS(SLFMM_BWD)
#undef SLFMM_BWD
      }
    }
    SLFMM_HIP(hipGetLastError());
    normalize_w(ctx, W, n, Ht, pe, K, eps);
    // held-out error of this iteration
    slfmm_timed tm(ctx, "nmf_cv");
    k_nmf_cv<<<cvblk, 256, 0, s>>>(Y, ld, n, pe, W, Ht, K, nmiss, sumg, cfg.seed + it, cfg.chunk_size, cv_thresh,
                                   cv_lists, cv_counts, cv_jl, it, cvpart);
    k_nmf_cv_final<<<1, 256, 0, s>>>(cvpart, (int)cvblk, cv + 3 * it);
    SLFMM_HIP(hipGetLastError());
  }
}

void relu_rows(hipStream_t s, const double* F, uint64_t ldf, uint64_t p, int K, double* Ht) {
  if (!p) return;
  k_relu_rows<<<(unsigned)cdiv(p * K, 256), 256, 0, s>>>(F, ldf, p, K, 0.0, Ht);
  SLFMM_HIP(hipGetLastError());
}

// ---- GWAS-phase CV of the association pass (src/testing.rs:423-487) ---------
// Chunk q (chunk_size consecutive SNPs) draws one uniform per OBSERVED call
// from the stream seeded seed + q, in column-then-row order; a draw below
// cv_rate holds the call out.  Per SNP the held-out column is mean-filled
// (leave-out mean mu_m), H = max(0, W^+ y_m) and the held-out calls are
// predicted by W H.  W^+ y_m follows from the precomputed W^+ y_mf (mean
// fill with the full mean mu) by the corrections on the changed entries:
//   held-out k: (mu_m - g_k) Wp_k;   missing k: (mu_m - mu) Wp_k (summed: SWp).
__global__ void k_chunk_obs_base(const int32_t* __restrict__ nmiss, uint64_t n, uint64_t p, uint64_t chunk,
                                 uint64_t* __restrict__ base) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t j0 = q * chunk;
  if (j0 >= p) return;
  const uint64_t j1 = std::min(p, j0 + chunk);
  uint64_t b = 0;
  for (uint64_t j = j0; j < j1; ++j) {
    base[j] = b;
    b += n - (uint64_t)nmiss[j];
  }
}

// One thread per SNP.  The first pass draws every observed call and records
// the sample indices of the first kGwasList held-out calls in LDS; when a
// SNP holds out no more than that, the other two passes walk the list
// instead of drawing again (same calls, same order).
constexpr int kGwasList = 16;
__global__ void __launch_bounds__(256) k_gwas_cv(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t n,
                                                 uint64_t p, const double* __restrict__ W,
                                                 const double* __restrict__ Wp, int K,
                                                 const double* __restrict__ Hraw, uint64_t ldh,
                                                 const double* __restrict__ SWp, const int32_t* __restrict__ nmiss,
                                                 const int32_t* __restrict__ sumg, const uint64_t* __restrict__ base,
                                                 uint64_t seed, uint64_t chunk, uint64_t cv_thresh,
                                                 double* __restrict__ part) {
  __shared__ double r0[256], r1[256];
  __shared__ unsigned long long r2[256];
  __shared__ uint32_t held[kGwasList][256];
  const int t = threadIdx.x;
  const uint64_t j = (uint64_t)blockIdx.x * 256 + t;
  double e1 = 0.0, e2 = 0.0;
  unsigned long long cnt = 0;
  if (j < p) {
    // draw rank of the stream seeded seed + j / chunk: base[j] + (observed calls before)
    const uint64_t ctr0 = stream_counter(seed + j / chunk, base[j]);
    const uint8_t* rowb = packed + j * ld;
    const uint4* row = reinterpret_cast<const uint4*>(rowb);
    const uint64_t nkb = cdiv(n, 64);
    // fn(k, code) for every held-out call in sample order, by drawing again
    auto redraw = [&](auto&& fn) {
      uint64_t ctr = ctr0;
      for (uint64_t kb = 0; kb < nkb; ++kb) {
        const uint4 q = row[kb];
        const uint32_t wd[4] = {q.x, q.y, q.z, q.w};
        const int lim = (int)std::min<uint64_t>(64, n - kb * 64);
        for (int kk = 0; kk < lim; ++kk) {
          const uint32_t code = code_of(wd[kk >> 4], kk & 15);
          if (code == 1u) continue;
          if ((mix_counter(ctr) >> 11) < cv_thresh) fn(kb * 64 + kk, code);
          ctr += kStreamStep;
        }
      }
    };
    // pass A: held-out count and dosage sum, the first kGwasList indices kept
    uint64_t m = 0;
    double sgm = 0.0;
    redraw([&](uint64_t k, uint32_t code) {
      if (m < kGwasList) held[m][t] = (uint32_t)k;
      ++m;
      sgm += dosage_of(code);
    });
    auto each_held = [&](auto&& fn) {
      if (m <= kGwasList) {
        for (uint64_t i = 0; i < m; ++i) {
          const uint32_t k = held[i][t];
          fn((uint64_t)k, (uint32_t)(rowb[k >> 2] >> (2 * (k & 3))) & 3u);
        }
      } else {
        redraw(fn);
      }
    };
    if (m > 0) {
      const int64_t nobs = (int64_t)n - nmiss[j];
      const double mu = nobs > 0 ? (double)sumg[j] / (double)nobs : 0.0;
      const int64_t keep = nobs - (int64_t)m;
      const double mum = keep > 0 ? ((double)sumg[j] - sgm) / (double)keep : 0.0;
      double h[kNmfMaxK];
      for (int r = 0; r < K; ++r) h[r] = Hraw[j * ldh + r] + (mum - mu) * SWp[j * K + r];
      // pass B: corrections of the held-out calls
      each_held([&](uint64_t k, uint32_t code) {
        const double dlt = mum - dosage_of(code);
        for (int r = 0; r < K; ++r) h[r] += Wp[k * K + r] * dlt;
      });
      for (int r = 0; r < K; ++r) h[r] = h[r] > 0.0 ? h[r] : 0.0;
      // pass C: errors of the held-out calls
      each_held([&](uint64_t k, uint32_t code) {
        double pred = 0.0;
        for (int r = 0; r < K; ++r) pred += W[k * K + r] * h[r];
        const double g = dosage_of(code);
        e1 += fabs(g - pred);
        e2 += fabs(g - mum);
        ++cnt;
      });
    }
  }
  r0[t] = e1; r1[t] = e2; r2[t] = cnt;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (t < s2) { r0[t] += r0[t + s2]; r1[t] += r1[t + s2]; r2[t] += r2[t + s2]; }
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x * 3 + 0] = r0[0];
    part[blockIdx.x * 3 + 1] = r1[0];
    part[blockIdx.x * 3 + 2] = (double)r2[0];
  }
}

void gwas_cv(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p, const double* W,
             const double* Wp, int K, const double* Hraw, uint64_t ldh, const double* SWp, const int32_t* nmiss,
             const int32_t* sumg, uint64_t seed, uint64_t chunk, double cv_rate, double* out3) {
  hipStream_t s = ctx->stream;
  SLFMM_CHECK(chunk >= 1, SLFMM_E_ARG, "chunk_size must be positive");
  if (!p) {
    launch_fill(s, out3, 0.0, 3);
    return;
  }
  uint64_t* base = ctx->buf<uint64_t>("gcv_base", p);
  const uint64_t nq = cdiv(p, chunk);
  k_chunk_obs_base<<<(unsigned)cdiv(nq, 64), 64, 0, s>>>(nmiss, n, p, chunk, base);
  const unsigned nblk = (unsigned)cdiv(p, 256);
  double* part = ctx->buf<double>("gcv_part", (uint64_t)nblk * 3);
  k_gwas_cv<<<nblk, 256, 0, s>>>(packed, ld, n, p, W, Wp, K, Hraw, ldh, SWp, nmiss, sumg, base, seed, chunk,
                                 u01_threshold(cv_rate), part);
  k_nmf_cv_final<<<1, 256, 0, s>>>(part, (int)nblk, out3);
  SLFMM_HIP(hipGetLastError());
}

void nmf_transpose_h(hipStream_t s, const double* Ht, uint64_t pe, int K, double* H) {
  k_transpose_kp<<<(unsigned)cdiv(pe * K, 256), 256, 0, s>>>(Ht, pe, K, H);
  SLFMM_HIP(hipGetLastError());
}

}  // namespace slfmm
