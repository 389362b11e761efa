// Steps 0-2: thin precompute and the streaming randomized SVD on the GPU.
//
// The reference forms M = D_lambda Q^T (n x n, src/precompute.rs:66-69) and
// multiplies it into every forward-pass chunk (src/rsvd.rs:93-94).  With
// N = Q M = I - Q_d diag(1 - d_lambda) Q_d^T (identity plus rank d, Q_d the
// thin left singular vectors of X) every step has an exact low-rank form:
//   sketch     Z = Y^T B0                 (B0 = M^T Omega, src/rsvd.rs:51)
//   forward    W = Y Q_z;  Q_W = qr(N W)  (reference: qr(M Y Q_z) = Q^T qr(N W))
//   backward   Z = Y^T (N Q_W)            (reference: Y^T M^T Q_a)
//   recovery   U_hat = N^{-1} U_{NW}[:, :K] (reference: Q D^{-1} U_small[:, :K])
// so U_hat equals the reference's up to column signs, with no n x n array.
#include "internal.h"
#include "rng.h"

namespace slfmm {

// ---- forward pass: W[k][c] = sum_j v_jk * sc_j * Qz[j][c] --------------------
// One workgroup per (512-sample tile, SNP chunk, 16-column block); thread t
// takes samples 512 x + t and 512 x + 256 + t.  The SNP-side values (16
// columns of a Q_z row, mu_j, sc_j) are the same for every thread, so they
// come in by scalar loads and feed the FMAs from SGPRs; the codes of 32 SNPs
// x 512 samples (and, with FILL, their loadings H_j) are staged in LDS.
// Partial sums per chunk are reduced in a fixed order afterwards.  With FILL
// the missing calls take W_k . H_j - mu_j (NMF in-RAM imputation).  Rows are
// read 128 bytes per tile, within the 128 readable bytes past the last row.
// Measured at n = 5000, p = 2M, l = 18 with an NMF fill: 42.5 ms, against
// 53 ms with the Q_z rows broadcast from LDS and 62 ms with one 32-column
// block (no prefetch room in the SGPRs); 16 columns keep two SNPs' rows in
// flight.
template <int FK>
__global__ void __launch_bounds__(256) k_acc_f64(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t n,
                                                 uint64_t p, uint64_t snps_per_chunk, const double* __restrict__ Qz,
                                                 uint64_t ldq, int l, const double* __restrict__ mu,
                                                 const double* __restrict__ sc, double* __restrict__ part,
                                                 uint64_t ldp, FillSpec fs) {
  constexpr int NC = 16;
  __shared__ uint32_t sw[32][32];
  constexpr bool FILL = FK > 0;
  constexpr int KR = FILL ? FK : 1;   // fill rank (fs.K)
  __shared__ double hs[FILL ? 32 : 1][KR];
  const int t = threadIdx.x;
  const uint64_t k0 = (uint64_t)blockIdx.x * 512 + t, k1 = k0 + 256;
  const uint64_t j0 = (uint64_t)blockIdx.y * snps_per_chunk;
  const uint64_t j1 = std::min(p, j0 + snps_per_chunk);
  const int cb = blockIdx.z;
  const int ncol = min(NC, l - cb * NC);
  double wk0[KR], wk1[KR];
  if constexpr (FILL) {
#pragma unroll
    for (int r = 0; r < KR; ++r) {
      wk0[r] = k0 < n ? fs.W[k0 * KR + r] : 0.0;
      wk1[r] = k1 < n ? fs.W[k1 * KR + r] : 0.0;
    }
  }
  double acc0[NC], acc1[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) { acc0[c] = 0.0; acc1[c] = 0.0; }
  const int sh = 2 * (t & 15), w0 = t >> 4;
  for (uint64_t jb = j0; jb < j1; jb += 32) {
    const int nj = (int)std::min<uint64_t>(32, j1 - jb);
    __syncthreads();
    for (int e = t; e < 32 * 32; e += 256) {
      const int jj = e >> 5, wd = e & 31;
      sw[jj][wd] = jj < nj ? reinterpret_cast<const uint32_t*>(packed + (jb + jj) * ld + (uint64_t)blockIdx.x * 128)[wd]
                           : 0u;
    }
    if constexpr (FILL) {
      for (int e = t; e < 32 * KR; e += 256) {
        const int jj = e / KR, r = e % KR;
        hs[jj][r] = jj < nj ? fs.Ht[(jb + jj) * KR + r] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll 2   // two SNPs' Q_z rows (64 SGPRs) in flight
    for (int jj = 0; jj < nj; ++jj) {
      const uint64_t j = jb + jj;
      const uint32_t c0 = (sw[jj][w0] >> sh) & 3u, c1 = (sw[jj][16 + w0] >> sh) & 3u;
      const double m = mu[j], sj = sc[j];
      double miss0 = 0.0, miss1 = 0.0;
      if constexpr (FILL) {
        if (c0 == 1u || c1 == 1u) {
          double f0 = 0.0, f1 = 0.0;
#pragma unroll
          for (int r = 0; r < KR; ++r) {
            const double h = hs[jj][r];
            f0 += wk0[r] * h;
            f1 += wk1[r] * h;
          }
          miss0 = f0 - m;
          miss1 = f1 - m;
        }
      }
      const double v0 = ((c0 & 2u) ? ((c0 & 1u) ? 2.0 - m : 1.0 - m) : ((c0 & 1u) ? miss0 : -m)) * sj;
      const double v1 = ((c1 & 2u) ? ((c1 & 1u) ? 2.0 - m : 1.0 - m) : ((c1 & 1u) ? miss1 : -m)) * sj;
      const double* q = Qz + j * ldq + cb * NC;   // ldq >= NC (cb + 1): padded columns are read, never stored
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        acc0[c] += v0 * q[c];
        acc1[c] += v1 * q[c];
      }
    }
  }
  if (k0 < n)
    for (int c = 0; c < ncol; ++c) part[((uint64_t)blockIdx.y * n + k0) * ldp + cb * NC + c] = acc0[c];
  if (k1 < n)
    for (int c = 0; c < ncol; ++c) part[((uint64_t)blockIdx.y * n + k1) * ldp + cb * NC + c] = acc1[c];
}

__global__ void k_acc_reduce(const double* __restrict__ part, uint64_t n, int l, uint64_t ldp, int nchunk,
                             double* W, uint64_t ldw) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * (uint64_t)l) return;
  uint64_t k = e / l;
  int c = (int)(e % l);
  double s = 0.0;
  for (int g = 0; g < nchunk; ++g) s += part[((uint64_t)g * n + k) * ldp + c];
  W[k * ldw + c] = s;
}

static void accumulate_f64(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                           const double* Qz, uint64_t ldq, uint64_t l, const double* mu, const double* sc,
                           double* W, uint64_t ldw, const FillSpec* fill) {
  hipStream_t s = ctx->stream;
  const bool fl = fill && fill->W;
  constexpr int NC = 16;   // columns per block
  const uint64_t ntile = cdiv(n, 512), ncb = cdiv(l, NC);
  SLFMM_CHECK(ldq >= NC * ncb, SLFMM_E_ARG, "forward pass: Q_z row stride below its column blocks");
  uint64_t nchunk = std::max<uint64_t>(1, std::min<uint64_t>(cdiv(p, 256), cdiv(2048, ntile * ncb)));
  const uint64_t spc = cdiv(p, nchunk);
  nchunk = cdiv(p, spc);
  const uint64_t ldp = ncb * NC;
  double* part = ctx->buf<double>("rs_accpart", nchunk * n * ldp);
  const dim3 grid((unsigned)ntile, (unsigned)nchunk, (unsigned)ncb);
  if (fl) {
    SLFMM_CHECK(fill->K >= 1 && fill->K <= kNmfMaxK, SLFMM_E_ARG, "NMF rank must be 1..16");
    switch (fill->K) {
#define SLFMM_ACC(R) \
  case R: k_acc_f64<R><<<grid, 256, 0, s>>>(packed, ld, n, p, spc, Qz, ldq, (int)l, mu, sc, part, ldp, *fill); break;
      SLFMM_NMF_RANKS(SLFMM_ACC)
#undef SLFMM_ACC
    }
  } else {
    k_acc_f64<0><<<grid, 256, 0, s>>>(packed, ld, n, p, spc, Qz, ldq, (int)l, mu, sc, part, ldp, FillSpec{});
  }
  k_acc_reduce<<<(unsigned)cdiv(n * l, 256), 256, 0, s>>>(part, n, (int)l, ldp, (int)nchunk, W, ldw);
  SLFMM_HIP(hipGetLastError());
}

// ---- low-rank operator N (and N^{-1}) ----------------------------------------
// V (rows x c, ldv) <- N V  or  N^{-1} V:  V + Q_d S (Q_d^T V) with
// S = diag(-(1 - d_lambda)) (N) or diag(1 / d_lambda - 1) (N^{-1}),
// the scaling applied while the GEMM stages Q_d^T V.
static void apply_N(slfmm_ctx* ctx, const double* Qd, uint64_t d, const double* dlam, double* V, uint64_t ldv,
                    uint64_t c, uint64_t rows, bool inverse) {
  double* C = ctx->buf<double>("rs_NC", d * c);
  ts_gram(ctx, Qd, d, d, V, ldv, c, rows, C);
  ts_gemm(ctx->stream, Qd, d, d, C, c, c, rows, 1.0, 1.0, V, ldv, nullptr, dlam, inverse);
}

__global__ void k_normal(double* out, uint64_t n, int l, uint64_t ld, uint64_t seed) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * (uint64_t)l) return;
  out[(e / l) * ld + (e % l)] = normal01(seed, e);
}

void launch_normal(hipStream_t s, double* out, uint64_t n, uint64_t l, uint64_t ld, uint64_t seed) {
  if (!n || !l) return;
  k_normal<<<(unsigned)cdiv(n * l, 256), 256, 0, s>>>(out, n, (int)l, ld, seed);
  SLFMM_HIP(hipGetLastError());
}

__global__ void k_pad_rows(const double* __restrict__ src, uint64_t lds, uint64_t rows, int cols,
                           double* dst, uint64_t ldd, uint64_t rows_pad) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows_pad * ldd) return;
  uint64_t r = e / ldd;
  int c = (int)(e % ldd);
  dst[e] = (r < rows && c < cols) ? src[r * lds + c] : 0.0;
}

__global__ void k_transpose_small(const double* in, int m, double* out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m * m) out[(e % m) * m + e / m] = in[e];
}

// ---- thin precompute -----------------------------------------------------------
// X = Q_d diag(sigma) R^T from the eigendecomposition of X^T X (one-sided
// Jacobi on the d x d Gram), d_lambda = sqrt(lambda / (lambda + sigma^2))
// (src/precompute.rs:52-59), ridge_inv = safe_inv(X^T X + lambda I).
__global__ void k_prec_finish(const double* ev, const double* V, int d, double lambda, double* sigma,
                              double* dlam, double* M, double* Gr) {
  if (threadIdx.x != 0) return;
  double smax = ev[0] > 0.0 ? sqrt(ev[0]) : 0.0;
  for (int j = 0; j < d; ++j) {
    double sg = ev[j] > 0.0 ? sqrt(ev[j]) : 0.0;
    sigma[j] = sg;
    double s2 = sg * sg;
    dlam[j] = sqrt(lambda / (lambda + s2));
    double inv = (sg > 1e-12 * smax && sg > 0.0) ? 1.0 / sg : 0.0;
    for (int i = 0; i < d; ++i) M[i * d + j] = V[i * d + j] * inv;
  }
  for (int i = 0; i < d; ++i) Gr[i * d + i] += lambda;
}

void precompute_dev(slfmm_ctx* ctx, const double* d_xs, uint64_t n, uint64_t d, double lambda,
                    double* d_sigma, double* d_dlam, double* d_qd, double* d_ridge) {
  SLFMM_CHECK(lambda > 0.0, SLFMM_E_LAMBDA,
              "Ridge penalty λ must be positive, got " + std::to_string(lambda) +
                  ". λ=0 makes the D_λ diagonal singular (d_λ[j] = 0/(0+σ²) = 0, d_λ_inv = ∞). "
                  "Use a small positive value like 1e-10.");
  SLFMM_CHECK(d >= 1 && d <= (uint64_t)kMaxD, SLFMM_E_ARG, "need 1 <= d <= 16 covariates");
  hipStream_t s = ctx->stream;
  double* G = ctx->buf<double>("pc_G", d * d);
  ts_gram(ctx, d_xs, d, d, d_xs, d, d, n, G);
  double* V = ctx->buf<double>("pc_V", d * d);
  double* ev = ctx->buf<double>("pc_ev", d);
  jacobi_svd_small(ctx, G, d, V, ev);
  double* M = ctx->buf<double>("pc_M", d * d);
  k_prec_finish<<<1, 64, 0, s>>>(ev, V, (int)d, lambda, d_sigma, d_dlam, M, G);
  ts_gemm(s, d_xs, d, d, M, d, d, n, 1.0, 0.0, d_qd, d);
  safe_inv_dev(ctx, G, d, d_ridge,
               ctx->check_slot("Warning: X^T X + λI (ridge) is singular, adding epsilon diagonal regularization",
                               SLFMM_E_SINGULAR, "X^T X + λI (ridge) inversion failed even with epsilon regularization"));
}

// Rows h_idx[0..n_idx) of g as a contiguous packed block (kSlack zero bytes
// after the last row); g's own buffer when the subset is every SNP in order.
const uint8_t* gather_subset(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* h_idx, uint64_t n_idx,
                             const char* tag, uint64_t* pe) {
  hipStream_t s = ctx->stream;
  bool identity = (h_idx == nullptr);
  if (h_idx) {
    identity = (n_idx == g->p);
    for (uint64_t i = 0; identity && i < n_idx; ++i) identity = (h_idx[i] == (int64_t)i);
  }
  if (identity && !g->host) {
    *pe = g->p;
    return g->d;
  }
  if (!identity)
    for (uint64_t i = 0; i < n_idx; ++i)
      SLFMM_CHECK(h_idx[i] >= 0 && (uint64_t)h_idx[i] < g->p, SLFMM_E_ARG, "estimation SNP index out of range");
  if (g->host) {   // rows of a host-resident panel: copied through pinned staging
    const uint64_t m = identity ? g->p : n_idx;
    uint8_t* Yg = ctx->buf<uint8_t>(tag, m * g->ld + kSlack);
    SLFMM_HIP(hipMemsetAsync(Yg, 0, m * g->ld + kSlack, s));
    host_rows_to_device(ctx, g, identity ? nullptr : h_idx, m, Yg);
    *pe = m;
    return Yg;
  }
  int64_t* didx = ctx->buf<int64_t>(std::string(tag) + "_idx", n_idx);
  if (ctx->staged_idx != h_idx) ctx->upload(didx, h_idx, n_idx * 8);
  uint8_t* Yg = ctx->buf<uint8_t>(tag, n_idx * g->ld + kSlack);
  launch_fill(s, reinterpret_cast<double*>(Yg + n_idx * g->ld), 0.0, kSlack / 8);
  geno_gather(ctx, g, didx, n_idx, Yg);
  *pe = n_idx;
  return Yg;
}

// ---- randomized SVD ---------------------------------------------------------------
void rsvd_run(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* h_idx, uint64_t n_idx, const double* d_xs,
              uint64_t d, const slfmm_config& cfg, const double* d_b0, double* d_uhat, const FillSpec* fill) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n, K = cfg.k, l = cfg.k + cfg.oversampling;
  SLFMM_CHECK(K >= 1 && l <= 64, SLFMM_E_ARG, "need 1 <= K and K + oversampling <= 64");
  SLFMM_CHECK(l <= n, SLFMM_E_ARG, "K + oversampling must not exceed the number of samples");
  double* sigma = ctx->buf<double>("rs_sigma", d);
  double* dlam = ctx->buf<double>("rs_dlam", d);
  double* Qd = ctx->buf<double>("rs_Qd", n * d);
  double* ridge = ctx->buf<double>("rs_ridge", d * d);
  precompute_dev(ctx, d_xs, n, d, cfg.lambda, sigma, dlam, Qd, ridge);

  // estimation SNPs as a contiguous packed block
  uint64_t pe = 0;
  const uint8_t* Y = gather_subset(ctx, g, h_idx, n_idx, "rs_Y", &pe);
  SLFMM_CHECK(pe >= l, SLFMM_E_ARG, "need at least K + oversampling estimation SNPs");
  int32_t* nmiss = ctx->buf<int32_t>("rs_nmiss", pe);
  int32_t* sumg = ctx->buf<int32_t>("rs_sumg", pe);
  int32_t* nhom = ctx->buf<int32_t>("rs_nhom", pe);
  double* mu = ctx->buf<double>("rs_mu", pe);
  double* sc = ctx->buf<double>("rs_sc", pe);
  const bool filled = fill && fill->W;
  if (filled) {   // NMF in-RAM fill of the estimation rows: moments of the filled rows
    fill_moments(ctx, Y, g->ld, pe, *fill, cfg.norm, nullptr, mu, sc, nullptr, nullptr,
                 cfg.precision == SLFMM_PREC_I8);
  } else {
    launch_stats(s, Y, g->ld, n, pe, nmiss, sumg, nhom);
    launch_mean_scale(s, nmiss, sumg, n, pe, cfg.norm, mu, sc);
  }

  // the int8 sketch carries at most 32 columns; wider sketches run the FP64
  // kernels.  With an NMF fill the sketch/backward and forward passes take
  // the filled int8 passes where their plans reach (fill_i8, fill_fwd_i8).
  const bool mixed = cfg.precision == SLFMM_PREC_I8 && l <= 32 && !filled;
  const bool fill_i8 = cfg.precision == SLFMM_PREC_I8 && filled && fill_functionals_i8_supported((int)l, fill->K);
  const bool fill_fwd_i8 = cfg.precision == SLFMM_PREC_I8 && filled && fill_forward_i8_supported((int)l, fill->K);
  if (cfg.precision == SLFMM_PREC_I8 && !filled && !mixed)
    fp64_stage(ctx, "randomized SVD", "K + oversampling = " + std::to_string(l) + " > 32 columns");
  if (cfg.precision == SLFMM_PREC_I8 && filled && !(fill_i8 && fill_fwd_i8))
    fp64_stage(ctx, "randomized SVD over the NMF-filled rows",
               "K + oversampling = " + std::to_string(l) + " with nmf-k = " + std::to_string(fill->K) +
                   " is past the int8 fill plans");
  uint8_t* YT = nullptr;
  uint64_t ldt = 0;
  // The sample-major copy for the forward passes is first read by the first
  // forward pass, after the first sketch and its QR: for a large copy (at
  // least 64 MB) in a single fit outside allocation-free calls it is made on
  // the side stream, overlapping that chain, and joined back before the first
  // forward pass (or when this scope is left early, so that a graph capture
  // always sees the join).  Measured: 0.15 ms less per fit at n = 10 000 and
  // 100 000 estimation SNPs; in a batch of small fits, whose streams already
  // overlap, the fork and join cost more than they hide (40.7 -> 60.3 ms).
  struct SideJoin {
    slfmm_ctx* ctx = nullptr;
    void join() {
      if (ctx) SLFMM_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
      ctx = nullptr;
    }
    ~SideJoin() {
      if (ctx) (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
    }
  } yt_join;
  if (mixed || fill_fwd_i8) {
    ldt = rup(cdiv(pe, 4), 16);
    YT = ctx->buf<uint8_t>("rs_YT", n * ldt + kSlack);
    SLFMM_CHECK(ldt % 16 == 0 && g->ld % 16 == 0, SLFMM_E_ARG, "transpose strides must be multiples of 16 bytes");
    hipStream_t ts = s;
    if (!ctx->arena && !ctx->batch_worker && n * ldt >= (64ull << 20)) {
      ctx->ensure_side_stream();
      ts = ctx->side_stream;
      SLFMM_HIP(hipEventRecord(ctx->ev_fork, s));
      SLFMM_HIP(hipStreamWaitEvent(ts, ctx->ev_fork, 0));
      yt_join.ctx = ctx;
    }
    launch_fill(ts, reinterpret_cast<double*>(YT + n * ldt), 0.0, kSlack / 8);   // the transpose writes every row byte
    launch_transpose(ts, Y, g->ld, n, pe, YT, ldt);
    if (ts != s) SLFMM_HIP(hipEventRecord(ctx->ev_join, ts));
  }

  const uint64_t npad = rup(n, 64);
  const int nrb = (int)cdiv(l, kRowBlk);
  const uint64_t lda = (uint64_t)nrb * kRowBlk;
  double* B = ctx->buf<double>("rs_B", n * l);
  double* A = ctx->buf<double>("rs_A", npad * lda);
  double* Z = ctx->buf<double>("rs_Z", pe * lda);
  double* W = ctx->buf<double>("rs_W", n * l);
  if (d_b0) {
    copy2d(s, d_b0, l, B, l, n, l);
  } else {
    launch_normal(s, B, n, l, l, cfg.seed);
    apply_N(ctx, Qd, d, dlam, B, l, l, n, false);
  }
  auto project = [&]() {   // Z = Y^T B, scaled per SNP
    slfmm_timed tm(ctx, "rsvd_project");
    if (mixed) {
      rsvd_project_i8(ctx, Y, g->ld, n, pe, B, l, cfg.norm, Z, lda);
    } else if (fill_i8) {
      fill_functionals_i8(ctx, Y, g->ld, n, pe, B, l, (int)l, fill->W, fill->Ht, fill->K, mu, sc, Z, lda);
    } else {
      k_pad_rows<<<(unsigned)cdiv(npad * lda, 256), 256, 0, s>>>(B, l, n, (int)l, A, lda, npad);
      launch_func_f64(s, Y, g->ld, pe, npad, A, lda, (int)l, mu, Z, lda, sc, fill);
    }
  };
  // The bases between power iterations carry only their column span forward,
  // so one CholeskyQR pass orthonormalises them; the basis of the final
  // projection, whose singular vectors are the result, gets CholeskyQR2.
  double* Zq = ctx->buf<double>("rs_Zq", pe * lda);
  auto orth_z = [&]() {   // Zq = orthonormal basis of span(Z)
    cholqr1(ctx, Z, lda, l, pe, Zq, lda,
            ctx->check_slot("", SLFMM_E_NUMERIC, "QR of the RSVD sketch failed (rank deficient)"));
  };
  auto forward = [&]() {   // W = N Y Q_z
    yt_join.join();
    {
      slfmm_timed tm(ctx, "rsvd_forward");
      if (mixed) rsvd_accumulate_i8(ctx, YT, ldt, n, pe, Zq, lda, l, mu, sc, W, l);
      else if (fill_fwd_i8)
        fill_forward_i8(ctx, YT, ldt, n, pe, Zq, lda, (int)l, mu, sc, fill->W, fill->Ht, fill->K, W, l);
      else accumulate_f64(ctx, Y, g->ld, n, pe, Zq, lda, l, mu, sc, W, l, fill);
    }
    apply_N(ctx, Qd, d, dlam, W, l, l, n, false);
  };
  project();                                           // src/rsvd.rs:56-75
  orth_z();
  for (uint64_t it = 0; it < cfg.n_power_iter; ++it) {  // src/rsvd.rs:79-133
    forward();
    cholqr1(ctx, W, l, l, n, B, l, ctx->check_slot("", SLFMM_E_NUMERIC, "QR of the power-iteration block failed"));
    apply_N(ctx, Qd, d, dlam, B, l, l, n, false);
    project();
    orth_z();
  }
  forward();                                           // src/rsvd.rs:138-157
  double* R = ctx->buf<double>("rs_R", l * l);
  double* Rt = ctx->buf<double>("rs_Rt", l * l);
  double* Vs = ctx->buf<double>("rs_V", l * l);
  double* sv = ctx->buf<double>("rs_sv", l);
  // every pass may shift: W is rank deficient when l = n (its columns are
  // centred) or when the estimation panel has fewer than l polymorphic SNPs,
  // and only its top K directions are kept (src/rsvd.rs:160-173)
  cholqr2(ctx, W, l, l, n, R, ctx->check_slot("", SLFMM_E_NUMERIC, "QR of the final projection failed"), true);
  k_transpose_small<<<(unsigned)cdiv(l * l, 256), 256, 0, s>>>(R, (int)l, Rt);
  jacobi_svd_small(ctx, Rt, l, Vs, sv);                // left singular vectors of R
  ts_gemm(s, W, l, l, Vs, l, K, n, 1.0, 0.0, d_uhat, K);
  apply_N(ctx, Qd, d, dlam, d_uhat, K, K, n, true);    // src/rsvd.rs:168-172
}

}  // namespace slfmm
