// Association pass (Steps 3-4), GIF and calibration on the GPU.
//
// The reference applies two dense n x n operators per chunk
// (src/testing.rs:360, :374).  Here every per-SNP output is a function of
// m = K + d linear functionals f = Q^T y of the genotype column, where Q is
// an orthonormal basis of the centred design [U_c | X] (CholeskyQR2 on the
// GPU, [U_c|X] = Q T), plus integer genotype counts (sum, sum of squares,
// missing).  With s_full = C^T y = E f for C = [1|X|U] (1^T y = 0 because
// y is centred):
//   coef_X = Ginv[1..d, :] E f         t_j = coef_j / sqrt(RSS/df * Ginv_jj)
//   RSS    = TSS - |f|^2                (TSS exact from integer counts)
//   B      = ridge_inv (E_X - X^T U (U^T U)^{-1} E_U) f
//   SS_cov = |L_x^T E_X f|^2,  (X^T X)^{-1} = L_x L_x^T
// which equal src/testing.rs:359-404 in exact arithmetic.  If [U_c|X] is
// rank deficient the basis falls back to [U_c|X] itself (T = I) and RSS
// uses the quadratic form of the (safe_inv-regularised) Ginv, as the
// reference does.
#include "internal.h"
#include "epilogue.h"

namespace slfmm {


// ---------------------------------------------------------------------------
// FP64 functional kernel:  F[j][32*rb + r] = sum_k A[k][32*rb + r] * v_jk
// with v = g - mu_j on observed calls and 0 on missing calls (mean
// imputation after centring, src/bed.rs:339-343).  Tile: 256 SNPs x 32 rows;
// wave w -> rows 16*(w&1)..+16 of the row block, SNP group w>>1; lane ->
// 2 SNPs.  A k-block of 64 samples of A is staged in LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_func_f64(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t p,
                                                  uint64_t npad, const double* __restrict__ A, uint64_t lda,
                                                  const double* __restrict__ mu, double* __restrict__ F,
                                                  uint64_t ldF, const double* __restrict__ outscale) {
  __shared__ double As[64][33];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int half = w & 1, grp = w >> 1;
  const int rb = blockIdx.y;
  const uint64_t j0 = (uint64_t)blockIdx.x * 256 + grp * 128 + lane;
  const uint64_t j1 = j0 + 64;
  const bool ok0 = j0 < p, ok1 = j1 < p;
  const double m0 = ok0 ? mu[j0] : 0.0, m1 = ok1 ? mu[j1] : 0.0;
  const double l00 = -m0, l02 = 1.0 - m0, l03 = 2.0 - m0;
  const double l10 = -m1, l12 = 1.0 - m1, l13 = 2.0 - m1;
  double acc0[16], acc1[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) { acc0[r] = 0.0; acc1[r] = 0.0; }
  const uint4* row0 = reinterpret_cast<const uint4*>(packed + (ok0 ? j0 : 0) * ld);
  const uint4* row1 = reinterpret_cast<const uint4*>(packed + (ok1 ? j1 : 0) * ld);
  const uint64_t nkb = npad / 64;
  for (uint64_t kb = 0; kb < nkb; ++kb) {
    __syncthreads();
    for (int e = t; e < 64 * 32; e += 256) {
      int kk = e >> 5, r = e & 31;
      As[kk][r] = A[(kb * 64 + kk) * lda + rb * 32 + r];
    }
    uint4 q0 = ok0 ? row0[kb] : make_uint4(0, 0, 0, 0);
    uint4 q1 = ok1 ? row1[kb] : make_uint4(0, 0, 0, 0);
    __syncthreads();
    uint32_t w0[4] = {q0.x, q0.y, q0.z, q0.w}, w1[4] = {q1.x, q1.y, q1.z, q1.w};
#pragma unroll 4
    for (int kk = 0; kk < 64; ++kk) {
      uint32_t c0 = (w0[kk >> 4] >> (2 * (kk & 15))) & 3u;
      uint32_t c1 = (w1[kk >> 4] >> (2 * (kk & 15))) & 3u;
      double v0 = (c0 & 2u) ? ((c0 & 1u) ? l03 : l02) : ((c0 & 1u) ? 0.0 : l00);
      double v1 = (c1 & 2u) ? ((c1 & 1u) ? l13 : l12) : ((c1 & 1u) ? 0.0 : l10);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        double a = As[kk][half * 16 + r];
        acc0[r] += a * v0;
        acc1[r] += a * v1;
      }
    }
  }
  const double s0 = (outscale && ok0) ? outscale[j0] : 1.0, s1 = (outscale && ok1) ? outscale[j1] : 1.0;
  if (ok0)
#pragma unroll
    for (int r = 0; r < 16; ++r) F[j0 * ldF + rb * 32 + half * 16 + r] = acc0[r] * s0;
  if (ok1)
#pragma unroll
    for (int r = 0; r < 16; ++r) F[j1 * ldF + rb * 32 + half * 16 + r] = acc1[r] * s1;
}

// Host launcher shared with the RSVD projection (Z = Y^T B).
void launch_func_f64(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t p, uint64_t npad,
                     const double* A, uint64_t lda, int nrb, const double* mu, double* F, uint64_t ldF,
                     const double* outscale) {
  if (p == 0) return;
  dim3 grid((unsigned)cdiv(p, 256), (unsigned)nrb);
  k_func_f64<<<grid, 256, 0, s>>>(packed, ld, p, npad, A, lda, mu, F, ldF, outscale);
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Plan: small matrices of the per-SNP epilogue, computed by one thread (the
// inputs are at most 65 x 65).
//   layout of `plan`: P (d x m) | Vd (d) | Rm (d x m) | Fm (d x m) | L (m x m)
// ---------------------------------------------------------------------------
__global__ void k_plan(const double* __restrict__ Gfull, const double* __restrict__ Ginv,
                       const double* __restrict__ UtUinv, const double* __restrict__ XtXinv,
                       const double* __restrict__ ridge, const double* __restrict__ T, int d, int K,
                       int fallback, double* plan, double* work, int* info) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int m = K + d, mc = 1 + d + K;
  double* P = plan;
  double* Vd = P + d * m;
  double* Rm = Vd + d;
  double* Fm = Rm + d * m;
  double* L = Fm + d * m;
  double* E = work;                 // mc x m
  double* Mid = E + mc * m;         // d x K
  double* D = Mid + d * K;          // d x m
  double* Lx = D + d * m;           // d x d
  double* S = Lx + d * d;           // m x m
  // E: s_full = E f.  Basis order [U_c (K) | X (d)], s_basis = T^T f.
  for (int r = 0; r < m; ++r) E[0 * m + r] = 0.0;
  for (int j = 0; j < d; ++j)
    for (int r = 0; r < m; ++r) E[(1 + j) * m + r] = T[r * m + (K + j)];
  for (int u = 0; u < K; ++u)
    for (int r = 0; r < m; ++r) E[(1 + d + u) * m + r] = T[r * m + u];
  for (int j = 0; j < d; ++j) {
    for (int r = 0; r < m; ++r) {
      double s = 0.0;
      for (int c = 0; c < mc; ++c) s += Ginv[(1 + j) * mc + c] * E[c * m + r];
      P[j * m + r] = s;
    }
    Vd[j] = Ginv[(1 + j) * mc + (1 + j)];
  }
  // Mid = X^T U (U^T U)^{-1};  D = E_X - Mid E_U;  Rm = ridge D
  for (int j = 0; j < d; ++j)
    for (int u = 0; u < K; ++u) {
      double s = 0.0;
      for (int v = 0; v < K; ++v) s += Gfull[(1 + j) * mc + (1 + d + v)] * UtUinv[v * K + u];
      Mid[j * K + u] = s;
    }
  for (int j = 0; j < d; ++j)
    for (int r = 0; r < m; ++r) {
      double s = E[(1 + j) * m + r];
      for (int u = 0; u < K; ++u) s -= Mid[j * K + u] * E[(1 + d + u) * m + r];
      D[j * m + r] = s;
    }
  for (int j = 0; j < d; ++j)
    for (int r = 0; r < m; ++r) {
      double s = 0.0;
      for (int i = 0; i < d; ++i) s += ridge[j * d + i] * D[i * m + r];
      Rm[j * m + r] = s;
    }
  // Lx: Cholesky (lower) of (X^T X)^{-1};  Fm = Lx^T E_X
  int bad = 0;
  for (int j = 0; j < d; ++j) {
    double s = XtXinv[j * d + j];
    for (int k = 0; k < j; ++k) s -= Lx[j * d + k] * Lx[j * d + k];
    if (!(s > 0.0)) { bad = 1; s = 0.0; }
    double ljj = sqrt(s);
    Lx[j * d + j] = ljj;
    for (int i = j + 1; i < d; ++i) {
      double v = XtXinv[i * d + j];
      for (int k = 0; k < j; ++k) v -= Lx[i * d + k] * Lx[j * d + k];
      Lx[i * d + j] = ljj > 0.0 ? v / ljj : 0.0;
    }
    for (int i = 0; i < j; ++i) Lx[i * d + j] = 0.0;
  }
  for (int c = 0; c < d; ++c)
    for (int r = 0; r < m; ++r) {
      double s = 0.0;
      for (int i = 0; i < d; ++i) s += Lx[i * d + c] * E[(1 + i) * m + r];
      Fm[c * m + r] = s;
    }
  // Fallback RSS quadratic form S = E^T Ginv E = L L^T
  if (fallback) {
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        double s = 0.0;
        for (int c = 0; c < mc; ++c) {
          double gc = 0.0;
          for (int e = 0; e < mc; ++e) gc += Ginv[c * mc + e] * E[e * m + b];
          s += E[c * m + a] * gc;
        }
        S[a * m + b] = s;
      }
    for (int j = 0; j < m; ++j) {
      double s = S[j * m + j];
      for (int k = 0; k < j; ++k) s -= L[j * m + k] * L[j * m + k];
      if (!(s > 0.0)) { bad |= 2; s = 0.0; }
      double ljj = sqrt(s);
      L[j * m + j] = ljj;
      for (int i = j + 1; i < m; ++i) {
        double v = S[i * m + j];
        for (int k = 0; k < j; ++k) v -= L[i * m + k] * L[j * m + k];
        L[i * m + j] = ljj > 0.0 ? v / ljj : 0.0;
      }
      for (int i = 0; i < j; ++i) L[i * m + j] = 0.0;
    }
  }
  *info = bad;
}

// ---------------------------------------------------------------------------
// Per-SNP epilogue (shared by the FP64 path and the fused mixed path).
// ---------------------------------------------------------------------------
template <int MAXM>
__global__ void __launch_bounds__(128) k_epilogue(EpiArgs a, const double* __restrict__ F, uint64_t ldF) {
  extern __shared__ double sPlan[];
  const int plan_len = a.d * a.m * 3 + a.d + (a.fallback ? a.m * a.m : 0);
  for (int e = threadIdx.x; e < plan_len; e += blockDim.x) sPlan[e] = a.plan[e];
  __syncthreads();
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.p) return;
  double fu[MAXM];
#pragma unroll
  for (int r = 0; r < MAXM; ++r) fu[r] = r < a.m ? F[j * ldF + r] : 0.0;
  snp_epilogue<MAXM>(a, sPlan, j, fu, a.n_miss[j], a.sum_g[j], a.n_hom[j]);
}

// ---------------------------------------------------------------------------
// t^2 histogram (src/testing.rs:136-149): LDS sub-histogram for bins below
// 16384, global atomics beyond, integer adds only (order independent).
// ---------------------------------------------------------------------------
constexpr int kLdsBins = 16384;

__global__ void __launch_bounds__(256) k_hist(const double* __restrict__ tstat, uint64_t p, int d,
                                              uint64_t snps_per_blk, unsigned long long* hist) {
  __shared__ uint32_t h[kLdsBins];
  const int trait = blockIdx.y;
  for (int b = threadIdx.x; b < kLdsBins; b += 256) h[b] = 0;
  __syncthreads();
  const uint64_t j0 = (uint64_t)blockIdx.x * snps_per_blk, j1 = std::min(p, j0 + snps_per_blk);
  unsigned long long* hg = hist + (uint64_t)trait * kBins;
  for (uint64_t j = j0 + threadIdx.x; j < j1; j += 256) {
    double tv = tstat[j * d + trait];
    double tsq = tv * tv;
    if (!isfinite(tsq)) continue;
    double q = tsq / kBinWidth;
    int bin = q >= (double)(kBins - 1) ? kBins - 1 : (int)q;
    if (bin < kLdsBins) atomicAdd(&h[bin], 1u);
    else atomicAdd(&hg[bin], 1ull);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kLdsBins; b += 256)
    if (h[b]) atomicAdd(&hg[b], (unsigned long long)h[b]);
}

// Median walk + GIF (src/testing.rs:159-224), one workgroup for all traits.
// For even N the reference ends on the first bin whose cumulative count
// exceeds N/2, i.e. the bin of sorted position N/2.
__global__ void __launch_bounds__(256) k_gif(const unsigned long long* __restrict__ hist, int d, double* gif_pt,
                                             double* gif_avg) {
  __shared__ unsigned long long seg[256];
  __shared__ unsigned long long tot;
  __shared__ int fbin[2];
  __shared__ double raw[kMaxD];
  const int t = threadIdx.x;
  const int per = (kBins + 255) / 256;
  for (int tr = 0; tr < d; ++tr) {
    const unsigned long long* h = hist + (uint64_t)tr * kBins;
    unsigned long long s = 0;
    const int b0 = t * per, b1 = min(kBins, b0 + per);
    for (int b = b0; b < b1; ++b) s += h[b];
    seg[t] = s;
    __syncthreads();
    if (t == 0) {
      unsigned long long c = 0;
      for (int i = 0; i < 256; ++i) { unsigned long long v = seg[i]; seg[i] = c; c += v; }
      tot = c;
      fbin[0] = fbin[1] = -1;
    }
    __syncthreads();
    const unsigned long long N = tot;
    if (N > 0) {
      const bool even = (N % 2) == 0;
      const unsigned long long target = even ? N / 2 - 1 : N / 2;
      unsigned long long c = seg[t];
      for (int b = b0; b < b1; ++b) {
        unsigned long long c2 = c + h[b];
        if (c <= target && c2 > target) fbin[0] = b;
        if (c <= target + 1 && c2 > target + 1) fbin[1] = b;
        c = c2;
      }
    }
    __syncthreads();
    if (t == 0) {
      double med = 0.0;
      if (N > 0) {
        const bool even = (N % 2) == 0;
        const double w = kBinWidth;
        if (!even) med = ((double)fbin[0] + 0.5) * w;
        else med = (((double)fbin[0] + 0.5) * w + ((double)fbin[1] + 0.5) * w) / 2.0;
      }
      double r = med < 1e-10 ? 1.0 : med / 0.4549;
      raw[tr] = r;
      gif_pt[tr] = r > 1.0 ? r : 1.0;
    }
    __syncthreads();
  }
  if (t == 0) {
    double s = 0.0;
    for (int tr = 0; tr < d; ++tr) s += raw[tr];
    *gif_avg = s / (double)d;
  }
}

// p = 2 Phi(-|t / sqrt(GIF)|) = erfc(|z| / sqrt 2)  (src/testing.rs:705-709)
__global__ void k_calibrate(const double* __restrict__ tstat, uint64_t p, int d, const double* __restrict__ gif_pt,
                            double* pval) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p * (uint64_t)d) return;
  int tr = (int)(e % d);
  double gs = sqrt(gif_pt[tr]);
  double z = tstat[e] / gs;
  pval[e] = 2.0 * (0.5 * erfc(fabs(z) / 1.4142135623730951));
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
__global__ void k_build_design(const double* __restrict__ xs, const double* __restrict__ u,
                               const double* __restrict__ umean, uint64_t n, int d, int K,
                               double* cc, uint64_t ldcc, double* cf, uint64_t ldcf) {
  // cc (n x ldcc): [U_c | X] basis (centred U);  cf (n x ldcf): [1 | X | U]
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < K; ++k) cc[i * ldcc + k] = u[i * K + k] - umean[k];
  for (int j = 0; j < d; ++j) cc[i * ldcc + K + j] = xs[i * d + j];
  cf[i * ldcf] = 1.0;
  for (int j = 0; j < d; ++j) cf[i * ldcf + 1 + j] = xs[i * d + j];
  for (int k = 0; k < K; ++k) cf[i * ldcf + 1 + d + k] = u[i * K + k];
}

__global__ void k_extract_block(const double* __restrict__ G, int ld, int r0, int c0, int rows, int cols,
                                double* out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  out[e] = G[(r0 + e / cols) * ld + c0 + e % cols];
}

__global__ void k_pad_operand(const double* __restrict__ Q, uint64_t ldq, uint64_t n, int m,
                              double* A, uint64_t lda, uint64_t npad) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= npad * lda) return;
  uint64_t i = e / lda;
  int c = (int)(e % lda);
  A[e] = (i < n && c < m) ? Q[i * ldq + c] : 0.0;
}

struct AssocPlan {
  int d = 0, K = 0, m = 0, fallback = 0;
  double* Q = nullptr;      // n x m orthonormal basis (or [U_c|X] itself)
  double* plan = nullptr;   // P | Vd | Rm | Fm | L
};

// Builds the basis Q and the epilogue plan; xs, u, ridge are device pointers.
static AssocPlan make_plan(slfmm_ctx* ctx, uint64_t n, const double* d_xs, uint64_t d, const double* d_u,
                           uint64_t K, const double* d_ridge) {
  hipStream_t s = ctx->stream;
  AssocPlan pl;
  pl.d = (int)d; pl.K = (int)K; pl.m = (int)(K + d);
  const uint64_t m = K + d, mc = 1 + d + K;
  SLFMM_CHECK(m <= (uint64_t)kMaxM && d <= (uint64_t)kMaxD, SLFMM_E_ARG,
              "K + d must be <= 64 and d <= 16");
  double* umean = ctx->buf<double>("as_umean", std::max<uint64_t>(K, 1));
  if (K) launch_colmeans(s, d_u, K, n, K, umean);
  double* cc = ctx->buf<double>("as_cc", n * m);
  double* cf = ctx->buf<double>("as_cf", n * mc);
  k_build_design<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(d_xs, d_u, umean, n, (int)d, (int)K, cc, m, cf, mc);
  SLFMM_HIP(hipGetLastError());
  double* Gfull = ctx->buf<double>("as_Gfull", mc * mc);
  ts_gram(ctx, cf, mc, mc, cf, mc, mc, n, Gfull);
  double* Ginv = ctx->buf<double>("as_Ginv", mc * mc);
  const int ginv_inf = safe_inv_dev(ctx, Gfull, mc, Ginv);
  if (ginv_inf == 1) fprintf(stderr, "Warning: C^T C  where C = [1 | X | U_hat] is singular, adding epsilon diagonal regularization\n");
  SLFMM_CHECK(ginv_inf != 2, SLFMM_E_SINGULAR, "C^T C inversion failed even with epsilon regularization");
  int inf = 0;
  double* UtU = ctx->buf<double>("as_UtU", std::max<uint64_t>(K * K, 1));
  double* UtUinv = ctx->buf<double>("as_UtUinv", std::max<uint64_t>(K * K, 1));
  if (K) {
    k_extract_block<<<(unsigned)cdiv(K * K, 256), 256, 0, s>>>(Gfull, (int)mc, (int)(1 + d), (int)(1 + d), (int)K, (int)K, UtU);
    inf = safe_inv_dev(ctx, UtU, K, UtUinv);
    if (inf == 1) fprintf(stderr, "Warning: U_hat^T U_hat is singular, adding epsilon diagonal regularization\n");
    SLFMM_CHECK(inf != 2, SLFMM_E_SINGULAR, "U_hat^T U_hat inversion failed even with epsilon regularization");
  }
  double* XtX = ctx->buf<double>("as_XtX", d * d);
  double* XtXinv = ctx->buf<double>("as_XtXinv", d * d);
  k_extract_block<<<(unsigned)cdiv(d * d, 256), 256, 0, s>>>(Gfull, (int)mc, 1, 1, (int)d, (int)d, XtX);
  inf = safe_inv_dev(ctx, XtX, d, XtXinv);
  if (inf == 1) fprintf(stderr, "Warning: X^T X for sequential R^2 is singular, adding epsilon diagonal regularization\n");
  SLFMM_CHECK(inf != 2, SLFMM_E_SINGULAR, "X^T X inversion failed even with epsilon regularization");
  // Orthonormal basis of [U_c | X]; keep [U_c|X] if it is rank deficient or
  // if C^T C needed regularisation (then the reference's own Gram form rules).
  pl.Q = ctx->buf<double>("as_Q", n * m);
  double* T = ctx->buf<double>("as_T", m * m);
  SLFMM_HIP(hipMemcpyAsync(pl.Q, cc, n * m * 8, hipMemcpyDeviceToDevice, s));
  bool ok = (ginv_inf == 0) && cholqr2(ctx, pl.Q, m, m, n, T);
  if (!ok) {
    pl.fallback = 1;
    SLFMM_HIP(hipMemcpyAsync(pl.Q, cc, n * m * 8, hipMemcpyDeviceToDevice, s));
    std::vector<double> I(m * m, 0.0);
    for (uint64_t i = 0; i < m; ++i) I[i * m + i] = 1.0;
    SLFMM_HIP(hipMemcpyAsync(T, I.data(), m * m * 8, hipMemcpyHostToDevice, s));
    SLFMM_HIP(hipStreamSynchronize(s));
  }
  const uint64_t plan_len = d * m * 3 + d + m * m;
  pl.plan = ctx->buf<double>("as_plan", plan_len);
  double* work = ctx->buf<double>("as_work", mc * m + d * K + d * m + d * d + m * m + 16);
  int* dinfo = ctx->buf<int>("plan_info", 1);
  k_plan<<<1, 64, 0, s>>>(Gfull, Ginv, UtUinv, XtXinv, d_ridge, T, (int)d, (int)K, pl.fallback, pl.plan, work, dinfo);
  SLFMM_HIP(hipGetLastError());
  return pl;
}

void launch_func_i8(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_Q, uint64_t m, const EpiArgs& ea,
                    bool fused_epilogue);   // assoc_i8.hip

void assoc_run(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_xs, uint64_t d, const double* d_u,
               uint64_t k, double lambda, const slfmm_config& cfg, double* d_beta, double* d_t,
               double* d_r2, double* d_p, uint64_t* d_hist, double* d_gif_pt, double* d_gif) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n, p = g->p;
  SLFMM_CHECK(n > 1 + d + k, SLFMM_E_DF,
              "Insufficient degrees of freedom: n=" + std::to_string(n) + " samples but 1+d+K=1+" +
                  std::to_string(d) + "+" + std::to_string(k) + "=" + std::to_string(1 + d + k) +
                  ". Need n > 1 + d + K for valid t-tests. Reduce K or add more samples.");
  SLFMM_CHECK(lambda > 0.0, SLFMM_E_LAMBDA, "Ridge penalty λ must be positive");
  // ridge_inv = (X^T X + lambda I)^{-1}  (src/precompute.rs:73-78)
  double* xtx = ctx->buf<double>("as_xtxr", d * d);
  ts_gram(ctx, d_xs, d, d, d_xs, d, d, n, xtx);
  std::vector<double> h(d * d);
  SLFMM_HIP(hipMemcpyAsync(h.data(), xtx, d * d * 8, hipMemcpyDeviceToHost, s));
  SLFMM_HIP(hipStreamSynchronize(s));
  for (uint64_t j = 0; j < d; ++j) h[j * d + j] += lambda;
  SLFMM_HIP(hipMemcpyAsync(xtx, h.data(), d * d * 8, hipMemcpyHostToDevice, s));
  double* ridge = ctx->buf<double>("as_ridge", d * d);
  int inf = safe_inv_dev(ctx, xtx, d, ridge);
  if (inf == 1) fprintf(stderr, "Warning: X^T X + λI (ridge) is singular, adding epsilon diagonal regularization\n");
  SLFMM_CHECK(inf != 2, SLFMM_E_SINGULAR, "X^T X + λI (ridge) inversion failed even with epsilon regularization");

  AssocPlan pl = make_plan(ctx, n, d_xs, d, d_u, k, ridge);
  const uint64_t m = pl.m;
  int32_t* nmiss = ctx->buf<int32_t>("as_nmiss", p);
  int32_t* sumg = ctx->buf<int32_t>("as_sumg", p);
  int32_t* nhom = ctx->buf<int32_t>("as_nhom", p);
  EpiArgs ea;
  ea.n = n; ea.p = p; ea.d = (int)d; ea.K = (int)k; ea.m = (int)m; ea.fallback = pl.fallback;
  ea.norm = cfg.norm; ea.df = (double)(n - 1 - d - k); ea.plan = pl.plan;
  ea.n_miss = nmiss; ea.sum_g = sumg; ea.n_hom = nhom;
  ea.beta = d_beta; ea.tstat = d_t; ea.r2 = d_r2;
  if (cfg.precision == SLFMM_PREC_I8) {
    slfmm_timed tm(ctx, "assoc_stream");
    launch_func_i8(ctx, g, pl.Q, m, ea, true);
  } else {
    launch_stats(s, g->d, g->ld, n, p, nmiss, sumg, nhom);
    double* mu = ctx->buf<double>("as_mu", p);
    double* sc = ctx->buf<double>("as_sc", p);
    launch_mean_scale(s, nmiss, sumg, n, p, cfg.norm, mu, sc);
    const uint64_t npad = rup(n, 64);
    const int nrb = (int)cdiv(m, kRowBlk);
    const uint64_t lda = (uint64_t)nrb * kRowBlk;
    double* A = ctx->buf<double>("as_A64", npad * lda);
    k_pad_operand<<<(unsigned)cdiv(npad * lda, 256), 256, 0, s>>>(pl.Q, m, n, (int)m, A, lda, npad);
    double* F = ctx->buf<double>("as_F", p * lda);
    {
      slfmm_timed tm(ctx, "assoc_stream");
      launch_func_f64(s, g->d, g->ld, p, npad, A, lda, nrb, mu, F, lda, nullptr);
    }
    const int plan_len = (int)(d * m * 3 + d + (pl.fallback ? m * m : 0));
    const size_t shm = plan_len * sizeof(double);
    if (m <= 32) k_epilogue<32><<<(unsigned)cdiv(p, 128), 128, shm, s>>>(ea, F, lda);
    else k_epilogue<64><<<(unsigned)cdiv(p, 128), 128, shm, s>>>(ea, F, lda);
    SLFMM_HIP(hipGetLastError());
  }
  // histogram, GIF, calibration
  unsigned long long* hist = d_hist ? (unsigned long long*)d_hist
                                    : ctx->buf<unsigned long long>("as_hist", d * (uint64_t)kBins);
  SLFMM_HIP(hipMemsetAsync(hist, 0, d * (uint64_t)kBins * 8, s));
  const uint64_t spb = 1ull << 16;
  if (p) {
    k_hist<<<dim3((unsigned)cdiv(p, spb), (unsigned)d), 256, 0, s>>>(d_t, p, (int)d, spb, hist);
    SLFMM_HIP(hipGetLastError());
  }
  k_gif<<<1, 256, 0, s>>>(hist, (int)d, d_gif_pt, d_gif);
  if (p) k_calibrate<<<(unsigned)cdiv(p * d, 256), 256, 0, s>>>(d_t, p, (int)d, d_gif_pt, d_p);
  SLFMM_HIP(hipGetLastError());
}

}  // namespace slfmm
