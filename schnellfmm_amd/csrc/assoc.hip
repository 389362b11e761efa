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
// FP64 functional kernel:  F[j][NH RW rb + r] = sum_k A[k][NH RW rb + r] v_jk
// with v = g - mu_j on observed calls and 0 on missing calls (mean
// imputation after centring, src/bed.rs:339-343).  Tile: (4 / NH) x 128
// SNPs x NH RW rows; with NH = 2, wave w -> rows RW (w & 1)..+RW of the row
// block and SNP group w >> 1; with NH = 1 every wave takes all RW rows and
// its own SNP group; lane -> 2 SNPs.  (RW, NH) follows the operand's width
// ((8, 1), (8, 2), (12, 2), then blocks of (16, 2)), so narrow operands pay
// neither for padding rows nor for decoding codes for a few FMAs.  The 16 A values of a sample are the same for every lane of the
// wave, so they come in by scalar loads (two samples' rows in flight) and
// feed the FMAs from SGPRs: no LDS staging and no barrier (the LDS
// broadcast of one value per two FMAs bounded the earlier form).
// ---------------------------------------------------------------------------
// With a fill rank FK > 0, missing calls take the value W_k . H_j - mu_j (NMF
// imputation after centring, rank FK = fs.K): the 64 W rows of the block are
// staged in LDS (read only where a call is missing), and each lane keeps the
// FK loadings of its two SNPs in registers.
template <int FK, int RW, int NH>
__global__ void __launch_bounds__(256) k_func_f64(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t p,
                                                  uint64_t npad, const double* __restrict__ A, uint64_t lda,
                                                  const double* __restrict__ mu, double* __restrict__ F,
                                                  uint64_t ldF, const double* __restrict__ outscale, FillSpec fs) {
  constexpr bool FILL = FK > 0;
  constexpr int KR = FILL ? FK : 1;
  __shared__ double Ws[FILL ? 64 : 1][KR + 1];
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);   // wave-uniform: A rows by scalar loads
  const int half = NH == 2 ? (w & 1) : 0, grp = NH == 2 ? (w >> 1) : w;
  const int rb = blockIdx.y;
  const uint64_t j0 = (uint64_t)blockIdx.x * (512 / NH) + grp * 128 + lane;
  const uint64_t j1 = j0 + 64;
  const bool ok0 = j0 < p, ok1 = j1 < p;
  const double m0 = ok0 ? mu[j0] : 0.0, m1 = ok1 ? mu[j1] : 0.0;
  const double l00 = -m0, l02 = 1.0 - m0, l03 = 2.0 - m0;
  const double l10 = -m1, l12 = 1.0 - m1, l13 = 2.0 - m1;
  double h0[KR], h1[KR];
  if constexpr (FILL) {
#pragma unroll
    for (int r = 0; r < KR; ++r) {
      h0[r] = ok0 ? fs.Ht[j0 * KR + r] : 0.0;
      h1[r] = ok1 ? fs.Ht[j1 * KR + r] : 0.0;
    }
  }
  double acc0[RW], acc1[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) { acc0[r] = 0.0; acc1[r] = 0.0; }
  const uint4* row0 = reinterpret_cast<const uint4*>(packed + (ok0 ? j0 : 0) * ld);
  const uint4* row1 = reinterpret_cast<const uint4*>(packed + (ok1 ? j1 : 0) * ld);
  const double* Acol = A + rb * NH * RW + half * RW;
  const uint64_t nkb = npad / 64;
  for (uint64_t kb = 0; kb < nkb; ++kb) {
    const uint4 q0 = ok0 ? row0[kb] : make_uint4(0, 0, 0, 0);
    const uint4 q1 = ok1 ? row1[kb] : make_uint4(0, 0, 0, 0);
    if constexpr (FILL) {   // the block's W rows, read only where a call is missing
      __syncthreads();
      for (int e = t; e < 64 * KR; e += 256) {
        const int kk = e / KR, r = e % KR;
        const uint64_t k = kb * 64 + kk;
        Ws[kk][r] = k < fs.n ? fs.W[k * KR + r] : 0.0;
      }
      __syncthreads();
    }
#pragma unroll 1
    for (int wq = 0; wq < 4; ++wq) {   // one 16-code word per SNP
    const uint32_t x0 = wq == 0 ? q0.x : wq == 1 ? q0.y : wq == 2 ? q0.z : q0.w;
    const uint32_t x1 = wq == 0 ? q1.x : wq == 1 ? q1.y : wq == 2 ? q1.z : q1.w;
#pragma unroll 2   // two samples' A rows (64 SGPRs) in flight
    for (int c = 0; c < 16; ++c) {
      const uint64_t k = kb * 64 + 16 * wq + c;
      const uint32_t c0 = (x0 >> (2 * c)) & 3u;
      const uint32_t c1 = (x1 >> (2 * c)) & 3u;
      double miss0 = 0.0, miss1 = 0.0;
      if constexpr (FILL) {
        if (c0 == 1u || c1 == 1u) {
          double f0 = 0.0, f1 = 0.0;
#pragma unroll
          for (int r = 0; r < KR; ++r) {
            const double wr = Ws[16 * wq + c][r];
            f0 += wr * h0[r];
            f1 += wr * h1[r];
          }
          miss0 = f0 - m0;
          miss1 = f1 - m1;
        }
      }
      const double v0 = (c0 & 2u) ? ((c0 & 1u) ? l03 : l02) : ((c0 & 1u) ? miss0 : l00);
      const double v1 = (c1 & 2u) ? ((c1 & 1u) ? l13 : l12) : ((c1 & 1u) ? miss1 : l10);
      const double* ak = Acol + k * lda;
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const double a = ak[r];
        acc0[r] += a * v0;
        acc1[r] += a * v1;
      }
    }
    }
  }
  const double s0 = (outscale && ok0) ? outscale[j0] : 1.0, s1 = (outscale && ok1) ? outscale[j1] : 1.0;
  if (ok0)
#pragma unroll
    for (int r = 0; r < RW; ++r) F[j0 * ldF + rb * NH * RW + half * RW + r] = acc0[r] * s0;
  if (ok1)
#pragma unroll
    for (int r = 0; r < RW; ++r) F[j1 * ldF + rb * NH * RW + half * RW + r] = acc1[r] * s1;
}

// Host launcher shared with the RSVD projection (Z = Y^T B) and the NMF
// functionals; `fill` (optional) selects NMF imputation of missing calls.
// Columns [0, m) of F are written (A and F hold at least rup(m, 32)).
template <int RW, int NH>
static void launch_func_rw(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t p, uint64_t npad,
                           const double* A, uint64_t lda, int m, const double* mu, double* F, uint64_t ldF,
                           const double* outscale, const FillSpec* fill) {
  const dim3 grid((unsigned)cdiv(p, 512 / NH), (unsigned)cdiv(m, NH * RW));
  if (fill && fill->W) {
    SLFMM_CHECK(fill->K >= 1 && fill->K <= kNmfMaxK, SLFMM_E_ARG, "NMF rank must be 1..16");
    switch (fill->K) {
#define SLFMM_FUNC(R)                                                                                         \
  case R:                                                                                                     \
    k_func_f64<R, RW, NH><<<grid, 256, 0, s>>>(packed, ld, p, npad, A, lda, mu, F, ldF, outscale, *fill);     \
    break;
      SLFMM_NMF_RANKS(SLFMM_FUNC)
#undef SLFMM_FUNC
    }
  } else {
    k_func_f64<0, RW, NH><<<grid, 256, 0, s>>>(packed, ld, p, npad, A, lda, mu, F, ldF, outscale, FillSpec{});
  }
}

void launch_func_f64(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t p, uint64_t npad,
                     const double* A, uint64_t lda, int m, const double* mu, double* F, uint64_t ldF,
                     const double* outscale, const FillSpec* fill) {
  if (p == 0) return;
  SLFMM_CHECK(m >= 1 && lda >= rup((uint64_t)m, (uint64_t)kRowBlk) && ldF >= (uint64_t)m, SLFMM_E_ARG,
              "FP64 functionals: operand or output stride below the operand width");
  // row blocks of NH RW: one block up to 24 rows, blocks of 32 beyond (within rup(m, 32))
  // measured at n = 5000, p = 2M: m = 8 without a fill 8.4 -> 6.4 ms from
  // (4, 2) to (8, 1); m = 10 with a fill 19.3 ms at (8, 2), 21.2 at (16, 1)
  if (m <= 8) launch_func_rw<8, 1>(s, packed, ld, p, npad, A, lda, m, mu, F, ldF, outscale, fill);
  else if (m <= 16) launch_func_rw<8, 2>(s, packed, ld, p, npad, A, lda, m, mu, F, ldF, outscale, fill);
  else if (m <= 24) launch_func_rw<12, 2>(s, packed, ld, p, npad, A, lda, m, mu, F, ldF, outscale, fill);
  else launch_func_rw<16, 2>(s, packed, ld, p, npad, A, lda, m, mu, F, ldF, outscale, fill);
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Moments of NMF-filled rows (decode with fill, src/bed.rs:211-235): one
// thread per row; S1 = sum of the fills, S2 = sum of their squares, and with
// Wp the sum of the Wp rows over the missing calls.
// ---------------------------------------------------------------------------
// One thread per SNP row; the W (and Wp) rows of a sample are the same for
// every thread, so they come in by scalar loads (the rank K is a template
// parameter), and the fill's sums are taken on every call with the
// observed ones adding zero: no per-call branch, and the sums keep the
// per-entry order.
template <int K>
__global__ void __launch_bounds__(256) k_fill_stats(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t rows,
                                                    FillSpec fs, const double* __restrict__ Wp,
                                                    double* __restrict__ S1, double* __restrict__ S2,
                                                    double* __restrict__ SWp) {
  const int t = threadIdx.x;
  const uint64_t j = (uint64_t)blockIdx.x * 256 + t;
  const bool ok = j < rows;
  double h[K], sp[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    h[r] = ok ? fs.Ht[j * K + r] : 0.0;
    sp[r] = 0.0;
  }
  double s1 = 0.0, s2 = 0.0;
  const uint4* row = reinterpret_cast<const uint4*>(packed + (ok ? j : 0) * ld);
  const uint64_t n = fs.n, nkb = cdiv(n, 64);
  for (uint64_t kb = 0; kb < nkb; ++kb) {
    const uint4 q = ok ? row[kb] : make_uint4(0, 0, 0, 0);
#pragma unroll 1
    for (int wq = 0; wq < 4; ++wq) {
      const uint32_t x = wq == 0 ? q.x : wq == 1 ? q.y : wq == 2 ? q.z : q.w;
      const uint64_t kw = kb * 64 + 16 * wq;
      const int cn = (int)std::min<uint64_t>(16, kw < n ? n - kw : 0);   // samples of this word
#pragma unroll 2
      for (int c = 0; c < cn; ++c) {
        const uint64_t k = kw + c;
        const bool miss = ((x >> (2 * c)) & 3u) == 1u;
        double f = 0.0;
#pragma unroll
        for (int r = 0; r < K; ++r) f += fs.W[k * K + r] * h[r];
        s1 += miss ? f : 0.0;
        s2 += miss ? f * f : 0.0;
        if (Wp)
#pragma unroll
          for (int r = 0; r < K; ++r) sp[r] += miss ? Wp[k * K + r] : 0.0;
      }
    }
  }
  if (!ok) return;
  S1[j] = s1;
  S2[j] = s2;
  if (SWp)
    for (int r = 0; r < K; ++r) SWp[j * K + r] = sp[r];
}

__global__ void k_fill_moments(const int32_t* __restrict__ nmiss, const int32_t* __restrict__ sumg,
                               const int32_t* __restrict__ nhom, const double* __restrict__ S1,
                               const double* __restrict__ S2, uint64_t n, uint64_t rows, int norm,
                               double* __restrict__ mu, double* __restrict__ sc, double* __restrict__ tss) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= rows) return;
  const double nm = (double)nmiss[j], sg = (double)sumg[j], nobs = (double)n - nm;
  const double sg2 = sg + 2.0 * (double)nhom[j];   // sum of g^2 over observed calls
  const double m = n ? (sg + S1[j]) / (double)n : 0.0;
  const double s = epi_scale(m, norm);
  const double obs = sg2 - 2.0 * m * sg + nobs * m * m;
  const double fil = S2[j] - 2.0 * m * S1[j] + nm * m * m;
  mu[j] = m;
  sc[j] = s;
  if (tss) tss[j] = s * s * (obs + fil);
}

void fill_moments(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t rows, const FillSpec& fs, int norm,
                  const double* Wp, double* mu, double* sc, double* tss, double* SWp, bool i8) {
  hipStream_t s = ctx->stream;
  SLFMM_CHECK(fs.K >= 1 && fs.K <= kNmfMaxK, SLFMM_E_ARG, "NMF rank must be 1..16");
  int32_t* nmiss = ctx->buf<int32_t>("fl_nmiss", rows);
  int32_t* sumg = ctx->buf<int32_t>("fl_sumg", rows);
  int32_t* nhom = ctx->buf<int32_t>("fl_nhom", rows);
  double* S1 = ctx->buf<double>("fl_s1", rows);
  double* S2 = ctx->buf<double>("fl_s2", rows);
  launch_stats(s, packed, ld, fs.n, rows, nmiss, sumg, nhom);
  if (rows) {
    if (i8 && fs.K <= kNmfMaxK8) {
      fill_stats_i8(ctx, packed, ld, fs.n, rows, fs.W, fs.Ht, fs.K, Wp, S1, S2, SWp);
    } else {
      switch (fs.K) {
#define SLFMM_FST(R) \
  case R: k_fill_stats<R><<<(unsigned)cdiv(rows, 256), 256, 0, s>>>(packed, ld, rows, fs, Wp, S1, S2, SWp); break;
        SLFMM_NMF_RANKS(SLFMM_FST)
#undef SLFMM_FST
      }
    }
    k_fill_moments<<<(unsigned)cdiv(rows, 256), 256, 0, s>>>(nmiss, sumg, nhom, S1, S2, fs.n, rows, norm, mu, sc, tss);
  }
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Plan: small matrices of the per-SNP epilogue (inputs at most 65 x 65),
// one workgroup.  Layout of `plan` (plan_len(d, m) doubles):
//   P (d x m) | Vd (d) | Rm (d x m) | Fm (d x m) | L (m x m) | fallback flag
// The basis falls back to [U_c|X] itself (T = I, RSS by the quadratic form
// L L^T = E^T Ginv E) when C^T C needed regularisation or when CholeskyQR
// of [U_c|X] failed; both outcomes are read from device status words.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int plan_fallback(const int* ginv_st, const int* cq_st) {
  return (*ginv_st != 0 || *cq_st >= 2) ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_plan(const double* __restrict__ Gfull, const double* __restrict__ Ginv,
                                              const double* __restrict__ UtUinv, const double* __restrict__ XtXinv,
                                              const double* __restrict__ ridge, const double* __restrict__ T, int d,
                                              int K, const int* ginv_st, const int* cq_st, double* plan,
                                              double* work) {
  __shared__ double sT[kMaxM * kMaxM];
  __shared__ double sMid[kMaxD * kMaxM];
  __shared__ double sD[kMaxD * kMaxM];
  __shared__ double sLx[kMaxD * kMaxD];
  const int t = threadIdx.x;
  const int m = K + d, mc = 1 + d + K;
  const int fb = plan_fallback(ginv_st, cq_st);
  for (int e = t; e < m * m; e += 256) sT[e] = fb ? ((e / m == e % m) ? 1.0 : 0.0) : T[e];
  __syncthreads();
  // E (mc x m): s_full = E f, basis order [U_c (K) | X (d)], s_basis = T^T f
  auto E = [&](int c, int r) -> double {
    if (c == 0) return 0.0;
    if (c <= d) return sT[r * m + K + (c - 1)];
    return sT[r * m + (c - 1 - d)];
  };
  double* P = plan;
  double* Vd = P + d * m;
  double* Rm = Vd + d;
  double* Fm = Rm + d * m;
  double* L = Fm + d * m;
  double* flag = L + m * m;
  for (int e = t; e < d * m; e += 256) {
    const int j = e / m, r = e % m;
    double s = 0.0;
    for (int c = 0; c < mc; ++c) s += Ginv[(1 + j) * mc + c] * E(c, r);
    P[e] = s;
  }
  if (t < d) Vd[t] = Ginv[(1 + t) * mc + (1 + t)];
  // Mid = X^T U (U^T U)^{-1}
  for (int e = t; e < d * K; e += 256) {
    const int j = e / K, u = e % K;
    double s = 0.0;
    for (int v = 0; v < K; ++v) s += Gfull[(1 + j) * mc + (1 + d + v)] * UtUinv[v * K + u];
    sMid[e] = s;
  }
  __syncthreads();
  // D = E_X - Mid E_U
  for (int e = t; e < d * m; e += 256) {
    const int j = e / m, r = e % m;
    double s = E(1 + j, r);
    for (int u = 0; u < K; ++u) s -= sMid[j * K + u] * E(1 + d + u, r);
    sD[e] = s;
  }
  if (t == 0) {   // Lx: lower Cholesky factor of (X^T X)^{-1}
    for (int j = 0; j < d; ++j) {
      double s = XtXinv[j * d + j];
      for (int k = 0; k < j; ++k) s -= sLx[j * d + k] * sLx[j * d + k];
      const double ljj = s > 0.0 ? sqrt(s) : 0.0;
      sLx[j * d + j] = ljj;
      for (int i = j + 1; i < d; ++i) {
        double v = XtXinv[i * d + j];
        for (int k = 0; k < j; ++k) v -= sLx[i * d + k] * sLx[j * d + k];
        sLx[i * d + j] = ljj > 0.0 ? v / ljj : 0.0;
      }
      for (int i = 0; i < j; ++i) sLx[i * d + j] = 0.0;
    }
  }
  __syncthreads();
  for (int e = t; e < d * m; e += 256) {
    const int j = e / m, r = e % m;
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += ridge[j * d + i] * sD[i * m + r];
    Rm[e] = s;
    double f = 0.0;   // Fm = Lx^T E_X
    for (int i = 0; i < d; ++i) f += sLx[i * d + j] * E(1 + i, r);
    Fm[e] = f;
  }
  if (!fb) {
    for (int e = t; e < m * m; e += 256) L[e] = 0.0;
  } else {
    // S = E^T Ginv E = L L^T  (GE = Ginv E staged in work)
    double* GE = work;          // mc x m
    double* S = work + mc * m;  // m x m
    for (int e = t; e < mc * m; e += 256) {
      const int c = e / m, b = e % m;
      double g = 0.0;
      for (int q = 0; q < mc; ++q) g += Ginv[c * mc + q] * E(q, b);
      GE[e] = g;
    }
    __syncthreads();
    for (int e = t; e < m * m; e += 256) {
      const int a = e / m, b = e % m;
      double v = 0.0;
      for (int c = 0; c < mc; ++c) v += E(c, a) * GE[c * m + b];
      S[e] = v;
    }
    __syncthreads();
    if (t == 0) {
      for (int j = 0; j < m; ++j) {
        double v = S[j * m + j];
        for (int k = 0; k < j; ++k) v -= L[j * m + k] * L[j * m + k];
        const double ljj = v > 0.0 ? sqrt(v) : 0.0;
        L[j * m + j] = ljj;
        for (int i = j + 1; i < m; ++i) {
          double w = S[i * m + j];
          for (int k = 0; k < j; ++k) w -= L[i * m + k] * L[j * m + k];
          L[i * m + j] = ljj > 0.0 ? w / ljj : 0.0;
        }
        for (int i = 0; i < j; ++i) L[i * m + j] = 0.0;
      }
    }
  }
  if (t == 0) *flag = (double)fb;
}

// Q <- [U_c|X] when the plan falls back to it.
__global__ void k_select_basis(const int* ginv_st, const int* cq_st, const double* __restrict__ cc,
                               double* __restrict__ Q, uint64_t count) {
  if (!plan_fallback(ginv_st, cq_st)) return;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (uint64_t)gridDim.x * blockDim.x)
    Q[e] = cc[e];
}

// ---------------------------------------------------------------------------
// Per-SNP epilogue (shared by the FP64 path and the fused mixed path).
// ---------------------------------------------------------------------------
template <int MAXM>
__global__ void __launch_bounds__(128) k_epilogue(EpiArgs a, const double* __restrict__ F, uint64_t ldF) {
  extern __shared__ double sPlan[];
  const int len = plan_len(a.d, a.m);
  for (int e = threadIdx.x; e < len; e += blockDim.x) sPlan[e] = a.plan[e];
  __syncthreads();
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.p) return;
  double fu[MAXM];
#pragma unroll
  for (int r = 0; r < MAXM; ++r) fu[r] = r < a.m ? F[j * ldF + r] : 0.0;
  snp_epilogue<MAXM>(a, sPlan, j, fu, a.n_miss[j], a.sum_g[j], a.n_hom[j]);
}

// ---------------------------------------------------------------------------
// t^2 histogram (src/testing.rs:136-149) in two integer passes:
//  k_hist_part: each workgroup counts a SNP range into LDS bins [0, LB) per
//    trait (values beyond go straight to the global histogram by atomics,
//    they are rare), then stores its LDS counts as a partial row;
//  k_hist_sum:  adds the partial rows into the histogram, 16 slices of rows
//    in parallel (64-bit integer atomics);
//  k_hist_chunks: per-chunk totals (1024 bins per chunk) for the median search.
// Integer adds only, so the counts do not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int kChunkBins = 1024;
constexpr int kHistChunks = (kBins + kChunkBins - 1) / kChunkBins;

__global__ void __launch_bounds__(1024) k_hist_part(const double* __restrict__ tstat, uint64_t p, int d,
                                                    uint64_t snps_per_blk, int LB, uint32_t* __restrict__ part,
                                                    unsigned long long* __restrict__ hist) {
  extern __shared__ uint32_t h[];   // d x LB
  const int t = threadIdx.x;
  for (int b = t; b < d * LB; b += 1024) h[b] = 0;
  __syncthreads();
  const uint64_t j0 = (uint64_t)blockIdx.x * snps_per_blk;
  const uint64_t j1 = std::min(p, j0 + snps_per_blk);
  // trait of element e = e mod d, advanced by 1024 mod d per step (no 64-bit modulo in the loop)
  const int step = 1024 % d;
  int tr = (int)((j0 * d + t) % d);
  // four elements per pass, their loads issued together (elements past the
  // range read as NaN and are skipped like any non-finite statistic)
  const uint64_t e1 = j1 * d;
  for (uint64_t e = j0 * d + t; e < e1; e += 4096) {
    double tv[4];
    int trk[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      tv[k] = e + 1024 * k < e1 ? tstat[e + 1024 * k] : __builtin_nan("");
      trk[k] = tr;
      tr = tr + step >= d ? tr + step - d : tr + step;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double tsq = tv[k] * tv[k];
      if (!isfinite(tsq)) continue;
      const double q = tsq / kBinWidth;
      const int bin = q >= (double)(kBins - 1) ? kBins - 1 : (int)q;
      if (bin < LB) atomicAdd(&h[trk[k] * LB + bin], 1u);
      else atomicAdd(&hist[(uint64_t)trk[k] * kBins + bin], 1ull);
    }
  }
  __syncthreads();
  uint32_t* row = part + (uint64_t)blockIdx.x * d * LB;
  for (int b = t; b < d * LB; b += 1024) row[b] = h[b];
}

// hist[tr][b] += sum of the partial rows g in [z G, z G + G) for bins b < LB
// (integer atomics: the totals do not depend on the order).
__global__ void __launch_bounds__(256) k_hist_sum(const uint32_t* __restrict__ part, int nblk, int d, int LB,
                                                  int rows_per_z, unsigned long long* __restrict__ hist) {
  const int b = blockIdx.x * 256 + threadIdx.x, tr = blockIdx.y;
  if (b >= LB) return;
  const int g0 = blockIdx.z * rows_per_z, g1 = min(nblk, g0 + rows_per_z);
  unsigned long long v = 0;
  for (int g = g0; g < g1; ++g) v += part[((uint64_t)g * d + tr) * LB + b];
  if (v) atomicAdd(&hist[(uint64_t)tr * kBins + b], v);
}

// Per-chunk totals of the final histogram (kChunkBins bins per chunk).
__global__ void __launch_bounds__(256) k_hist_chunks(const unsigned long long* __restrict__ hist,
                                                     unsigned long long* __restrict__ chunk_tot) {
  __shared__ unsigned long long red[256];
  const int t = threadIdx.x, tr = blockIdx.y, ch = blockIdx.x;
  unsigned long long local = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int b = ch * kChunkBins + i * 256 + t;
    if (b < kBins) local += hist[(uint64_t)tr * kBins + b];
  }
  red[t] = local;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (t < s2) red[t] += red[t + s2];
    __syncthreads();
  }
  if (t == 0) chunk_tot[tr * kHistChunks + ch] = red[0];
}

// Median walk + GIF (src/testing.rs:159-224), one workgroup for all traits.
// For even N the reference ends on the first bin whose cumulative count
// exceeds N/2, i.e. the bin of sorted position N/2; for odd N the bin of
// position (N-1)/2.  The chunk holding a position is found from the chunk
// totals, then the bin inside it by a workgroup scan.
// Exclusive prefix sum of one value per thread over a 256-thread block (exact
// integer sums, so the order of the additions does not matter); total gets
// the block sum.  Every thread of the block must call it.
__device__ unsigned long long block_exclusive_scan(unsigned long long v, unsigned long long* wsum,
                                                   unsigned long long& total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  unsigned long long base = 0;
  for (int i = 0; i < w; ++i) base += wsum[i];
  total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();   // wsum may be reused by the next call
  return base + x - v;
}

__global__ void __launch_bounds__(256) k_gif(const unsigned long long* __restrict__ hist,
                                             const unsigned long long* __restrict__ chunk_tot, int d,
                                             double* gif_pt, double* gif_avg) {
  __shared__ unsigned long long cum[kHistChunks + 1];
  __shared__ int fbin[2];
  __shared__ double raw[kMaxD];
  __shared__ unsigned long long wsum[4];
  static_assert(kHistChunks < 256, "one chunk total per thread");
  const int t = threadIdx.x;
  for (int tr = 0; tr < d; ++tr) {
    const unsigned long long* h = hist + (uint64_t)tr * kBins;
    {   // cum[i] = sum of the chunk totals before chunk i; cum[kHistChunks] = N
      unsigned long long tot = 0;
      const unsigned long long e =
          block_exclusive_scan(t < kHistChunks ? chunk_tot[tr * kHistChunks + t] : 0ull, wsum, tot);
      if (t <= kHistChunks) cum[t] = e;
    }
    if (t == 0) fbin[0] = fbin[1] = -1;
    __syncthreads();
    const unsigned long long N = cum[kHistChunks];
    const bool even = (N % 2) == 0;
    const int npos = N == 0 ? 0 : (even ? 2 : 1);
    for (int q = 0; q < npos; ++q) {
      const unsigned long long target = (even ? N / 2 - 1 : N / 2) + q;
      int ch = 0;   // chunk whose cumulative range holds target (uniform)
      while (!(cum[ch] <= target && target < cum[ch + 1])) ++ch;
      unsigned long long v[4], s = 0;
      const int b0 = ch * kChunkBins + 4 * t;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = (b0 + i < kBins) ? h[b0 + i] : 0ull;
        s += v[i];
      }
      unsigned long long tot = 0;
      unsigned long long c = cum[ch] + block_exclusive_scan(s, wsum, tot);   // count below bin b0
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const unsigned long long c2 = c + v[i];
        if (c <= target && c2 > target) fbin[q] = b0 + i;
        c = c2;
      }
      __syncthreads();
    }
    if (t == 0) {
      double med = 0.0;
      if (N > 0) {
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
// One thread per SNP, looping over its d traits (no 64-bit modulo).  The
// kernel is bound by the FP64 erfc (a flat, coalesced element loop with four
// loads in flight per thread measured no faster).
__global__ void k_calibrate(const double* __restrict__ tstat, uint64_t p, int d, const double* __restrict__ gif_pt,
                            double* pval) {
  __shared__ double gsq[kMaxD];   // sqrt(GIF) per trait, once per workgroup
  if ((int)threadIdx.x < d) gsq[threadIdx.x] = sqrt(gif_pt[threadIdx.x]);
  __syncthreads();
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p) return;
  for (int tr = 0; tr < d; ++tr) {
    const double gs = gsq[tr];
    const uint64_t e = j * d + tr;
    const double z = tstat[e] / gs;
    pval[e] = 2.0 * (0.5 * erfc(fabs(z) / 1.4142135623730951));
  }
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

void pad_operand(hipStream_t s, const double* Q, uint64_t ldq, uint64_t n, int m, double* A, uint64_t lda,
                 uint64_t npad) {
  k_pad_operand<<<(unsigned)cdiv(npad * lda, 256), 256, 0, s>>>(Q, ldq, n, m, A, lda, npad);
  SLFMM_HIP(hipGetLastError());
}

struct AssocPlan {
  int d = 0, K = 0, m = 0;
  double* Q = nullptr;      // n x m orthonormal basis (or [U_c|X] itself)
  double* plan = nullptr;   // P | Vd | Rm | Fm | L | fallback flag
};

// Builds the basis Q and the epilogue plan; xs, u, ridge are device pointers.
static AssocPlan make_plan(slfmm_ctx* ctx, uint64_t n, const double* d_xs, uint64_t d, const double* d_u,
                           uint64_t K, const double* d_ridge) {
  hipStream_t s = ctx->stream;
  AssocPlan pl;
  pl.d = (int)d; pl.K = (int)K; pl.m = (int)(K + d);
  const uint64_t m = K + d, mc = 1 + d + K;
  SLFMM_CHECK(m <= (uint64_t)kMaxM && mc <= (uint64_t)kMaxM && d <= (uint64_t)kMaxD, SLFMM_E_ARG,
              "1 + d + K must be <= 64 and d <= 16");
  double* umean = ctx->buf<double>("as_umean", std::max<uint64_t>(K, 1));
  if (K) colmeans(ctx, d_u, K, n, K, umean);
  double* cc = ctx->buf<double>("as_cc", n * m);
  double* cf = ctx->buf<double>("as_cf", n * mc);
  k_build_design<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(d_xs, d_u, umean, n, (int)d, (int)K, cc, m, cf, mc);
  SLFMM_HIP(hipGetLastError());
  double* Gfull = ctx->buf<double>("as_Gfull", mc * mc);
  ts_gram(ctx, cf, mc, mc, cf, mc, mc, n, Gfull);
  double* Ginv = ctx->buf<double>("as_Ginv", mc * mc);
  int* st_ginv = ctx->check_slot(
      "Warning: C^T C  where C = [1 | X | U_hat] is singular, adding epsilon diagonal regularization",
      SLFMM_E_SINGULAR, "C^T C inversion failed even with epsilon regularization");
  safe_inv_dev(ctx, Gfull, mc, Ginv, st_ginv);
  double* UtU = ctx->buf<double>("as_UtU", std::max<uint64_t>(K * K, 1));
  double* UtUinv = ctx->buf<double>("as_UtUinv", std::max<uint64_t>(K * K, 1));
  if (K) {
    k_extract_block<<<(unsigned)cdiv(K * K, 256), 256, 0, s>>>(Gfull, (int)mc, (int)(1 + d), (int)(1 + d), (int)K,
                                                               (int)K, UtU);
    safe_inv_dev(ctx, UtU, K, UtUinv,
                 ctx->check_slot("Warning: U_hat^T U_hat is singular, adding epsilon diagonal regularization",
                                 SLFMM_E_SINGULAR, "U_hat^T U_hat inversion failed even with epsilon regularization"));
  }
  double* XtX = ctx->buf<double>("as_XtX", d * d);
  double* XtXinv = ctx->buf<double>("as_XtXinv", d * d);
  k_extract_block<<<(unsigned)cdiv(d * d, 256), 256, 0, s>>>(Gfull, (int)mc, 1, 1, (int)d, (int)d, XtX);
  safe_inv_dev(ctx, XtX, d, XtXinv,
               ctx->check_slot("Warning: X^T X for sequential R^2 is singular, adding epsilon diagonal regularization",
                               SLFMM_E_SINGULAR, "X^T X inversion failed even with epsilon regularization"));
  // Orthonormal basis of [U_c | X] = Q T (CholeskyQR2); the plan falls back
  // to [U_c|X] itself if that fails or if C^T C needed regularisation.
  pl.Q = ctx->buf<double>("as_Q", n * m);
  double* T = ctx->buf<double>("as_T", m * m);
  copy2d(s, cc, m, pl.Q, m, n, m);
  int* st_cq = ctx->check_slot("", SLFMM_OK, "");
  cholqr2(ctx, pl.Q, m, m, n, T, st_cq);
  k_select_basis<<<(unsigned)std::min<uint64_t>(cdiv(n * m, 256), 2048), 256, 0, s>>>(st_ginv, st_cq, cc, pl.Q, n * m);
  pl.plan = ctx->buf<double>("as_plan", plan_len((int)d, (int)m));
  double* work = ctx->buf<double>("as_work", mc * m + m * m);
  k_plan<<<1, 256, 0, s>>>(Gfull, Ginv, UtUinv, XtXinv, d_ridge, T, (int)d, (int)K, st_ginv, st_cq, pl.plan, work);
  SLFMM_HIP(hipGetLastError());
  return pl;
}

void launch_func_i8(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_Q, uint64_t m, const EpiArgs& ea,
                    bool fused_epilogue);   // assoc_i8.hip

void assoc_run(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_xs, uint64_t d, const double* d_u,
               uint64_t k, double lambda, const slfmm_config& cfg, double* d_beta, double* d_t,
               double* d_r2, double* d_p, uint64_t* d_hist, double* d_gif_pt, double* d_gif, const NmfAssoc* nmf) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n, p = g->p;
  SLFMM_CHECK(n > 1 + d + k, SLFMM_E_DF,
              "Insufficient degrees of freedom: n=" + std::to_string(n) + " samples but 1+d+K=1+" +
                  std::to_string(d) + "+" + std::to_string(k) + "=" + std::to_string(1 + d + k) +
                  ". Need n > 1 + d + K for valid t-tests. Reduce K or add more samples.");
  SLFMM_CHECK(lambda > 0.0, SLFMM_E_LAMBDA, "Ridge penalty λ must be positive");
  SLFMM_CHECK(d >= 1 && d <= (uint64_t)kMaxD, SLFMM_E_ARG, "need 1 <= d <= 16 covariates");
  // ridge_inv = (X^T X + lambda I)^{-1}  (src/precompute.rs:73-78)
  double* xtx = ctx->buf<double>("as_xtx", d * d);
  double* xtxr = ctx->buf<double>("as_xtxr", d * d);
  ts_gram(ctx, d_xs, d, d, d_xs, d, d, n, xtx);
  add_diag(s, xtx, d, lambda, xtxr);
  double* ridge = ctx->buf<double>("as_ridge", d * d);
  safe_inv_dev(ctx, xtxr, d, ridge,
               ctx->check_slot("Warning: X^T X + λI (ridge) is singular, adding epsilon diagonal regularization",
                               SLFMM_E_SINGULAR, "X^T X + λI (ridge) inversion failed even with epsilon regularization"));

  AssocPlan pl = make_plan(ctx, n, d_xs, d, d_u, k, ridge);
  const uint64_t m = pl.m;
  const bool with_nmf = nmf && nmf->W;
  // the int8 kernels carry at most 32 functionals (K + d); beyond that the FP64 kernels run.
  // With an NMF fill the functionals come from the filled int8 pass where its
  // plans reach (nmf_i8), and the counts, moments and epilogue stay FP64.
  const bool i8 = cfg.precision == SLFMM_PREC_I8 && m <= 32 && !with_nmf;
  const bool nmf_i8 = cfg.precision == SLFMM_PREC_I8 && with_nmf && fill_functionals_i8_supported((int)m, nmf->K);
  if (cfg.precision == SLFMM_PREC_I8 && !with_nmf && !i8)
    fp64_stage(ctx, "association", "K + d = " + std::to_string(m) + " > 32 columns");
  if (cfg.precision == SLFMM_PREC_I8 && with_nmf && !nmf_i8)
    fp64_stage(ctx, "association over the NMF-filled rows",
               "K + d = " + std::to_string(m) + " with nmf-k = " + std::to_string(nmf->K) +
                   " is past the int8 fill plans");
  // per-SNP integer counts of the FP64 kernels (the int8 kernels count in registers)
  int32_t* nmiss = i8 ? nullptr : ctx->buf<int32_t>("as_nmiss", p);
  int32_t* sumg = i8 ? nullptr : ctx->buf<int32_t>("as_sumg", p);
  int32_t* nhom = i8 ? nullptr : ctx->buf<int32_t>("as_nhom", p);
  EpiArgs ea;
  ea.n = n; ea.p = p; ea.d = (int)d; ea.K = (int)k; ea.m = (int)m;
  ea.norm = cfg.norm; ea.df = (double)(n - 1 - d - k); ea.plan = pl.plan;
  ea.n_miss = nmiss; ea.sum_g = sumg; ea.n_hom = nhom;
  ea.beta = d_beta; ea.tstat = d_t; ea.r2 = d_r2;
  SLFMM_CHECK(!(with_nmf && g->host), SLFMM_E_ARG, "NMF imputation needs a device-resident genotype panel");
  // FP64 operand: the basis padded to whole row blocks (shared by every row range)
  const uint64_t npad = rup(n, 64);
  const int nrb = (int)cdiv(m, kRowBlk);
  const uint64_t lda = (uint64_t)nrb * kRowBlk;
  double* A = nullptr;
  double* mu = nullptr;
  double* sc = nullptr;
  if (!i8) {
    A = ctx->buf<double>("as_A64", npad * lda);
    k_pad_operand<<<(unsigned)cdiv(npad * lda, 256), 256, 0, s>>>(pl.Q, m, n, (int)m, A, lda, npad);
    mu = ctx->buf<double>("as_mu", p);
    sc = ctx->buf<double>("as_sc", p);
  }
  const size_t shm = plan_len((int)d, (int)m) * sizeof(double);
  auto epilogue = [&](const EpiArgs& e, const double* F) {
    if (m <= 32) k_epilogue<32><<<(unsigned)cdiv(e.p, 128), 128, shm, s>>>(e, F, lda);
    else k_epilogue<64><<<(unsigned)cdiv(e.p, 128), 128, shm, s>>>(e, F, lda);
    SLFMM_HIP(hipGetLastError());
  };
  // Association of the SNP rows [j0, j0 + v.p) held in v (per-SNP outputs offset by j0).
  auto rows_pass = [&](const slfmm_geno& v, uint64_t j0, double* F) {
    if (v.p == 0) return;
    EpiArgs e = ea;
    e.p = v.p;
    e.beta = d_beta + j0 * d; e.tstat = d_t + j0 * d; e.r2 = d_r2 + j0 * 3;
    if (i8) {
      launch_func_i8(ctx, &v, pl.Q, m, e, true);
      return;
    }
    e.n_miss = nmiss + j0; e.sum_g = sumg + j0; e.n_hom = nhom + j0;
    launch_stats(s, v.d, v.ld, n, v.p, nmiss + j0, sumg + j0, nhom + j0);
    launch_mean_scale(s, nmiss + j0, sumg + j0, n, v.p, cfg.norm, mu + j0, sc + j0);
    {
      slfmm_timed tm(ctx, "assoc_stream");
      launch_func_f64(s, v.d, v.ld, v.p, npad, A, lda, (int)m, mu + j0, F, lda, nullptr, nullptr);
    }
    epilogue(e, F);
  };
  if (g->host) {
    // Host-resident panel: chunks of rows cross PCIe through two pinned
    // staging buffers into two device buffers on the copy stream while the
    // compute stream works on the previous chunk.  Chunk c uses buffer c & 1:
    //   copy stream:    [wait consumed(b)] copy(c) -> copied(b)
    //   compute stream: [wait copied(b)] clear tails, association -> consumed(b)
    //   host:           before refilling staging buffer b, wait copied(b).
    ctx->ensure_copy_stream();
    const uint64_t cap = std::max<uint64_t>(1, std::min(host_chunk_rows(g), p));
    const uint64_t cbytes = cap * g->ld + kSlack;
    uint8_t* cb[2] = {ctx->buf<uint8_t>("as_chunk0", cbytes), ctx->buf<uint8_t>("as_chunk1", cbytes)};
    double* F = i8 ? nullptr : ctx->buf<double>("as_F", cap * lda);
    for (int b = 0; b < 2; ++b) {
      SLFMM_HIP(hipMemsetAsync(cb[b], 0, cbytes, s));   // row bytes past ceil(n/4) stay zero
      SLFMM_HIP(hipEventRecord(ctx->ev_consumed[b], s));
    }
    const uint64_t nchunks = cdiv(p, cap);
    auto upload = [&](uint64_t c) {
      const int b = (int)(c & 1);
      const uint64_t j0 = c * cap, j1 = std::min(p, j0 + cap);
      SLFMM_HIP(hipEventSynchronize(ctx->ev_copied[b]));   // staging buffer b is free
      SLFMM_HIP(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_consumed[b], 0));
      host_chunk_upload(ctx, g, j0, j1, b, cb[b]);
    };
    if (nchunks) upload(0);
    for (uint64_t c = 0; c < nchunks; ++c) {
      if (c + 1 < nchunks) upload(c + 1);
      const int b = (int)(c & 1);
      const uint64_t j0 = c * cap, pc = std::min(p, j0 + cap) - j0;
      SLFMM_HIP(hipStreamWaitEvent(s, ctx->ev_copied[b], 0));
      geno_clear_tail(s, cb[b], g->ld, pc, n);
      if (pc < cap) SLFMM_HIP(hipMemsetAsync(cb[b] + pc * g->ld, 0, kSlack, s));
      slfmm_geno v = *g;
      v.host = nullptr; v.map_base = nullptr; v.d = cb[b]; v.p = pc; v.owned = false;
      rows_pass(v, j0, F);
      SLFMM_HIP(hipEventRecord(ctx->ev_consumed[b], s));
    }
    for (int b = 0; b < 2; ++b) SLFMM_HIP(hipEventSynchronize(ctx->ev_copied[b]));
  } else if (!with_nmf) {
    rows_pass(*g, 0, i8 ? nullptr : ctx->buf<double>("as_F", p * lda));
  } else {
    launch_stats(s, g->d, g->ld, n, p, nmiss, sumg, nhom);
    launch_mean_scale(s, nmiss, sumg, n, p, cfg.norm, mu, sc);
    // NMF on-the-fly imputation (src/parallel.rs:252-282): H_j = max(0, W^+ y_j)
    // of the mean-filled raw column, missing calls filled with W_k . H_j
    FillSpec fs;
    const int K = nmf->K;
    const uint64_t ldh = rup(K, kRowBlk);
    double* Hraw = ctx->buf<double>("as_nmf_hraw", p * ldh);
    double* Ht = ctx->buf<double>("as_nmf_ht", p * K);
    if (cfg.precision == SLFMM_PREC_I8) raw_functionals_i8(ctx, g->d, g->ld, n, p, mu, nmf->Wp, K, Hraw, ldh);
    else raw_functionals(ctx, g->d, g->ld, n, p, mu, nmf->Wp, K, K, Hraw, ldh);
    relu_rows(s, Hraw, ldh, p, K, Ht);
    fs.W = nmf->W; fs.Ht = Ht; fs.K = K; fs.n = n;
    double* fsc = ctx->buf<double>("as_nmf_sc", p);
    double* ftss = ctx->buf<double>("as_nmf_tss", p);
    double* fmu = ctx->buf<double>("as_nmf_mu", p);
    const bool cv = nmf->cv_rate > 0.0 && nmf->cv_out;
    double* SWp = cv ? ctx->buf<double>("as_nmf_swp", p * K) : nullptr;
    fill_moments(ctx, g->d, g->ld, p, fs, cfg.norm, cv ? nmf->Wp : nullptr, fmu, fsc, ftss, SWp,
                 cfg.precision == SLFMM_PREC_I8);
    if (cv)
      gwas_cv(ctx, g->d, g->ld, n, p, nmf->W, nmf->Wp, K, Hraw, ldh, SWp, nmiss, sumg, nmf->cv_seed, nmf->chunk,
              nmf->cv_rate, nmf->cv_out);
    EpiArgs e = ea;
    e.fill_sc = fsc;
    e.fill_tss = ftss;
    double* F = ctx->buf<double>("as_F", p * lda);
    {
      slfmm_timed tm(ctx, "assoc_stream");
      if (nmf_i8) fill_functionals_i8(ctx, g->d, g->ld, n, p, pl.Q, m, (int)m, nmf->W, Ht, K, fmu, nullptr, F, lda);
      else launch_func_f64(s, g->d, g->ld, p, npad, A, lda, (int)m, fmu, F, lda, nullptr, &fs);
    }
    epilogue(e, F);
  }
  // histogram, GIF, calibration
  unsigned long long* hist = d_hist ? (unsigned long long*)d_hist
                                    : ctx->buf<unsigned long long>("as_hist", d * (uint64_t)kBins);
  unsigned long long* ctot = ctx->buf<unsigned long long>("as_hist_chunks", d * (uint64_t)kHistChunks);
  launch_fill(s, reinterpret_cast<double*>(hist), 0.0, d * (uint64_t)kBins);   // zero bits
  const int LB = (int)std::min<uint64_t>(16384, 32768 / d);   // LDS bins per trait (128 KB)
  const uint64_t nblk = std::max<uint64_t>(1, std::min<uint64_t>(256, cdiv(p, 4096)));
  const uint64_t spb = std::max<uint64_t>(1, cdiv(p, nblk));
  uint32_t* part = ctx->buf<uint32_t>("as_hist_part", nblk * d * LB);
  SLFMM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_hist_part),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(d * LB * sizeof(uint32_t))));
  k_hist_part<<<(unsigned)nblk, 1024, d * LB * sizeof(uint32_t), s>>>(d_t, p, (int)d, spb, LB, part, hist);
  const int zs = 16, rows_z = (int)cdiv(nblk, zs);
  k_hist_sum<<<dim3((unsigned)cdiv(LB, 256), (unsigned)d, (unsigned)cdiv(nblk, rows_z)), 256, 0, s>>>(
      part, (int)nblk, (int)d, LB, rows_z, hist);
  k_hist_chunks<<<dim3(kHistChunks, (unsigned)d), 256, 0, s>>>(hist, ctot);
  k_gif<<<1, 256, 0, s>>>(hist, ctot, (int)d, d_gif_pt, d_gif);
  if (p) k_calibrate<<<(unsigned)cdiv(p, 256), 256, 0, s>>>(d_t, p, (int)d, d_gif_pt, d_p);
  SLFMM_HIP(hipGetLastError());
}

}  // namespace slfmm
