// Dense FP64 helpers that stay on the GPU: tall-skinny Gram / GEMM and
// column reductions with deterministic reductions, and one-workgroup
// solvers for the small matrices of the pipeline (m <= 64): Cholesky
// (CholeskyQR2 of the sketch), LU inverse with the reference's safe_inv
// fallback, and one-sided Jacobi (small SVD / symmetric eigendecomposition).
// Nothing here synchronises with the host: solver outcomes go to device
// status words that the caller reads once per API call.
#include "internal.h"
#include "colreduce.h"

namespace slfmm {

constexpr int kSm = 64;   // max order of the one-workgroup solvers

__global__ void k_fill(double* p, double v, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void launch_fill(hipStream_t s, double* p, double v, uint64_t count) {
  if (!count) return;
  k_fill<<<(unsigned)std::min<uint64_t>(cdiv(count, 256), 4096), 256, 0, s>>>(p, v, count);
  SLFMM_HIP(hipGetLastError());
}

// ---- copies ----------------------------------------------------------------
__global__ void k_copy2d(const double* __restrict__ src, uint64_t lds, double* __restrict__ dst, uint64_t ldd,
                         uint64_t rows, uint64_t cols, const int* gate) {
  if (gate && *gate != 1) return;
  const uint64_t total = rows * cols;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x)
    dst[(e / cols) * ldd + e % cols] = src[(e / cols) * lds + e % cols];
}

void copy2d(hipStream_t s, const double* src, uint64_t lds, double* dst, uint64_t ldd, uint64_t rows, uint64_t cols,
            const int* gate) {
  if (!rows || !cols) return;
  k_copy2d<<<(unsigned)std::min<uint64_t>(cdiv(rows * cols, 256), 4096), 256, 0, s>>>(src, lds, dst, ldd, rows, cols,
                                                                                       gate);
  SLFMM_HIP(hipGetLastError());
}

__global__ void k_add_diag(const double* __restrict__ In, int m, double lambda, double* __restrict__ out) {
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) out[e] = In[e] + ((e / m == e % m) ? lambda : 0.0);
}

void add_diag(hipStream_t s, const double* In, uint64_t m, double lambda, double* out) {
  k_add_diag<<<1, 256, 0, s>>>(In, (int)m, lambda, out);
  SLFMM_HIP(hipGetLastError());
}

// ---- tall-skinny Gram: out = A^T B ------------------------------------------
// Pass 1: per-workgroup partial Gram of a row range (32-row tiles in LDS).
// Pass 2: one wave per output entry sums the partials in a fixed order.
__global__ void __launch_bounds__(256) k_gram_partial(const double* __restrict__ A, uint64_t lda, int a,
                                                      const double* __restrict__ B, uint64_t ldb, int b,
                                                      uint64_t rows, uint64_t rpb, double* __restrict__ part,
                                                      const int* gate) {
  if (gate && *gate != 1) return;
  __shared__ double sa[32][kSm + 1];
  __shared__ double sb[32][kSm + 1];
  const int t = threadIdx.x;
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
  const uint64_t r1 = std::min(rows, r0 + rpb);
  const int nout = a * b;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (uint64_t rb = r0; rb < r1; rb += 32) {
    const int nr = (int)std::min<uint64_t>(32, r1 - rb);
    __syncthreads();
    {   // 4 rows x 64 columns per pass (a, b <= 64)
      const int c = t & 63;
      for (int rr = t >> 6; rr < 32; rr += 4) {
        if (c < a) sa[rr][c] = rr < nr ? A[(rb + rr) * lda + c] : 0.0;
        if (c < b) sb[rr][c] = rr < nr ? B[(rb + rr) * ldb + c] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      int e = t + 256 * q;
      if (e < nout) {
        int i = e / b, j = e % b;
        double s = acc[q];
        for (int rr = 0; rr < nr; ++rr) s += sa[rr][i] * sb[rr][j];
        acc[q] = s;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    int e = t + 256 * q;
    if (e < nout) part[(uint64_t)blockIdx.x * nout + e] = acc[q];
  }
}

// Same partial Gram for a, b <= 32 with register blocking: each wave takes
// every fourth row of a 32-row tile, each lane a 4 x 4 block of the 32 x 32
// output (8 LDS reads per 16 FMAs), and the four waves' sums are added in a
// fixed order at the end.  The next tile's loads are issued into registers
// before the current tile's products, so their latency overlaps the
// arithmetic; with B the same matrix as A (same pointer, stride and width)
// only A is loaded.
__global__ void __launch_bounds__(256) k_gram_partial32(const double* __restrict__ A, uint64_t lda, int a,
                                                        const double* __restrict__ B, uint64_t ldb, int b,
                                                        uint64_t rows, uint64_t rpb, double* __restrict__ part,
                                                        const int* gate) {
  if (gate && *gate != 1) return;
  __shared__ __align__(16) double sa[32][32];
  __shared__ __align__(16) double sb[32][32];
  __shared__ double red[4][32][33];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, ti = lane >> 3, tj = lane & 7;
  const bool same = A == B && lda == ldb && a == b;
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
  const uint64_t r1 = std::min(rows, r0 + rpb);
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
  // thread t stages column t & 31 of rows (t >> 5) + 8k, k < 4; columns
  // past a (b) and rows past the range are zero
  const int c = t & 31, rq = t >> 5;
  double va[4], vb[4];
  auto fetch = [&](uint64_t rb) {
    const int nr = (int)std::min<uint64_t>(32, r1 - rb);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int rr = rq + 8 * k;
      va[k] = (rr < nr && c < a) ? A[(rb + rr) * lda + c] : 0.0;
      vb[k] = (!same && rr < nr && c < b) ? B[(rb + rr) * ldb + c] : 0.0;
    }
  };
  if (r0 < r1) fetch(r0);
  for (uint64_t rb = r0; rb < r1; rb += 32) {
    const int nr = (int)std::min<uint64_t>(32, r1 - rb);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sa[rq + 8 * k][c] = va[k];
      sb[rq + 8 * k][c] = same ? va[k] : vb[k];
    }
    __syncthreads();
    if (rb + 32 < r1) fetch(rb + 32);
    for (int rr = wave; rr < nr; rr += 4) {
      double av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) { av[x] = sa[rr][4 * ti + x]; bv[x] = sb[rr][4 * tj + x]; }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] += av[x] * bv[y];
    }
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) red[wave][4 * ti + x][4 * tj + y] = acc[x][y];
  __syncthreads();
  const int nout = a * b;
  for (int e = t; e < nout; e += 256) {
    const int i = e / b, j = e % b;
    part[(uint64_t)blockIdx.x * nout + e] = ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
  }
}

__global__ void __launch_bounds__(256) k_gram_reduce(const double* __restrict__ part, int nblk, int nout,
                                                     double* __restrict__ out, const int* gate) {
  if (gate && *gate != 1) return;
  const int lane = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= nout) return;   // uniform per wave
  double s = 0.0;
  for (int g = lane; g < nblk; g += 64) s += part[(uint64_t)g * nout + e];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[e] = s;
}

void ts_gram(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t a, const double* B,
             uint64_t ldb, uint64_t b, uint64_t rows, double* out, const int* gate) {
  SLFMM_CHECK(a <= (uint64_t)kSm && b <= (uint64_t)kSm, SLFMM_E_ARG, "ts_gram: too many columns");
  hipStream_t s = ctx->stream;
  if (rows == 0) { launch_fill(s, out, 0.0, a * b); return; }
  // one 32-row tile per block up to tuning.gram_max_blocks blocks
  uint64_t nblk = std::min<uint64_t>(cdiv(rows, 32), std::max<uint32_t>(1u, ctx->tuning.gram_max_blocks));
  uint64_t rpb = cdiv(rows, nblk);
  nblk = cdiv(rows, rpb);
  double* part = ctx->buf<double>("gram_part", nblk * a * b);
  if (a <= 32 && b <= 32)
    k_gram_partial32<<<(unsigned)nblk, 256, 0, s>>>(A, lda, (int)a, B, ldb, (int)b, rows, rpb, part, gate);
  else
    k_gram_partial<<<(unsigned)nblk, 256, 0, s>>>(A, lda, (int)a, B, ldb, (int)b, rows, rpb, part, gate);
  k_gram_reduce<<<(unsigned)cdiv(a * b, 4), 256, 0, s>>>(part, (int)nblk, (int)(a * b), out, gate);
  SLFMM_HIP(hipGetLastError());
}

// ---- tall-skinny GEMM: out = alpha A M + beta out -------------------------
// Each pass of a workgroup stages rpb = 256 / cw rows of A in LDS with
// coalesced loads; thread (row, column) then forms its dot product from LDS.
__global__ void __launch_bounds__(256) k_ts_gemm(const double* __restrict__ A, uint64_t lda, int m,
                                                 const double* __restrict__ M, uint64_t ldm, int q,
                                                 uint64_t rows, double alpha, double beta, double* out,
                                                 uint64_t ldo, const int* gate, const double* dlam, int inverse) {
  if (gate && *gate != 1) return;
  __shared__ double sm[kSm * kSm];
  __shared__ double sa[8][kSm + 1];
  for (int e = threadIdx.x; e < m * q; e += 256) {
    double v = M[(uint64_t)(e / q) * ldm + (e % q)];
    if (dlam) {   // row t of M scaled by -(1 - d_t), or 1 / d_t - 1 for the inverse
      const double dl = dlam[e / q];
      v *= inverse ? (1.0 / dl - 1.0) : -(1.0 - dl);
    }
    sm[e] = v;
  }
  // thread = (row, column) on a 256 / cw x cw grid, cw = 32 or 64 >= q
  const int cw = q <= 32 ? 32 : 64;
  const int c = threadIdx.x & (cw - 1);
  const int rpb = 256 / cw;
  const int rr = threadIdx.x / cw;
  for (uint64_t r0 = (uint64_t)blockIdx.x * rpb; r0 < rows; r0 += (uint64_t)gridDim.x * rpb) {
    __syncthreads();
    for (int e = threadIdx.x; e < rpb * m; e += 256) {
      const int i = e / m, t = e % m;
      sa[i][t] = r0 + i < rows ? A[(r0 + i) * lda + t] : 0.0;
    }
    __syncthreads();
    const uint64_t r = r0 + rr;
    if (c < q && r < rows) {
      double s = 0.0;
      for (int t = 0; t < m; ++t) s += sa[rr][t] * sm[t * q + c];
      double v = alpha * s;
      if (beta != 0.0) v += beta * out[r * ldo + c];
      out[r * ldo + c] = v;
    }
  }
}

// Same product for m, q <= 32: column c of M is held in registers by the
// threads of column c, and a pass of a workgroup stages 64 rows of A in LDS
// (eight coalesced loads per thread) and forms eight rows of the output per
// thread.  Each output element is the same dot product, summed in the same
// order (t = 0, 1, ...), as k_ts_gemm's; the terms past m are zero products
// of the zero-padded operands, so no branch on m splits the LDS reads.
constexpr int kTsRows = 64;
__global__ void __launch_bounds__(256) k_ts_gemm32(const double* __restrict__ A, uint64_t lda, int m,
                                                   const double* __restrict__ M, uint64_t ldm, int q,
                                                   uint64_t rows, double alpha, double beta, double* out,
                                                   uint64_t ldo, const int* gate, const double* dlam, int inverse) {
  if (gate && *gate != 1) return;
  __shared__ double sa[kTsRows][33];
  __shared__ double sm[32][33];
  const int c = threadIdx.x & 31, rq = threadIdx.x >> 5;
  // M staged through LDS (coalesced, all loads in flight together), then
  // column c into registers
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = rq + 8 * k;
    double v = 0.0;
    if (t < m && c < q) v = M[(uint64_t)t * ldm + c];
    sm[t][c] = v;
  }
  __syncthreads();
  if (dlam && threadIdx.x < 32) {   // row t of M scaled by -(1 - d_t), or 1 / d_t - 1 for the inverse
    const int t = threadIdx.x;
    if (t < m) {
      const double dl = dlam[t];
      const double f = inverse ? (1.0 / dl - 1.0) : -(1.0 - dl);
      for (int cc = 0; cc < q; ++cc) sm[t][cc] *= f;
    }
  }
  __syncthreads();
  double mc[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) mc[t] = sm[t][c];
  for (uint64_t r0 = (uint64_t)blockIdx.x * kTsRows; r0 < rows; r0 += (uint64_t)gridDim.x * kTsRows) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTsRows / 8; ++k) {
      const int i = rq + 8 * k;
      sa[i][c] = (c < m && r0 + i < rows) ? A[(r0 + i) * lda + c] : 0.0;
    }
    __syncthreads();
    if (c < q) {
#pragma unroll 2
      for (int k = 0; k < kTsRows / 8; ++k) {
        const int i = rq + 8 * k;
        const uint64_t r = r0 + i;
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 32; ++t) s += sa[i][t] * mc[t];   // zero past m on both sides
        if (r < rows) {
          double v = alpha * s;
          if (beta != 0.0) v += beta * out[r * ldo + c];
          out[r * ldo + c] = v;
        }
      }
    }
  }
}

void ts_gemm(hipStream_t s, const double* A, uint64_t lda, uint64_t m, const double* M,
             uint64_t ldm, uint64_t q, uint64_t rows, double alpha, double beta, double* out,
             uint64_t ldo, const int* gate, const double* dlam, bool inverse) {
  if (rows == 0 || q == 0) return;
  SLFMM_CHECK(m <= (uint64_t)kSm && q <= (uint64_t)kSm, SLFMM_E_ARG, "ts_gemm: too many columns");
  if (m <= 32 && q <= 32 && rows > 32768) {   // measured: below this the one-row-per-thread kernel is faster
    const uint64_t blocks = std::min<uint64_t>(cdiv(rows, kTsRows), 4096);
    k_ts_gemm32<<<(unsigned)blocks, 256, 0, s>>>(A, lda, (int)m, M, ldm, (int)q, rows, alpha, beta, out, ldo, gate,
                                                 dlam, inverse ? 1 : 0);
    SLFMM_HIP(hipGetLastError());
    return;
  }
  const uint64_t rpb = q <= 32 ? 8 : 4;
  uint64_t blocks = std::min<uint64_t>(cdiv(rows, rpb), 1024);   // M is staged once per workgroup
  k_ts_gemm<<<(unsigned)blocks, 256, 0, s>>>(A, lda, (int)m, M, ldm, (int)q, rows, alpha, beta, out, ldo, gate, dlam,
                                             inverse ? 1 : 0);
  SLFMM_HIP(hipGetLastError());
}

// ---- column reductions --------------------------------------------------------
namespace {
struct FnVal {
  const double* A; uint64_t lda;
  __device__ void operator()(uint64_t r, int c, double (&v)[1]) const { v[0] = A[r * lda + c]; }
};
struct GMean {
  double* out; uint64_t rows;
  __device__ void operator()(int c, const double (&v)[1]) const { out[c] = rows ? v[0] / (double)rows : 0.0; }
};
template <int NV>
struct FnAbsScaled {
  const double* A; uint64_t lda; const double* rs; const double* rm2;
  __device__ void operator()(uint64_t r, int c, double (&v)[NV]) const {
    double x = A[r * lda + c];
    if (rs) x *= rs[r];
    v[0] = fabs(x);
    if constexpr (NV == 2) v[1] = fabs(x * rm2[r]);
  }
};
template <int NV>
struct GScale127 {
  double* out; int m;
  __device__ void operator()(int c, const double (&v)[NV]) const {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k * m + c] = v[k] > 0.0 ? v[k] / 127.0 : 1.0;
  }
};
}  // namespace

void colmeans(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t rows, uint64_t m, double* means) {
  colreduce<1, kColSum>(ctx, FnVal{A, lda}, GMean{means, rows}, rows, (int)m);
}

void colabsmax_scale(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t rows, uint64_t m, const double* rs,
                     const double* rm2, double* out) {
  if (rm2) colreduce<2, kColMax>(ctx, FnAbsScaled<2>{A, lda, rs, rm2}, GScale127<2>{out, (int)m}, rows, (int)m);
  else colreduce<1, kColMax>(ctx, FnAbsScaled<1>{A, lda, rs, nullptr}, GScale127<1>{out, (int)m}, rows, (int)m);
}

// ---- one-workgroup Cholesky:  G + shift*I = R^T R,  outputs R and R^{-1} ---
// Attempt 0 is plain Cholesky with a relative pivot test of 1e-13 against
// the original diagonal.  If it breaks down and allow_shift is set, it is
// redone on G + s I with s = 11 (rows m + m (m + 1)) u tr(G) (shifted
// CholeskyQR).  R^{-1} is formed in place: its strictly upper part, stored
// transposed, fills the strictly lower triangle of the factor.
// Status: mode 0 writes 0 (plain), 1 (shifted) or 2 (failed); mode 1
// writes only 2, on failure.
__global__ void __launch_bounds__(256) k_chol(const double* __restrict__ G, int m, double rows, int allow_shift,
                                              double* R, double* Rinv, int* st, int mode, const int* gate) {
  if (gate && *gate != 1) return;
  extern __shared__ double A[];   // m*m
  __shared__ double diag0[kSm];
  __shared__ double rdiag[kSm];
  __shared__ double dinv[kSm];
  __shared__ double shift;
  const int t = threadIdx.x;
  if (t == 0) shift = 0.0;
  int attempt = 0, fail = 0;
  for (;; ++attempt) {
    __syncthreads();
    const double sh = shift;
    for (int e = t; e < m * m; e += 256) A[e] = G[e] + ((e / m == e % m) ? sh : 0.0);
    __syncthreads();
    if (t < m) diag0[t] = A[t * m + t];
    __syncthreads();
    const double tol = attempt == 0 ? 1e-13 : 0.0;
    fail = 0;
    for (int j = 0; j < m; ++j) {
      // every thread reads the pivot and takes the same decision; the
      // diagonal goes to rdiag so that A[j][j] is not written while read
      const double dj = A[j * m + j];
      if (!(dj > tol * diag0[j]) || !(dj < 1e308)) {
        fail = j + 1;
        break;
      }
      const double rjj = sqrt(dj);
      for (int i = j + 1 + t; i < m; i += 256) A[j * m + i] /= rjj;
      if (t == 0) rdiag[j] = rjj;
      __syncthreads();
      // trailing update on a 16 x 16 thread grid (no division by the runtime width)
      for (int i = j + 1 + (t >> 4); i < m; i += 16)
        for (int l = j + 1 + (t & 15); l < m; l += 16)
          if (l >= i) A[i * m + l] -= A[j * m + i] * A[j * m + l];
      __syncthreads();
    }
    if (!fail || attempt == 1 || !allow_shift) break;
    if (t == 0) {
      double tr = 0.0;
      for (int i = 0; i < m; ++i) tr += G[i * m + i];
      shift = 11.0 * (rows * m + (double)m * (m + 1)) * 1.1102230246251565e-16 * tr;
    }
  }
  if (fail) {
    if (t == 0) *st = 2;
    return;
  }
  if (t < m) {
    A[t * m + t] = rdiag[t];
    dinv[t] = 1.0 / rdiag[t];
  }
  __syncthreads();
  // Column c of R^{-1} by back substitution, 8 lanes per column:
  //   X[i][c] = -(sum_{i<k<=c} R[i][k] X[k][c]) / R[i][i],  X[c][c] = 1 / R[c][c],
  // lane `sub` summing the terms k = i+1+sub, i+9+sub, ...  X[k][c] is kept at
  // A[c][k], the strictly lower triangle, which the factor does not use.  The
  // 8 lanes of a column are in one wave and all hold the reduced sum.
  {
    const int grp = t >> 3, sub = t & 7;
    for (int c = grp; c < m; c += 32) {
      const double xcc = dinv[c];
      for (int i = c - 1; i >= 0; --i) {
        double s = 0.0;
        for (int k = i + 1 + sub; k <= c; k += 8) s += A[i * m + k] * (k == c ? xcc : A[c * m + k]);
        s += __shfl_xor(s, 4, 8);
        s += __shfl_xor(s, 2, 8);
        s += __shfl_xor(s, 1, 8);
        if (sub == 0) A[c * m + i] = -s * dinv[i];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // the group's next reads see X[i][c]
      }
    }
  }
  __syncthreads();
  for (int e = t; e < m * m; e += 256) {
    int i = e / m, l = e % m;
    R[e] = (l >= i) ? A[e] : 0.0;
    Rinv[e] = (l > i) ? A[l * m + i] : (l == i ? dinv[i] : 0.0);
  }
  if (t == 0 && mode == 0) *st = attempt;
}

__device__ __forceinline__ double lane_value(double v, int lane) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane), hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Orders one wave's LDS writes before its later LDS reads of other lanes'
// values (and the compiler's accesses around it).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same factorisation, shift rule, outputs and status words as k_chol
// for m <= 32, on one wave: lane l keeps column l of the working matrix in
// registers, padded to 32 x 32 with an identity block past m (pivots 1,
// zero couplings: the steps past m change nothing, and no step needs a
// branch on m).  Step j reads its pivot from lane j (readlane into scalar
// registers) and broadcasts row j of R through LDS (one write per lane,
// every lane reads the same values), so the trailing update is independent
// FMAs per lane.  The kernel runs 32 threads (one wave, lanes 32..63
// inactive).  The arithmetic of every element is k_chol's: R[j][l] =
// A[j][l] / r_jj, then A[i][l] -= R[j][i] R[j][l] for j < i <= l.  lds:
// 5 x 32 doubles (rows of R in two slots used by alternate steps, then the
// pivot references).
// Returns 0, or j + 1 for a failed pivot j; column l of R is left in col.
__device__ __forceinline__ int chol32_factor(const double* __restrict__ G, int m, double shift, double tol,
                                             double (&col)[32], double* lds) {
  // the lane index through an opaque move: the caller's retry loop then
  // keeps the per-step lane masks inside the loop instead of hoisting 64 of
  // them into scalar registers
  int l;
  asm volatile("v_mov_b32 %0, %1" : "=v"(l) : "v"((int)threadIdx.x));
  const bool lv = l < m;
#pragma unroll
  for (int i = 0; i < 32; ++i)
    col[i] = (lv && i < m) ? G[i * m + l] + (i == l ? shift : 0.0) : (i == l ? 1.0 : 0.0);
  double* d0s = lds + 4 * 32;   // pivot references (the padded diagonal reads 1)
  d0s[l] = lv ? G[l * m + l] + shift : 1.0;
  wave_lds_sync();
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    double* rrow = lds + 64 * (j & 1) + 32;
    const double dj = lane_value(col[j], j), d0 = d0s[j];   // pivot: lane j's diagonal element
    if (!(dj > tol * d0) || !(dj < 1e308)) return j + 1;   // wave-uniform
    const double rjj = sqrt(dj);
    const double r = l > j ? col[j] / rjj : (l == j ? rjj : 0.0);   // R[j][l]
    col[j] = r;
    rrow[l] = r;
    wave_lds_sync();
#pragma unroll
    for (int i = j + 1; i < 32; ++i) col[i] -= rrow[i] * r;
  }
  return 0;
}

__global__ void __launch_bounds__(32) k_chol32(const double* __restrict__ G, int m, double rows, int allow_shift,
                                               double* R, double* Rinv, int* st, int mode, const int* gate) {
  if (gate && *gate != 1) return;
  __shared__ double rb[5 * 32];
  __shared__ double Rs[32][33];   // R, row-major (padded rows)
  __shared__ double dinv_s[32];
  const int l = threadIdx.x;
  double col[32];
  int attempt = 0, fail = 0;
  double shift = 0.0, tol = 1e-13;
#pragma unroll 1
  for (;; ++attempt) {   // one copy of the factorisation code: attempt 1 is the shifted retry
    fail = chol32_factor(G, m, shift, tol, col, rb);
    if (!fail || attempt == 1 || !allow_shift) break;
    double tr = 0.0;
    for (int i = 0; i < m; ++i) tr += G[i * m + i];
    shift = 11.0 * (rows * m + (double)m * (m + 1)) * 1.1102230246251565e-16 * tr;
    tol = 0.0;
    wave_lds_sync();   // the first attempt's last reads of rb come before this attempt's writes
  }
  if (fail) {
    if (l == 0) *st = 2;
    return;
  }
  // R^{-1}: R goes to LDS (lane l writes column l), the reciprocal
  // diagonal is formed lane-parallel, and lane c forms column c of X by back
  // substitution, X[i] = -(sum_{k > i} R[i][k] X[k]) / R[i][i], reading the
  // rows of R as broadcasts.
#pragma unroll
  for (int i = 0; i < 32; ++i) Rs[i][l] = col[i];
  wave_lds_sync();
  dinv_s[l] = 1.0 / Rs[l][l];   // 1 on the padded diagonal
  wave_lds_sync();
  double X[32];
#pragma unroll
  for (int i = 31; i >= 0; --i) {
    double s = 0.0;
#pragma unroll
    for (int k = i + 1; k < 32; ++k) s += Rs[i][k] * X[k];   // X[k] = 0 past the column
    const double dinv = dinv_s[i];
    X[i] = i < l ? -s * dinv : (i == l ? dinv : 0.0);
  }
  if (l < m) {
#pragma unroll
    for (int i = 0; i < 32; ++i)
      if (i < m) {
        R[i * m + l] = i <= l ? col[i] : 0.0;
        Rinv[i * m + l] = X[i];
      }
    if (l == 0 && mode == 0) *st = attempt;
  }
}

static void launch_chol(hipStream_t s, const double* G, uint64_t m, uint64_t rows, bool allow_shift, double* R,
                        double* Rinv, int* st, int mode, const int* gate) {
  if (m <= 32)
    k_chol32<<<1, 32, 0, s>>>(G, (int)m, (double)rows, allow_shift ? 1 : 0, R, Rinv, st, mode, gate);
  else
    k_chol<<<1, 256, m * m * sizeof(double), s>>>(G, (int)m, (double)rows, allow_shift ? 1 : 0, R, Rinv, st, mode,
                                                  gate);
  SLFMM_HIP(hipGetLastError());
}

// With shift_all every pass may take the shift: a rank-deficient A then
// yields A = Q R with Q orthonormal to O(shift / s_i^2) in the directions of
// its singular values s_i (exact where s_i^2 >> shift), instead of failing.
void cholqr2(slfmm_ctx* ctx, double* A, uint64_t lda, uint64_t m, uint64_t rows, double* d_R, int* st,
             bool shift_all) {
  SLFMM_CHECK(m >= 1 && m <= (uint64_t)kSm, SLFMM_E_ARG, "cholqr2: 1..64 columns");
  hipStream_t s = ctx->stream;
  double* G = ctx->buf<double>("cqr_G", m * m);
  double* R1 = ctx->buf<double>("cqr_R1", m * m);
  double* R1i = ctx->buf<double>("cqr_R1i", m * m);
  double* R2 = ctx->buf<double>("cqr_R2", m * m);
  double* R2i = ctx->buf<double>("cqr_R2i", m * m);
  double* Rtmp = ctx->buf<double>("cqr_Rtmp", m * m);
  double* Rout = d_R ? d_R : ctx->buf<double>("cqr_Racc", m * m);
  double* T = ctx->buf<double>("cqr_T", rows * m);
  // pass 1 (shifted if plain Cholesky breaks down): A -> T
  ts_gram(ctx, A, lda, m, A, lda, m, rows, G);
  launch_chol(s, G, m, rows, true, R1, R1i, st, 0, nullptr);
  ts_gemm(s, A, lda, m, R1i, m, m, rows, 1.0, 0.0, T, m);
  // pass 2: T -> A
  ts_gram(ctx, T, m, m, T, m, m, rows, G);
  launch_chol(s, G, m, rows, shift_all, R2, R2i, st, 1, nullptr);
  ts_gemm(s, T, m, m, R2i, m, m, rows, 1.0, 0.0, A, lda);
  ts_gemm(s, R2, m, m, R1, m, m, m, 1.0, 0.0, Rout, m);
  // pass 3, only after a shifted first pass (*st == 1)
  ts_gram(ctx, A, lda, m, A, lda, m, rows, G, st);
  launch_chol(s, G, m, rows, shift_all, R2, R2i, st, 1, st);
  ts_gemm(s, A, lda, m, R2i, m, m, rows, 1.0, 0.0, T, m, st);
  copy2d(s, T, m, A, lda, rows, m, st);
  ts_gemm(s, R2, m, m, Rout, m, m, m, 1.0, 0.0, Rtmp, m, st);
  copy2d(s, Rtmp, m, Rout, m, m, m, st);
}

// One CholeskyQR pass (shifted if plain Cholesky breaks down): out = A R^{-1}.
// For the intermediate bases of the power iterations, where only the column
// span is carried forward: the loss of orthogonality, about cond(A)^2 u, does
// not change the span.
void cholqr1(slfmm_ctx* ctx, const double* A, uint64_t lda, uint64_t m, uint64_t rows, double* out, uint64_t ldo,
             int* st) {
  SLFMM_CHECK(m >= 1 && m <= (uint64_t)kSm, SLFMM_E_ARG, "cholqr1: 1..64 columns");
  hipStream_t s = ctx->stream;
  double* G = ctx->buf<double>("cqr_G", m * m);
  double* R1 = ctx->buf<double>("cqr_R1", m * m);
  double* R1i = ctx->buf<double>("cqr_R1i", m * m);
  ts_gram(ctx, A, lda, m, A, lda, m, rows, G);
  launch_chol(s, G, m, rows, true, R1, R1i, st, 0, nullptr);
  ts_gemm(s, A, lda, m, R1i, m, m, rows, 1.0, 0.0, out, ldo);
}

// ---- one-workgroup LU inverse with the safe_inv fallback -------------------
// src/testing.rs:996-1024: invert; if LU meets an exactly zero pivot, add
// eps = 1e-8 * max(max|diag|, 1e-300) to the diagonal and invert again.
__global__ void __launch_bounds__(256) k_safe_inv(const double* __restrict__ In, int m, double* Out, int* info) {
  extern __shared__ double A[];   // m*m LU factors, then Xs (m*m): column t of the inverse at Xs[i * m + t]
  double* Xs = A + m * m;
  __shared__ int piv[kSm];
  __shared__ int sing;
  __shared__ double eps;
  const int t = threadIdx.x;
  for (int attempt = 0; attempt < 2; ++attempt) {
    __syncthreads();
    for (int e = t; e < m * m; e += 256) A[e] = In[e] + ((attempt == 1 && e / m == e % m) ? eps : 0.0);
    if (t == 0) sing = 0;
    __syncthreads();
    for (int j = 0; j < m; ++j) {
      if (t == 0) {
        int p = j;
        double best = fabs(A[j * m + j]);
        for (int i = j + 1; i < m; ++i) {
          double v = fabs(A[i * m + j]);
          if (v > best) { best = v; p = i; }
        }
        piv[j] = p;
        if (A[p * m + j] == 0.0) sing = j + 1;
      }
      __syncthreads();
      if (sing) break;
      const int p = piv[j];
      if (p != j)
        for (int c = t; c < m; c += 256) {
          double tmp = A[j * m + c];
          A[j * m + c] = A[p * m + c];
          A[p * m + c] = tmp;
        }
      __syncthreads();
      const double inv = 1.0 / A[j * m + j];
      for (int i = j + 1 + t; i < m; i += 256) A[i * m + j] *= inv;
      __syncthreads();
      for (int i = j + 1 + (t >> 4); i < m; i += 16)   // 16 x 16 thread grid
        for (int l = j + 1 + (t & 15); l < m; l += 16) A[i * m + l] -= A[i * m + j] * A[j * m + l];
      __syncthreads();
    }
    if (!sing) {
      if (t < m) {   // column t of A^{-1}: P b, forward (unit L), backward (U); the column lives in LDS
        double* x = Xs + t;   // x[i] is Xs[i * m + t]
        for (int i = 0; i < m; ++i) x[i * m] = (i == t) ? 1.0 : 0.0;
        for (int j = 0; j < m; ++j) { int p = piv[j]; double tmp = x[j * m]; x[j * m] = x[p * m]; x[p * m] = tmp; }
        // (the inner loops are unrolled so that their LDS reads are issued
        // ahead of the dependent chain of subtractions; same order of terms)
        for (int i = 0; i < m; ++i) {
          double s = x[i * m];
#pragma unroll 8
          for (int k = 0; k < i; ++k) s -= A[i * m + k] * x[k * m];
          x[i * m] = s;
        }
        for (int i = m - 1; i >= 0; --i) {
          double s = x[i * m];
#pragma unroll 8
          for (int k = i + 1; k < m; ++k) s -= A[i * m + k] * x[k * m];
          x[i * m] = s / A[i * m + i];
        }
        for (int i = 0; i < m; ++i) Out[i * m + t] = x[i * m];
      }
      if (t == 0) *info = attempt;   // 0 = plain, 1 = regularised
      return;
    }
    if (t == 0) {
      double dmax = 0.0;
      for (int i = 0; i < m; ++i) dmax = fmax(dmax, fabs(In[i * m + i]));
      eps = 1e-8 * fmax(dmax, 1e-300);
    }
  }
  if (t == 0) *info = 2;
}

void safe_inv_dev(slfmm_ctx* ctx, const double* d_a, uint64_t m, double* d_out, int* d_st) {
  SLFMM_CHECK(m >= 1 && m <= (uint64_t)kSm, SLFMM_E_ARG, "safe_inv: order must be 1..64");
  const size_t lds = 2 * m * m * sizeof(double);
  if (lds > 48 * 1024)
    SLFMM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_safe_inv),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  k_safe_inv<<<1, 256, lds, ctx->stream>>>(d_a, (int)m, d_out, d_st);
  SLFMM_HIP(hipGetLastError());
}

// A double moved across lanes by a DPP pattern (two 32-bit moves).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum over an aligned group of 8 lanes without LDS: quad_perm swaps of
// lanes l ^ 1 and l ^ 2, then row_half_mirror (lane 7 - l of the group).
// Each step adds two equal-bit partial sums in commuted order, so every
// lane of the group ends with the same bits.  All 8 lanes must be active.
__device__ __forceinline__ double sum8(double x) {
  x += dpp_mov<0xB1>(x);
  x += dpp_mov<0x4E>(x);
  x += dpp_mov<0x141>(x);
  return x;
}

// ---- one-workgroup one-sided Jacobi -----------------------------------------
// Rotates column pairs of A (m x m) until they are mutually orthogonal;
// V accumulates the rotations, so A_in V = U diag(sv).  Pairs of a sweep are
// scheduled by the circle method: m/2 disjoint pairs per step, 8 lanes each.
// A and V are kept column by column in LDS (element (i, c) at c m + i), so
// the 8 lanes of a pair touch consecutive words.
__global__ void __launch_bounds__(256) k_jacobi(const double* __restrict__ In, int m, double* Vout, double* svout) {
  extern __shared__ double A[];   // m*m, then V (m*m)
  double* V = A + m * m;
  __shared__ double sv[kSm];
  __shared__ int order[kSm];
  __shared__ int rotated;
  const int t = threadIdx.x;
  const int mp = (m + 1) & ~1;
  for (int e = t; e < m * m; e += 256) {
    A[(e % m) * m + e / m] = In[e];
    V[e] = (e / m == e % m) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int grp = t >> 3, sub = t & 7;          // 32 groups of 8 lanes
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (t == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < mp - 1; ++r) {
      for (int pi = grp; pi < mp / 2; pi += 32) {
        int pos_a = pi, pos_b = mp - 1 - pi;
        int va = pos_a - 1 + r, vb = pos_b - 1 + r;   // (pos - 1 + r) mod (mp - 1), both terms below mp - 1
        if (va >= mp - 1) va -= mp - 1;
        if (vb >= mp - 1) vb -= mp - 1;
        int a = pos_a == 0 ? 0 : va + 1;
        int b = pos_b == 0 ? 0 : vb + 1;
        bool live = a < m && b < m;
        double al = 0.0, be = 0.0, ga = 0.0;
        if (live)
          for (int i = sub; i < m; i += 8) {
            double x = A[a * m + i], y = A[b * m + i];
            al += x * x; be += y * y; ga += x * y;
          }
        al = sum8(al);
        be = sum8(be);
        ga = sum8(ga);
        if (live && ga != 0.0 && fabs(ga) > 4e-16 * (double)m * sqrt(al * be)) {
          double zeta = (be - al) / (2.0 * ga);
          double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
          for (int i = sub; i < m; i += 8) {
            double x = A[a * m + i], y = A[b * m + i];
            A[a * m + i] = c * x - s * y;
            A[b * m + i] = s * x + c * y;
            x = V[a * m + i]; y = V[b * m + i];
            V[a * m + i] = c * x - s * y;
            V[b * m + i] = s * x + c * y;
          }
          if (sub == 0) rotated = 1;
        }
      }
      __syncthreads();
    }
    int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  if (t < m) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += A[t * m + i] * A[t * m + i];
    sv[t] = sqrt(s);
  }
  __syncthreads();
  if (t < m) {   // descending order: rank of column t (ties keep the column order; NaN sorts last)
    auto key = [&](int j) { const double x = sv[j]; return x != x ? -__builtin_huge_val() : x; };
    const double v = key(t);
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double u = key(j);
      rank += (u > v || (u == v && j < t)) ? 1 : 0;
    }
    order[rank] = t;   // the ranks are a permutation of 0 .. m-1
  }
  __syncthreads();
  for (int e = t; e < m * m; e += 256) {
    int i = e / m, c = e % m;
    Vout[e] = V[order[c] * m + i];
  }
  if (t < m) svout[t] = sv[order[t]];
}

void jacobi_svd_small(slfmm_ctx* ctx, const double* d_a, uint64_t m, double* d_v, double* d_sv) {
  SLFMM_CHECK(m >= 1 && m <= (uint64_t)kSm, SLFMM_E_ARG, "jacobi: order must be 1..64");
  k_jacobi<<<1, 256, 2 * m * m * sizeof(double), ctx->stream>>>(d_a, (int)m, d_v, d_sv);
  SLFMM_HIP(hipGetLastError());
}

}  // namespace slfmm
