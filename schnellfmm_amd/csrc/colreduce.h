// Deterministic column reductions of tall matrices (rows >> columns).
//
// Pass 1: workgroup b reduces a contiguous row range; its 256 threads are
// CW columns x (256 / CW) row lanes, so a wave reads whole matrix rows
// (coalesced) and every lane keeps one running value per column.  The row
// lanes are then combined in a fixed order in LDS.  Pass 2: one wave per
// column combines the per-workgroup partials, lane-strided and then by a
// fixed butterfly, so results do not depend on scheduling.
//
// F: __device__ void operator()(uint64_t row, int col, double (&v)[NV]) const
// G: __device__ void operator()(int col, const double (&v)[NV]) const
#pragma once
#include "internal.h"

namespace slfmm {

enum ColOp { kColSum = 0, kColMax = 1 };

template <int OP>
__device__ __forceinline__ double col_combine(double a, double b) {
  if constexpr (OP == kColSum) return a + b;
  else return fmax(a, b);
}

template <int CW, int NV, int OP, class F>
__global__ void __launch_bounds__(256) k_colpart(F f, uint64_t rows, int m, uint64_t rpb, double* __restrict__ part) {
  constexpr int RL = 256 / CW;
  __shared__ double red[NV][RL][CW];
  const int t = threadIdx.x, c = t % CW, rl = t / CW;
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb;
  const uint64_t r1 = r0 + rpb < rows ? r0 + rpb : rows;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  if (c < m) {
    for (uint64_t r = r0 + rl; r < r1; r += RL) {
      double v[NV];
      f(r, c, v);
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] = col_combine<OP>(acc[k], v[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) red[k][rl][c] = acc[k];
  __syncthreads();
  if (rl == 0 && c < m) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = red[k][0][c];
      for (int q = 1; q < RL; ++q) s = col_combine<OP>(s, red[k][q][c]);
      part[((uint64_t)blockIdx.x * m + c) * NV + k] = s;
    }
  }
}

template <int NV, int OP, class G>
__global__ void __launch_bounds__(256) k_colfin(const double* __restrict__ part, int nblk, int m, G g) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= m) return;   // uniform per wave
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = 0.0;
    for (int b = lane; b < nblk; b += 64) a = col_combine<OP>(a, part[((uint64_t)b * m + c) * NV + k]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = col_combine<OP>(a, __shfl_xor(a, off, 64));
    v[k] = a;
  }
  if (lane == 0) g(c, v);
}

template <int NV, int OP, class F, class G>
void colreduce(slfmm_ctx* ctx, F f, G g, uint64_t rows, int m) {
  SLFMM_CHECK(m >= 1 && m <= 64, SLFMM_E_ARG, "column reduction: 1..64 columns");
  uint64_t nblk = std::min<uint64_t>(std::max<uint64_t>(cdiv(rows, 256), 1), 1024);
  const uint64_t rpb = std::max<uint64_t>(cdiv(rows, nblk), 1);
  nblk = std::max<uint64_t>(cdiv(rows, rpb), 1);
  double* part = ctx->buf<double>("colred_part", nblk * m * NV);
  hipStream_t s = ctx->stream;
  if (m <= 32) k_colpart<32, NV, OP, F><<<(unsigned)nblk, 256, 0, s>>>(f, rows, m, rpb, part);
  else k_colpart<64, NV, OP, F><<<(unsigned)nblk, 256, 0, s>>>(f, rows, m, rpb, part);
  k_colfin<NV, OP, G><<<(unsigned)cdiv(m, 4), 256, 0, s>>>(part, (int)nblk, m, g);
  SLFMM_HIP(hipGetLastError());
}

}  // namespace slfmm
