// Diagnostic: sustained dense MFMA rate of the three operand formats that
// can carry a genotype stream exactly (int8, FP8 e4m3, BF16), measured on
// the whole chip.  The int8 kernels of this library were chosen over FP8 and
// BF16 hi+lo on error per MFMA pass; this measures the pass cost itself.
#include "internal.h"

namespace slfmm {

typedef int v4i_r __attribute__((ext_vector_type(4)));
typedef int v8i_r __attribute__((ext_vector_type(8)));
typedef int v16i_r __attribute__((ext_vector_type(16)));
typedef float v16f_r __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kRateChains = 4;   // independent accumulators per wave

__device__ __forceinline__ uint32_t rate_hash(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}

// KIND 0: v_mfma_i32_32x32x32_i8, 1: v_mfma_scale_f32_32x32x64_f8f6f4 with
// both operands FP8 e4m3 (unit scales), 2: v_mfma_f32_32x32x16_bf16.
// Operand bytes are fixed per lane (small magnitudes: no overflow, finite
// sums); the accumulators are stored so the loop is not removed.
template <int KIND>
__global__ void __launch_bounds__(256) k_mfma_rate(int iters, float* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  float sink = 0.0f;
  if constexpr (KIND == 0) {
    v4i_r a, b;
    for (int e = 0; e < 4; ++e) {
      a[e] = (int)(rate_hash(t * 8 + e) & 0x3F3F3F3Fu);
      b[e] = (int)(rate_hash(t * 8 + 4 + e) & 0x03030303u);
    }
    v16i_r acc[kRateChains];
    for (int c = 0; c < kRateChains; ++c)
      for (int e = 0; e < 16; ++e) acc[c][e] = 0;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int c = 0; c < kRateChains; ++c) acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[c], 0, 0, 0);
    for (int c = 0; c < kRateChains; ++c)
      for (int e = 0; e < 16; ++e) sink += (float)acc[c][e];
  } else if constexpr (KIND == 1) {
    v8i_r a, b;
    for (int e = 0; e < 8; ++e) {   // e4m3 bytes of magnitude <= 2 (exponent field <= 8)
      a[e] = (int)(rate_hash(t * 16 + e) & 0x3F3F3F3Fu);
      b[e] = (int)(rate_hash(t * 16 + 8 + e) & 0x38383838u);
    }
    v16f_r acc[kRateChains];
    for (int c = 0; c < kRateChains; ++c)
      for (int e = 0; e < 16; ++e) acc[c][e] = 0.0f;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int c = 0; c < kRateChains; ++c)
        acc[c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[c], 0, 0, 0, 127, 0, 127);
    for (int c = 0; c < kRateChains; ++c)
      for (int e = 0; e < 16; ++e) sink += acc[c][e];
  } else {
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
      a[e] = (__bf16)((float)(rate_hash(t * 16 + e) & 0xFFu) * (1.0f / 256.0f));
      b[e] = (__bf16)(float)(rate_hash(t * 16 + 8 + e) & 3u);
    }
    v16f_r acc[kRateChains];
    for (int c = 0; c < kRateChains; ++c)
      for (int e = 0; e < 16; ++e) acc[c][e] = 0.0f;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int c = 0; c < kRateChains; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
    for (int c = 0; c < kRateChains; ++c)
      for (int e = 0; e < 16; ++e) sink += acc[c][e];
  }
  out[t] = sink;
}

// Dense operations (2 per multiply-add) of one MFMA of each kind.
static double mfma_ops(int kind) {
  return kind == 0 ? 2.0 * 32 * 32 * 32 : (kind == 1 ? 2.0 * 32 * 32 * 64 : 2.0 * 32 * 32 * 16);
}

void mfma_rate(slfmm_ctx* ctx, int kind, uint64_t iters, double* tops) {
  SLFMM_CHECK(kind >= 0 && kind <= 2, SLFMM_E_ARG, "kind must be 0 (int8), 1 (fp8 e4m3) or 2 (bf16)");
  SLFMM_CHECK(iters >= 1 && iters <= (1u << 24), SLFMM_E_ARG, "iters must be 1..2^24");
  hipStream_t s = ctx->stream;
  int dev = 0, cus = 0;
  SLFMM_HIP(hipGetDevice(&dev));
  SLFMM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)cus * 2;   // 2 workgroups of 4 waves per CU: 2 waves per SIMD
  float* out = ctx->buf<float>("mfma_rate_out", (size_t)grid * 256);
  hipEvent_t e0, e1;
  SLFMM_HIP(hipEventCreate(&e0));
  SLFMM_HIP(hipEventCreate(&e1));
  auto launch = [&](int n) {
    if (kind == 0) k_mfma_rate<0><<<grid, 256, 0, s>>>(n, out);
    else if (kind == 1) k_mfma_rate<1><<<grid, 256, 0, s>>>(n, out);
    else k_mfma_rate<2><<<grid, 256, 0, s>>>(n, out);
  };
  launch((int)std::max<uint64_t>(1, iters / 8));   // warm-up (clocks, code object)
  SLFMM_HIP(hipEventRecord(e0, s));
  launch((int)iters);
  SLFMM_HIP(hipEventRecord(e1, s));
  SLFMM_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  SLFMM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SLFMM_HIP(hipGetLastError());
  const double ops = (double)grid * 4 * (double)iters * kRateChains * mfma_ops(kind);
  *tops = ops / ((double)ms * 1e-3) / 1e12;
}

}  // namespace slfmm
