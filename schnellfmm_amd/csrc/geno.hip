// Genotype layer: device layout of packed PLINK codes, integer per-SNP
// statistics, the bit-exact decode used by parity tests, and a seeded
// Balding-Nichols generator for benchmark-sized panels.
//
// Device layout.  SNP-major rows of 2-bit codes exactly as in a .bed file
// (sample i at byte i/4, bits 2*(i%4); 00 hom-ref, 01 missing, 10 het,
// 11 hom-alt; src/bed.rs:11-16), one row per SNP, row stride `ld` rounded up
// to 16 bytes so every row starts on a 16-byte boundary for dwordx4 loads;
// rows of 512 bytes or more are rounded up to 128 bytes, so that the
// streaming kernels' 64-byte row pieces never straddle a 128-byte cache line
// (a stream of them runs at 5.1 TB/s instead of 4.5 TB/s at n = 10 000).
// Codes past the last sample and bytes past ceil(n/4) are zero (hom-ref) so
// that whole-word arithmetic never needs a tail mask: a zero code adds
// nothing to any sum, and counts of observed calls are n - n_missing.
// kSlack zero bytes follow the last row (kernels load whole 16-byte segments
// of a sample block and may run past the row stride).
#include "internal.h"
#include "rng.h"

namespace slfmm {

uint64_t geno_ld(uint64_t n) {
  const uint64_t bps = cdiv(n, 4);
  return rup(bps, bps >= 512 ? 128 : 16);
}

void geno_alloc(slfmm_ctx* ctx, slfmm_geno* g) {
  g->ld = geno_ld(g->n);
  size_t bytes = (size_t)g->p * g->ld + kSlack;
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess)
    throw Error(SLFMM_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes for genotypes failed");
  g->d = (uint8_t*)p;
  g->owned = true;
  g->ctx = ctx;
  ctx->live_genos += 1;
  SLFMM_HIP(hipMemsetAsync(g->d, 0, bytes, ctx->stream));
}

// One thread per output byte: 4 kept samples of one SNP (sample_keep
// semantics of src/bed.rs:308-318: output row i decodes physical keep[i]).
__global__ void k_repack(const uint8_t* __restrict__ src, uint64_t bps_src, uint64_t p,
                         const int64_t* __restrict__ keep, uint64_t n_keep,
                         uint8_t* __restrict__ dst, uint64_t ld, uint64_t out_bytes) {
  uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t j = blockIdx.y + (uint64_t)blockIdx.z * 65535u;
  if (b >= out_bytes || j >= p) return;
  const uint8_t* row = src + j * bps_src;
  uint32_t out = 0;
  for (int s = 0; s < 4; ++s) {
    uint64_t i = 4 * b + s;
    if (i < n_keep) {
      int64_t ph = keep[i];
      uint32_t code = (row[ph >> 2] >> (2 * (ph & 3))) & 3u;
      out |= code << (2 * s);
    }
  }
  dst[j * ld + b] = (uint8_t)out;
}

// Zero the codes of the padding samples in the last byte of every row.
__global__ void k_clear_tail(uint8_t* d, uint64_t ld, uint64_t p, uint64_t n) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p || (n & 3) == 0) return;
  uint8_t mask = (uint8_t)((1u << (2 * (n & 3))) - 1u);
  d[j * ld + (n >> 2)] &= mask;
}

void geno_clear_tail(hipStream_t s, uint8_t* d, uint64_t ld, uint64_t p, uint64_t n) {
  if (p == 0 || (n & 3) == 0) return;
  k_clear_tail<<<(unsigned)cdiv(p, 256), 256, 0, s>>>(d, ld, p, n);
  SLFMM_HIP(hipGetLastError());
}

static dim3 grid_rows(uint64_t x_blocks, uint64_t p) {
  uint64_t y = p < 65535 ? p : 65535;
  uint64_t z = cdiv(p, 65535);
  return dim3((unsigned)x_blocks, (unsigned)y, (unsigned)z);
}

void geno_upload(slfmm_ctx* ctx, const uint8_t* packed, uint64_t n_phys, uint64_t n_snps,
                 const int64_t* keep, uint64_t n_keep, slfmm_geno* g) {
  const uint64_t bps = cdiv(n_phys, 4);
  if (keep) {
    for (uint64_t i = 0; i < n_keep; ++i)
      SLFMM_CHECK(keep[i] >= 0 && (uint64_t)keep[i] < n_phys, SLFMM_E_ARG,
                  "sample_keep index " + std::to_string(keep[i]) + " out of range");
  }
  g->n = keep ? n_keep : n_phys;
  g->p = n_snps;
  geno_alloc(ctx, g);
  if (n_snps == 0) return;
  hipStream_t s = ctx->stream;
  if (!keep) {
    SLFMM_HIP(hipMemcpy2DAsync(g->d, g->ld, packed, bps, bps, n_snps, hipMemcpyHostToDevice, s));
    geno_clear_tail(s, g->d, g->ld, n_snps, g->n);
  } else {
    // Stream the raw rows through a bounded staging buffer, repacking kept samples.
    const uint64_t rows_per = std::max<uint64_t>(1, (256ull << 20) / std::max<uint64_t>(bps, 1));
    uint8_t* stage = ctx->buf<uint8_t>("upload_stage", std::min(rows_per, n_snps) * bps);
    int64_t* dkeep = ctx->buf<int64_t>("upload_keep", n_keep);
    SLFMM_HIP(hipMemcpyAsync(dkeep, keep, n_keep * 8, hipMemcpyHostToDevice, s));
    const uint64_t out_bytes = cdiv(n_keep, 4);
    for (uint64_t r0 = 0; r0 < n_snps; r0 += rows_per) {
      uint64_t nr = std::min(rows_per, n_snps - r0);
      SLFMM_HIP(hipMemcpyAsync(stage, packed + r0 * bps, nr * bps, hipMemcpyHostToDevice, s));
      k_repack<<<grid_rows(cdiv(out_bytes, 256), nr), 256, 0, s>>>(stage, bps, nr, dkeep, n_keep,
                                                                   g->d + r0 * g->ld, g->ld, out_bytes);
      SLFMM_HIP(hipStreamSynchronize(s));   // staging buffer is reused
    }
  }
  SLFMM_HIP(hipGetLastError());
}

// One wave per gathered row (16 bytes per lane per pass), rows taken in a
// grid-stride loop by the waves of a fixed grid: a row is a few KB, so one
// workgroup per row left most of each short workgroup's time to its launch.
__global__ void __launch_bounds__(256) k_gather(const uint8_t* __restrict__ src, uint64_t ld,
                                                const int64_t* __restrict__ idx, uint64_t m,
                                                uint8_t* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = ld / 16;
  for (uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < m; r += (uint64_t)gridDim.x * 4) {
    const uint4* s = reinterpret_cast<const uint4*>(src + (uint64_t)idx[r] * ld);
    uint4* d = reinterpret_cast<uint4*>(dst + r * ld);
    for (uint64_t w = lane; w < nw; w += 64) d[w] = s[w];
  }
}

void geno_gather(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* d_idx, uint64_t m, uint8_t* dst) {
  if (m == 0) return;
  const unsigned blocks = (unsigned)std::min<uint64_t>(cdiv(m, 4), 8192);
  k_gather<<<blocks, 256, 0, ctx->stream>>>(g->d, g->ld, d_idx, m, dst);
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Integer statistics: one wave per SNP row.  From the lo/hi bit planes of a
// word of 16 codes: missing = lo & ~hi, hom-alt = lo & hi, dosage sum =
// popc(hi) + popc(lo & hi).  Exact; mean = sum_g / (n - n_miss) reproduces
// the reference's f64 sum of integer dosages bit for bit.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_stats(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t p,
                                               int32_t* n_miss, int32_t* sum_g, int32_t* n_hom) {
  const int lane = threadIdx.x & 63;
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= p) return;
  const uint4* row = reinterpret_cast<const uint4*>(packed + j * ld);
  uint32_t cm = 0, cs = 0, ch = 0;
  for (uint64_t w = lane; w < ld / 16; w += 64) {
    uint4 v = row[w];
    uint32_t ws[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t lo = ws[q] & 0x55555555u, hi = (ws[q] >> 1) & 0x55555555u;
      uint32_t hb = lo & hi;
      cm += __popc(lo & ~hi);
      cs += __popc(hi) + __popc(hb);
      ch += __popc(hb);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    cm += __shfl_xor(cm, off);
    cs += __shfl_xor(cs, off);
    ch += __shfl_xor(ch, off);
  }
  if (lane == 0) {
    n_miss[j] = (int32_t)cm;
    sum_g[j] = (int32_t)cs;
    n_hom[j] = (int32_t)ch;
  }
}

void launch_stats(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                  int32_t* n_miss, int32_t* sum_g, int32_t* n_hom) {
  (void)n;
  if (p == 0) return;
  k_stats<<<(unsigned)cdiv(p, 4), 256, 0, s>>>(packed, ld, p, n_miss, sum_g, n_hom);
  SLFMM_HIP(hipGetLastError());
}

// compute_scale (src/bed.rs:382-395).  This file is compiled with
// -ffp-contract=off so the operation sequence matches the reference.
__device__ __forceinline__ double dev_compute_scale(double mean, int norm) {
  double p = mean / 2.0;
  double twopq = 2.0 * p * (1.0 - p);
  if (norm == SLFMM_NORM_CENTER_ONLY) return twopq < 1e-20 ? 0.0 : 1.0;
  double denom = sqrt(twopq);
  return denom > 1e-10 ? 1.0 / denom : 0.0;
}

__global__ void k_mean_scale(const int32_t* n_miss, const int32_t* sum_g, uint64_t n, uint64_t p, int norm,
                             double* mu, double* sc) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p) return;
  int64_t nobs = (int64_t)n - n_miss[j];
  double mean = nobs > 0 ? (double)sum_g[j] / (double)nobs : 0.0;
  mu[j] = mean;
  sc[j] = dev_compute_scale(mean, norm);
}

void launch_mean_scale(hipStream_t s, const int32_t* n_miss, const int32_t* sum_g, uint64_t n,
                       uint64_t p, int norm, double* mu, double* sc) {
  if (p == 0) return;
  k_mean_scale<<<(unsigned)cdiv(p, 256), 256, 0, s>>>(n_miss, sum_g, n, p, norm, mu, sc);
  SLFMM_HIP(hipGetLastError());
}

// Bit-exact decode_single_snp (src/bed.rs:300-379): one thread per column,
// left-to-right sum exactly as the reference loop (order matters with fills).
__global__ void k_decode_f64(const uint8_t* __restrict__ packed, uint64_t ld, uint64_t n,
                             const int64_t* __restrict__ idx, uint64_t c, int norm,
                             const double* __restrict__ fill, double* __restrict__ out, int raw) {
  uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= c) return;
  const uint8_t* row = packed + (uint64_t)idx[col] * ld;
  const double table[4] = {0.0, __longlong_as_double(0x7ff8000000000000ll), 1.0, 2.0};
  if (raw) {
    for (uint64_t i = 0; i < n; ++i) out[i * c + col] = table[(row[i >> 2] >> (2 * (i & 3))) & 3u];
    return;
  }
  double sum = 0.0;
  uint32_t nvalid = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t code = (row[i >> 2] >> (2 * (i & 3))) & 3u;
    double v = table[code];
    if (code == 1u && fill) v = fill[i * c + col];
    if (!(v != v)) { sum += v; ++nvalid; }
  }
  double mean = nvalid > 0 ? sum / (double)nvalid : 0.0;
  double scale = dev_compute_scale(mean, norm);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t code = (row[i >> 2] >> (2 * (i & 3))) & 3u;
    double v = table[code];
    if (code == 1u && fill) v = fill[i * c + col];
    out[i * c + col] = (v != v) ? 0.0 : (v - mean) * scale;
  }
}

void launch_decode_f64(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t n,
                       const int64_t* idx, uint64_t c, int norm, const double* fill, double* out,
                       int raw) {
  if (c == 0) return;
  k_decode_f64<<<(unsigned)cdiv(c, 64), 64, 0, s>>>(packed, ld, n, idx, c, norm, fill, out, raw);
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Sample-major transpose: out row i (sample) holds the codes of SNPs
// 0..p-1, 4 SNPs per byte (SNP j at byte j/4, bits 2*(j%4)).  Tile of
// 128 SNPs x 128 samples: 16-byte loads of the SNP rows into LDS words,
// 16-byte stores of the sample rows.  Every byte [0, ldt) of rows i < n is
// written (codes past p are 0); ldt must be a multiple of 16.
// ---------------------------------------------------------------------------
// 16 x 16 two-bit codes held as 16 words (word r = row r, code c at bits
// 2c): four rounds of masked block swaps (8, 4, 2, 1 rows apart) leave word c
// holding column c, code r at bits 2r.
__device__ __forceinline__ void transpose16_2bit(uint32_t (&w)[16]) {
#pragma unroll
  for (int round = 0; round < 4; ++round) {
    const int s = 8 >> round;
    const int sh = 2 * s;
    const uint32_t mask = round == 0 ? 0x0000FFFFu : round == 1 ? 0x00FF00FFu : round == 2 ? 0x0F0F0F0Fu : 0x33333333u;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      if (r & s) continue;
      const uint32_t t = ((w[r] >> sh) ^ w[r + s]) & mask;
      w[r] ^= t << sh;
      w[r + s] ^= t;
    }
  }
}

// Tile of 256 SNPs x 256 samples: 256 SNP rows of 64 bytes are staged in LDS
// (four threads per row, so a wave moves 16 whole 64-byte row pieces), each
// thread transposes one 16 x 16 block of codes in registers, and the 256
// sample rows go back out the same way.  Transposed words are stored with
// their word index rotated by the sample block (LDS banks), and read back
// with the same rotation.
__global__ void __launch_bounds__(256) k_transpose(const uint8_t* __restrict__ in, uint64_t ld, uint64_t n,
                                                   uint64_t p, uint8_t* __restrict__ out, uint64_t ldt) {
  __shared__ uint32_t tile[256][17];
  const uint64_t j0 = (uint64_t)blockIdx.x * 256;   // SNP tile
  const uint64_t i0 = (uint64_t)blockIdx.y * 256;   // sample tile
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {   // piece e: SNP row e / 4, 16 bytes e % 4 (rows past p, bytes past ld: zero)
    const int e = t + 256 * k, row = e >> 2, q = e & 3;
    const uint64_t j = j0 + row, off = i0 / 4 + 16 * q;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (j < p && off + 16 <= ld) v = *reinterpret_cast<const uint4*>(in + j * ld + off);
    tile[row][4 * q + 0] = v.x; tile[row][4 * q + 1] = v.y; tile[row][4 * q + 2] = v.z; tile[row][4 * q + 3] = v.w;
  }
  __syncthreads();
  const int vb = t >> 4, wb = t & 15;   // SNP block vb, sample block wb
  uint32_t w[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) w[r] = tile[16 * vb + r][wb];
  transpose16_2bit(w);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 16; ++c) tile[16 * wb + c][(vb + wb) & 15] = w[c];   // sample row 16 wb + c, word vb
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {   // piece e: sample row e / 4, words 4 (e % 4) .. + 3
    const int e = t + 256 * k, row = e >> 2, q = e & 3, rot = row >> 4;
    const uint64_t i = i0 + row, ob = j0 / 4 + 16 * q;
    if (i < n && ob + 16 <= ldt)
      *reinterpret_cast<uint4*>(out + i * ldt + ob) =
          make_uint4(tile[row][(4 * q + 0 + rot) & 15], tile[row][(4 * q + 1 + rot) & 15],
                     tile[row][(4 * q + 2 + rot) & 15], tile[row][(4 * q + 3 + rot) & 15]);
  }
}

void launch_transpose(hipStream_t s, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                      uint8_t* out, uint64_t ldt) {
  if (p == 0 || n == 0) return;
  SLFMM_CHECK(ldt % 16 == 0 && ld % 16 == 0, SLFMM_E_ARG, "transpose strides must be multiples of 16 bytes");
  k_transpose<<<dim3((unsigned)cdiv(p, 256), (unsigned)cdiv(n, 256)), 256, 0, s>>>(packed, ld, n, p, out, ldt);
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Counter-based random numbers (SplitMix64 finaliser over a seeded counter).
// ---------------------------------------------------------------------------


// Marsaglia-Tsang gamma(a) with a bounded number of proposals.
__device__ double gamma_draw(double a, uint64_t seed, uint64_t base) {
  double boost = 1.0;
  if (a < 1.0) {
    boost = pow(u01(seed, base + 1001) + 1e-300, 1.0 / a);
    a += 1.0;
  }
  double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (int it = 0; it < 64; ++it) {
    double x = normal01(seed, base + 4 * it);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    double u = u01(seed, base + 4 * it + 3);
    if (log(u + 1e-300) < 0.5 * x * x + d - d * v + d * log(v)) return d * v * boost;
  }
  return a * boost;   // mean, if every proposal was rejected
}

__global__ void k_bn_freqs(float* freq, uint64_t p, uint32_t n_pops, double fst_lo, double fst_hi,
                           double maf_lo, double maf_hi, uint64_t seed) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p) return;
  double panc = maf_lo + (maf_hi - maf_lo) * u01(seed ^ 0xA5A5A5A5ull, j);
  for (uint32_t k = 0; k < n_pops; ++k) {
    double F = fst_lo + (fst_hi - fst_lo) * u01(seed ^ 0x5EEDull, k);
    double a = panc * (1.0 - F) / F, b = (1.0 - panc) * (1.0 - F) / F;
    uint64_t base = (j * n_pops + k) * 1024;
    double x = gamma_draw(a, seed + 17, base), y = gamma_draw(b, seed + 29, base);
    double f = x / (x + y);
    f = f < 1e-6 ? 1e-6 : (f > 1.0 - 1e-6 ? 1.0 - 1e-6 : f);
    freq[j * n_pops + k] = (float)f;
  }
}

// One thread per 32-bit word (16 samples) of one SNP row.
__global__ void k_bn_geno(const float* __restrict__ freq, uint8_t* packed, uint64_t ld, uint64_t n,
                          uint64_t p, uint32_t n_pops, double miss_rate, uint64_t seed) {
  uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t j = blockIdx.y + (uint64_t)blockIdx.z * 65535u;
  if (j >= p || w * 16 >= n) return;
  uint32_t word = 0;
  const uint32_t thr_miss = (uint32_t)(miss_rate * 2097152.0);
  for (int s = 0; s < 16; ++s) {
    uint64_t i = w * 16 + s;
    if (i >= n) break;
    uint32_t pop = (uint32_t)((i * n_pops) / n);
    float f = freq[j * n_pops + pop];
    uint32_t tf = (uint32_t)(f * 2097152.0f);
    uint64_t h = mix64(seed, j * n + i);
    uint32_t u1 = (uint32_t)(h & 0x1FFFFF), u2 = (uint32_t)((h >> 21) & 0x1FFFFF), u3 = (uint32_t)(h >> 43);
    uint32_t g = (u1 < tf) + (u2 < tf);
    uint32_t code = (u3 < thr_miss) ? 1u : (g == 0 ? 0u : (g == 1 ? 2u : 3u));
    word |= code << (2 * s);
  }
  reinterpret_cast<uint32_t*>(packed + j * ld)[w] = word;
}

// ---- generalised panel: admixture, inbred lines, Beta MAF, structured missingness
struct SynthDev {
  uint32_t n_pops;
  double fst_lo, fst_hi;
  int maf_beta;
  double maf_a, maf_b, maf_min;
  const float* ancestry;      // device n x n_pops, or null
  double het_rate;            // < 0: Hardy-Weinberg
  double miss_rate;
  const float* sample_miss;   // device n, or null
  double snp_miss_a, snp_miss_b;
  double miss_scale;
  uint64_t seed;
};

__global__ void k_bn_freqs_ex(float* freq, float* snp_miss, uint64_t p, SynthDev sp, uint64_t seed) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p) return;
  double panc;
  if (sp.maf_beta) {
    const double x = gamma_draw(sp.maf_a, seed + 101, j * 1024), y = gamma_draw(sp.maf_b, seed + 103, j * 1024);
    panc = x / (x + y);
    panc = panc < sp.maf_min ? sp.maf_min : (panc > 1.0 - sp.maf_min ? 1.0 - sp.maf_min : panc);
  } else {
    panc = sp.maf_a + (sp.maf_b - sp.maf_a) * u01(seed ^ 0xA5A5A5A5ull, j);
  }
  for (uint32_t k = 0; k < sp.n_pops; ++k) {
    double F = sp.fst_lo + (sp.fst_hi - sp.fst_lo) * u01(seed ^ 0x5EEDull, k);
    double a = panc * (1.0 - F) / F, b = (1.0 - panc) * (1.0 - F) / F;
    uint64_t base = (j * sp.n_pops + k) * 1024;
    double x = gamma_draw(a, seed + 17, base), y = gamma_draw(b, seed + 29, base);
    double f = x / (x + y);
    f = f < 1e-6 ? 1e-6 : (f > 1.0 - 1e-6 ? 1.0 - 1e-6 : f);
    freq[j * sp.n_pops + k] = (float)f;
  }
  double bm = 0.0;
  if (sp.snp_miss_a > 0.0) {
    const double x = gamma_draw(sp.snp_miss_a, seed + 211, j * 1024), y = gamma_draw(sp.snp_miss_b, seed + 223, j * 1024);
    bm = x / (x + y);
  }
  snp_miss[j] = (float)bm;
}

__global__ void k_bn_geno_ex(const float* __restrict__ freq, const float* __restrict__ snp_miss, uint8_t* packed,
                             uint64_t ld, uint64_t n, uint64_t p, SynthDev sp, uint64_t seed) {
  uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t j = blockIdx.y + (uint64_t)blockIdx.z * 65535u;
  if (j >= p || w * 16 >= n) return;
  uint32_t word = 0;
  const double bj = snp_miss[j];
  const uint32_t thr_het = sp.het_rate >= 0.0 ? (uint32_t)(sp.het_rate * 2097152.0) : 0u;
  for (int s = 0; s < 16; ++s) {
    uint64_t i = w * 16 + s;
    if (i >= n) break;
    float f;
    if (sp.ancestry) {
      float acc = 0.0f;
      for (uint32_t k = 0; k < sp.n_pops; ++k) acc += sp.ancestry[i * sp.n_pops + k] * freq[j * sp.n_pops + k];
      f = acc;
    } else {
      f = freq[j * sp.n_pops + (uint32_t)((i * sp.n_pops) / n)];
    }
    const double ai = sp.sample_miss ? (double)sp.sample_miss[i] : 0.0;
    double pm = sp.miss_scale * (1.0 - (1.0 - sp.miss_rate) * (1.0 - ai) * (1.0 - bj));
    pm = pm < 1.0 ? pm : 1.0;
    const uint32_t tf = (uint32_t)(f * 2097152.0f), tm = (uint32_t)(pm * 2097152.0);
    uint64_t h = mix64(seed, j * n + i);
    uint32_t u1 = (uint32_t)(h & 0x1FFFFF), u2 = (uint32_t)((h >> 21) & 0x1FFFFF), u3 = (uint32_t)(h >> 43);
    uint32_t g;
    if (sp.het_rate >= 0.0) g = u2 < thr_het ? 1u : 2u * (u1 < tf);
    else g = (u1 < tf) + (u2 < tf);
    uint32_t code = (u3 < tm) ? 1u : (g == 0 ? 0u : (g == 1 ? 2u : 3u));
    word |= code << (2 * s);
  }
  reinterpret_cast<uint32_t*>(packed + j * ld)[w] = word;
}

void launch_synth_ex(hipStream_t s, uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p, const slfmm_synth_spec& spec) {
  if (p == 0 || n == 0) return;
  SynthDev sp{spec.n_pops, spec.fst_lo, spec.fst_hi, spec.maf_beta, spec.maf_a, spec.maf_b, spec.maf_min, nullptr,
              spec.het_rate, spec.miss_rate, nullptr, spec.snp_miss_a, spec.snp_miss_b,
              spec.miss_scale > 0.0 ? spec.miss_scale : 1.0, spec.seed};
  const uint64_t chunk = 1ull << 22;   // SNPs per frequency batch (bounded scratch)
  float *freq = nullptr, *bmiss = nullptr, *anc = nullptr, *smiss = nullptr;
  try {
    SLFMM_HIP(hipMalloc(&freq, std::min(chunk, p) * spec.n_pops * sizeof(float)));
    SLFMM_HIP(hipMalloc(&bmiss, std::min(chunk, p) * sizeof(float)));
    if (spec.ancestry) {
      SLFMM_HIP(hipMalloc(&anc, n * spec.n_pops * sizeof(float)));
      SLFMM_HIP(hipMemcpyAsync(anc, spec.ancestry, n * spec.n_pops * sizeof(float), hipMemcpyHostToDevice, s));
      sp.ancestry = anc;
    }
    if (spec.sample_miss) {
      SLFMM_HIP(hipMalloc(&smiss, n * sizeof(float)));
      SLFMM_HIP(hipMemcpyAsync(smiss, spec.sample_miss, n * sizeof(float), hipMemcpyHostToDevice, s));
      sp.sample_miss = smiss;
    }
    for (uint64_t j0 = 0; j0 < p; j0 += chunk) {
      uint64_t nj = std::min(chunk, p - j0);
      const uint64_t sd = spec.seed + 7919 * (j0 / chunk);
      k_bn_freqs_ex<<<(unsigned)cdiv(nj, 256), 256, 0, s>>>(freq, bmiss, nj, sp, sd);
      k_bn_geno_ex<<<grid_rows(cdiv(cdiv(n, 16), 256), nj), 256, 0, s>>>(freq, bmiss, packed + j0 * ld, ld, n, nj, sp,
                                                                          spec.seed * 31 + j0);
      SLFMM_HIP(hipGetLastError());
    }
    SLFMM_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    hipFree(freq); hipFree(bmiss); hipFree(anc); hipFree(smiss);
    throw;
  }
  SLFMM_HIP(hipFree(freq));
  SLFMM_HIP(hipFree(bmiss));
  if (anc) SLFMM_HIP(hipFree(anc));
  if (smiss) SLFMM_HIP(hipFree(smiss));
}

void launch_synth_bn(hipStream_t s, uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p,
                     uint32_t n_pops, double fst_lo, double fst_hi, double maf_lo, double maf_hi,
                     double miss_rate, uint64_t seed) {
  if (p == 0 || n == 0) return;
  const uint64_t chunk = 1ull << 22;   // SNPs per frequency batch (bounded scratch)
  float* freq = nullptr;
  SLFMM_HIP(hipMalloc(&freq, std::min(chunk, p) * n_pops * sizeof(float)));
  for (uint64_t j0 = 0; j0 < p; j0 += chunk) {
    uint64_t nj = std::min(chunk, p - j0);
    k_bn_freqs<<<(unsigned)cdiv(nj, 256), 256, 0, s>>>(freq, nj, n_pops, fst_lo, fst_hi, maf_lo, maf_hi,
                                                      seed + 7919 * (j0 / chunk));
    k_bn_geno<<<grid_rows(cdiv(cdiv(n, 16), 256), nj), 256, 0, s>>>(freq, packed + j0 * ld, ld, n, nj, n_pops,
                                                                    miss_rate, seed * 31 + j0);
    SLFMM_HIP(hipGetLastError());
  }
  SLFMM_HIP(hipStreamSynchronize(s));
  SLFMM_HIP(hipFree(freq));
}

}  // namespace slfmm
