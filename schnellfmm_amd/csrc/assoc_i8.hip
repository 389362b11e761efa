// Mixed-precision genotype-stream kernels on gfx950 int8 MFMA.
//
// Genotype codes are exact small integers, so they enter the matrix cores
// as int8 with no rounding: per 16 samples of one SNP the lane decodes one
// 32-bit word of packed codes into two int8 operand fragments,
//   g'   = scaled dosage, missing -> fill code  (00->0, 01->c, 10->S, 11->2S)
//   miss = missing indicator                    (01->1, else 0)
// and v_mfma_i32_16x16x64_i8 accumulates exactly in int32.  The dense
// operand (an orthonormal basis Q of [U_c|X] for the association, the
// sketch B for the RSVD) is carried as 2 or 3 int8 terms per entry with a
// per-column scale, Q ~ sc (t0 + t1/254 + t2/254^2) (about 15 / 22
// significant bits) against g', and as its leading 1 or 2 terms against the
// missing indicator.  The recombined functional of every SNP is that of the
// centred column (g - mu on observed calls, 0 on missing calls;
// src/bed.rs:339-343), and the FP64 per-SNP epilogue of the association pass
// runs on it unchanged.
//
// SNP-side kernels (association pass, RSVD sketch) run
// v_mfma_i32_32x32x32_i8: lane l holds SNP column l % 32 of a 32-SNP tile
// and samples 16 (l / 32) .. + 15 of a k-step.  For a block of 256 samples,
// lane half h = l / 32 owns samples 128 h .. 128 h + 127 of its SNP row (32
// contiguous bytes, 8 words), word s serving k-step s; byte c of decoded
// register v holds sample 256 blk + 128 h + 16 s + 4 c + v.  The dense
// operand is pre-permuted into exactly this order (one 16-byte fragment per
// lane), so the same K index meets on both sides.  (The RSVD forward pass,
// on the sample-major copy, uses v_mfma_i32_16x16x64_i8 with the analogous
// order.)
#include "internal.h"
#include "colreduce.h"
#include "epilogue.h"

namespace slfmm {

typedef int v4i __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Quantisation of the dense operand
// ---------------------------------------------------------------------------
__device__ __forceinline__ int clamp127(double v) {
  int x = (int)v;
  return x > 127 ? 127 : (x < -127 ? -127 : x);
}

__device__ __forceinline__ int quant_term(double q, int t) {
  double h = rint(q);
  if (t == 0) return clamp127(h);
  double r1 = (q - h) * 254.0;
  double md = rint(r1);
  if (t == 1) return clamp127(md);
  return clamp127(rint((r1 - md) * 254.0));
}

// Row layout of the packed operand of the SNP-side kernels.  Row 32 rt + r
// of a v_mfma_i32_32x32x32_i8 result, r = 8 b + 4 h + v, lands in register
// 4 b + v of lane half h, so functional i (a column of the dense operand) is
// given to half h = i % 2 at position q = i / 2, and every int32 sum it
// needs ends in the registers of the lanes that recombine it (no transpose
// through LDS).  Within a half the rows of tile rt are slots s = 16 rt +
// 4 b + v:
//   g' tiles (rt < NT0):   terms 0 and 1 of position q at slots 2q, 2q + 1;
//                          term 2 (columns i >= k3) at slot 3 NQ - 1 - q
//   miss tiles (NT1):      term 0 of position q at slot q; term 1 (columns
//                          i >= km) at slot 2 NQ - 1 - q.
// One all-ones row sits at the last slot of half g1: in a miss tile (slot
// 16 NT1 - 1, the missing count) or, when the miss tiles are full, in a g'
// tile (slot 16 NT0 - 1, the sum of g').  All slot indices are compile-time
// functions of q, so the accumulators are addressed statically.
struct RowMap {
  int m, k3, km, NQ, NT0, NT1, g1, ones_g;
  static RowMap make(int m, int k3, int km) {
    RowMap r{m, k3, km, (m + 1) / 2, 0, 0, 0, 0};
    int topG = 0, topM = 0, tgs[2], tms[2];
    for (int h = 0; h < 2; ++h) {
      const int nq = h < m ? (m - h + 1) / 2 : 0;
      int qx = nq, qm = nq;   // first positions holding a column >= k3, >= km
      for (int q = nq - 1; q >= 0; --q) {
        if (h + 2 * q >= k3) qx = q;
        if (h + 2 * q >= km) qm = q;
      }
      int tg = 2 * nq - 1, tm = nq - 1;
      if (qx < nq) tg = std::max(tg, 3 * r.NQ - 1 - qx);
      if (qm < nq) tm = std::max(tm, 2 * r.NQ - 1 - qm);
      tgs[h] = tg; tms[h] = tm;
      topG = std::max(topG, tg);
      topM = std::max(topM, tm);
    }
    r.NT0 = (topG + 16) / 16;
    r.NT1 = (topM + 16) / 16;
    for (int h = 0; h < 2; ++h)
      if (tms[h] < 16 * r.NT1 - 1) { r.g1 = h; return r; }
    for (int h = 0; h < 2; ++h)
      if (tgs[h] < 16 * r.NT0 - 1) { r.g1 = h; r.ones_g = 1; return r; }
    ++r.NT1;   // room for the ones row
    return r;
  }
  // (column, term) of packed row `row`: column -1 for padding, -2 for the ones row
  __host__ __device__ int2 info(int row) const {
    const int rt = row >> 5, rr = row & 31, h = (rr >> 2) & 1, sl = 4 * (rr >> 3) + (rr & 3);
    if (rt < NT0) {
      const int s = 16 * rt + sl;
      if (ones_g && h == g1 && s == 16 * NT0 - 1) return make_int2(-2, 0);
      if (s < 2 * NQ) {
        const int i = h + 2 * (s >> 1);
        return i < m ? make_int2(i, s & 1) : make_int2(-1, 0);
      }
      const int q = 3 * NQ - 1 - s, i = h + 2 * q;
      return q >= 0 && i < m && i >= k3 ? make_int2(i, 2) : make_int2(-1, 0);
    }
    const int s = 16 * (rt - NT0) + sl;
    if (!ones_g && h == g1 && s == 16 * NT1 - 1) return make_int2(-2, 0);
    if (s < NQ) {
      const int i = h + 2 * s;
      return i < m ? make_int2(i, 0) : make_int2(-1, 0);
    }
    const int q = 2 * NQ - 1 - s, i = h + 2 * q;
    return q >= 0 && i < m && i >= km ? make_int2(i, 1) : make_int2(-1, 0);
  }
};

// Row layout of the operand of the NMF backward pass (MODE 2): the K
// columns of W on the dosage side (3 terms each: rows 3 r + t of the first
// NT0 tiles) and the P = K (K + 1) / 2 products W_r W_s (r <= s, pair
// p = r K - r (r - 1) / 2 + s - r) on the missing side (2 terms each: rows
// 2 p + t of the NT1 tiles after them).  Operand column K + p holds pair p.
struct NmfMap {
  int K, P, NT0, NT1;
  static NmfMap make(int K) {
    const int P = K * (K + 1) / 2;
    return NmfMap{K, P, (3 * K + 31) / 32, (2 * P + 31) / 32};
  }
  __host__ __device__ int2 info(int row) const {
    if (row < 32 * NT0) return row < 3 * K ? make_int2(row / 3, row % 3) : make_int2(-1, 0);
    const int r = row - 32 * NT0;
    return r < 2 * P ? make_int2(K + r / 2, r % 2) : make_int2(-1, 0);
  }
};

// Apack for the SNP-side kernels: rows = functional terms, K = samples.
// Fragment (blk, s, rt) is 1 KB (8 k-steps of 32 samples per 256-sample
// block): lane l holds row 32 rt + l % 32, byte 4v + c = sample
// 256 blk + 128 (l / 32) + 16 s + 4 c + v.
template <class Map>
__global__ void k_pack_rows(const double* __restrict__ Q, uint64_t ldq, uint64_t n, Map rmap,
                            const double* __restrict__ sc, int NT, uint64_t total, int8_t* __restrict__ out) {
  constexpr int KS = 8;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e & 15), lane = (int)((e >> 4) & 63);
    const uint64_t frag = e >> 10;                 // (blk*KS + s)*NT + rt
    const int rt = (int)(frag % NT);
    const uint64_t ks = frag / NT;
    const int s = (int)(ks % KS);
    const uint64_t blk = ks / KS;
    const int v = j >> 2, c = j & 3, h = lane >> 5;
    const uint64_t sample = blk * 256 + 128 * h + 16 * s + 4 * c + v;
    const int2 ri = rmap.info(32 * rt + (lane & 31));
    int val = 0;
    if (sample < n) {
      if (ri.x >= 0) val = quant_term(Q[sample * ldq + ri.x] / sc[ri.x], ri.y);
      else if (ri.x == -2) val = 1;
    }
    out[e] = (int8_t)val;
  }
}

// O_t = sum_k term_t(Q[k][i]) over the n samples (exact integers in FP64).
struct FnTerms {
  const double* Q; uint64_t ldq; const double* sc;
  __device__ void operator()(uint64_t r, int c, double (&v)[3]) const {
    const double q = Q[r * ldq + c] / sc[c];
    v[0] = (double)quant_term(q, 0);
    v[1] = (double)quant_term(q, 1);
    v[2] = (double)quant_term(q, 2);
  }
};
struct GStore3 {
  double* out;
  __device__ void operator()(int c, const double (&v)[3]) const {
    out[c * 3 + 0] = v[0]; out[c * 3 + 1] = v[1]; out[c * 3 + 2] = v[2];
  }
};

// Recombination weights of one packed operand, [3][32] doubles:
// wg_i = sc_i 254^-(ng-1) / S,  wm_i = sc_i 254^-(nm-1),
// C_i  = sc_i sum_{t < ng} 254^-t O_t   (ng = 3 for i >= k3, else 2; nm = 2 for
// i >= km, else 1; zero for i >= m).
constexpr int kWtab = 3 * 32;
__global__ void k_weights(int m, int k3, int km, double S, const double* __restrict__ sc, const double* __restrict__ O,
                          double* __restrict__ w) {
  const int i = threadIdx.x;
  if (i >= 32) return;
  double wg = 0.0, wm = 0.0, C = 0.0;
  if (i < m) {
    const bool x = i >= k3;
    wg = sc[i] * (x ? 1.0 / 64516.0 : 1.0 / 254.0) / S;
    wm = sc[i] * (i >= km ? 1.0 / 254.0 : 1.0);
    double c = O[i * 3 + 0] + O[i * 3 + 1] / 254.0;
    if (x) c += O[i * 3 + 2] / 64516.0;
    C = sc[i] * c;
  }
  w[i] = wg;
  w[32 + i] = wm;
  w[64 + i] = C;
}

struct Operand {
  int8_t* pack = nullptr;
  RowMap rmap{};
  uint64_t nblk = 0;
  double* wtab = nullptr;
  int S = 1;
  int fill = 0;   // per-SNP fill code in the g' fragment
};

// Scale of the dosage in the g' fragment (00 -> 0, 01 -> c, 10 -> S, 11 -> 2S):
// the int32 sums stay below 2^31 while 127 * 2S * n does.
// Scale S of the g' codes {0, c, S, 2S}: 127 (2S) n must stay below 2^31.
static int fill_scale(uint64_t n) {
  return n <= 262144 ? 32 : (n <= 524288 ? 16 : (n <= 1048576 ? 8 : (n <= 2000000 ? 4 : (n <= 4000000 ? 2 : 1))));
}

// Quantise Q (n x m, ldq) into the packed fragment layout: columns from k3
// on carry 3 terms on the g' side, columns from km on 2 terms on the miss
// side.  With fill, the g' fragment carries the per-SNP fill code.  Device
// work only.
static Operand prepare_operand(slfmm_ctx* ctx, const char* tag, const double* Q, uint64_t ldq, uint64_t n, int m,
                               int k3, int km, bool fill) {
  hipStream_t s = ctx->stream;
  SLFMM_CHECK(m >= 1 && m <= 32 && k3 >= 0 && k3 <= m, SLFMM_E_ARG, "operand shape");
  SLFMM_CHECK(n <= 8000000, SLFMM_E_ARG, "too many samples for the int8 kernels (use SLFMM_PREC_FP64)");
  Operand op;
  std::string T(tag);
  op.rmap = RowMap::make(m, k3, km);
  op.fill = fill ? 1 : 0;
  op.S = fill ? fill_scale(n) : 1;
  const int NT = op.rmap.NT0 + op.rmap.NT1;
  op.nblk = cdiv(n, 256);
  double* sc = ctx->buf<double>(T + "_sc", m);
  colabsmax_scale(ctx, Q, ldq, n, m, nullptr, nullptr, sc);
  double* O = ctx->buf<double>(T + "_O", m * 3);
  colreduce<3, kColSum>(ctx, FnTerms{Q, ldq, sc}, GStore3{O}, n, m);
  op.wtab = ctx->buf<double>(T + "_w", kWtab);
  k_weights<<<1, 32, 0, s>>>(m, k3, km, (double)op.S, sc, O, op.wtab);
  const uint64_t total = op.nblk * 8 * NT * 1024;
  op.pack = ctx->buf<int8_t>(T + "_pack", total);
  k_pack_rows<<<(unsigned)std::min<uint64_t>(cdiv(total, 256), 16384), 256, 0, s>>>(Q, ldq, n, op.rmap, sc, NT, total,
                                                                                  op.pack);
  SLFMM_HIP(hipGetLastError());
  return op;
}

// ---------------------------------------------------------------------------
// Genotype fragment decode
// ---------------------------------------------------------------------------
struct Frag {
  v4i g0, miss;
};

// One word = 16 two-bit codes of one SNP.  Register v of a fragment takes
// the codes at bit 8c + 2v (c = 0..3): one shift and mask leave one code per
// byte, and v_perm_b32 maps each byte through a 4-entry table (selector
// bytes 0..3 pick bytes of the second operand).
constexpr uint32_t kPermG0 = 0x02010000u;     // 00->0, 01->0, 10->1, 11->2
constexpr uint32_t kPermMiss = 0x00000100u;   // 01->1, else 0

// Plain dosage / missing fragments (RSVD forward pass) with the per-word
// counts popc(w) (dosage + 1 per missing call) and popc of the high bits
// (het + hom).
__device__ __forceinline__ Frag decode_word(uint32_t w, uint32_t& c_pop, uint32_t& c_hi) {
  c_pop += __popc(w);
  c_hi += __popc(w & 0xAAAAAAAAu);
  Frag f;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const uint32_t cv = (w >> (2 * v)) & 0x03030303u;
    f.g0[v] = (int)__builtin_amdgcn_perm(0u, kPermG0, cv);
    f.miss[v] = (int)__builtin_amdgcn_perm(0u, kPermMiss, cv);
  }
  return f;
}

template <int KS>
__device__ __forceinline__ void load_words(const uint8_t* p, uint32_t (&w)[KS]) {
  if constexpr (KS == 4) {
    uint4 q = *reinterpret_cast<const uint4*>(p);
    w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
  } else if constexpr (KS == 2) {
    uint2 q = *reinterpret_cast<const uint2*>(p);
    w[0] = q.x; w[1] = q.y;
  } else {
    w[0] = *reinterpret_cast<const uint32_t*>(p);
  }
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Asynchronous copy of one 1 KB fragment chunk (64 lanes x 16 B) from global
// memory straight into LDS (global_load_lds_dwordx4): lane l's 16 bytes land
// at dst + 16 l.  The destination is uniform across the wave.
__device__ __forceinline__ void chunk_to_lds(const unsigned char* src_chunk, unsigned char* dst_chunk, int lane) {
  __builtin_amdgcn_global_load_lds((const void*)(src_chunk + 16 * lane), (lds_ptr_t)(dst_chunk), 16, 0, 0);
}

constexpr int kOperandPrefetch = 6;   // operand fragments read ahead of their MFMAs

// vmcnt wait with a compile-time count (vmcnt[3:0]; expcnt and lgkmcnt at
// their maxima, i.e. not waited on).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N <= 15, "vmcnt immediate");
  __builtin_amdgcn_s_waitcnt(0x0F70 | N);
}

// Raw buffer descriptor over `bytes` bytes at `base` (stride 0, byte
// offsets; loads at or past `bytes` read zeros).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// LDS-DMA through a buffer descriptor: lane l's 16 bytes at byte offset
// voff + soff land at dst + 16 l (soff wave-uniform, so the per-chunk offset
// needs no vector register).  (The buffer form keeps the compiler's LDS-read
// waits partial: after a global_load_lds it waits for every outstanding LDS read.)
__device__ __forceinline__ void buf_to_lds(rsrc_t rs, uint32_t voff, uint32_t soff, unsigned char* dst) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_ptr_t)dst, 16, voff, soff, 0, 0);
}

// ---------------------------------------------------------------------------
// SNP-side kernel: functionals of every SNP row against a packed operand.
// MODE 0: fused association epilogue;  MODE 1: store s_j * f (RSVD sketch).
//
// Fill code.  Before the stream, each SNP gets an integer fill code c from
// its first 1024 samples (c = round(S mu_1024)), and the g' fragment decodes
// 00 -> 0, 01 -> c, 10 -> S, 11 -> 2S (one byte-permute table per lane).
// With q^f the g'-side quantisation (2 or 3 terms) and q^c the miss-side one
// (1 or 2 terms, the leading terms of the same expansion):
//   f_i = (1/S) sum_k q^f g'_k - mu sum_k q^f + (mu - c/S) sum_{miss} q^c
//       = sum_{obs} q^f (g - mu) + (mu - c/S) sum_{miss} (q^c - q^f),
// so the coarse miss side only meets the small weight |mu - c/S|.
//
// Workgroup: 4 waves x ST tiles of 32 SNPs, two workgroups per CU, so one
// workgroup's barriers and epilogue overlap the other's stream.  Blocks of
// 256 samples (8 k-steps of 32):
//   * the operand block (8 NT fragments) goes to a 2-slot LDS ring by
//     buffer-descriptor DMA, shared by the four waves;
//   * the 32 x 64 bytes of a tile's rows go to a one-block LDS slot of its
//     wave by the same DMA, 16 whole 64-byte row pieces per instruction;
//     each lane then reads back the 32 bytes of its own column and half.
// Rows past the last SNP, bytes past the row stride and blocks past the last
// read zeros through the buffer descriptor.  The 32x32x32 MFMA issues half
// as many instructions per MAC as 16x16x64, which leaves the VALU decode
// more issue cycles beside it.
// ---------------------------------------------------------------------------
struct SnpArgs {
  const uint8_t* packed;
  uint64_t ld, p, n;
  uint32_t rb;   // bytes of a row that are read: ceil(n/4) rounded up to 16 (the rest of the stride reads as zeros)
  int nblk, m, k3, km, g1, ones_g, norm, S, fill;
  const int8_t* A;
  const double* wtab;
  EpiArgs ea;
  double* Z;
  uint64_t ldZ;
  // MODE 2 (NMF backward pass): H^T rows updated in place, the operand's
  // column scales, W^T W (K x K) and the clamp eps
  double* Ht;
  const double* nmf_sc;
  const double* WtW;
  double eps;
  // MODE 3 (raw sums): int32 sum of packed row r for SNP j at raw[r * ldraw + j]
  int32_t* raw;
  uint64_t ldraw;
};

typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
constexpr int kSnpWaves = 4;

// Byte-permute table of the g' fragment from the SNP's first
// 256 kFillBlocks samples (lane half h holds samples 256 b + 128 h .. + 127
// of block b in w[b][0..1]; samples past n are not counted).  With the
// all-ones row in a g' tile the missing count is recovered from sum g' =
// S sg + c n_miss and popc = sg + n_miss, which needs c != S; the code then
// moves to the neighbour on the side of the estimate.  The code only sets
// how large the miss-side residual weight |mu - c/S| is: any c gives exact
// functionals.
constexpr int kFillBlocks = 4;
__device__ __forceinline__ uint32_t fill_table(const v4u (&w)[kFillBlocks][2], uint64_t n, int h, int S, int fill,
                                               int ones_g) {
  int sg = 0, miss = 0, valid = 0;
#pragma unroll
  for (int b = 0; b < kFillBlocks; ++b) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t x = w[b][i][v], lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
        sg += __popc(hi) + __popc(lo & hi);   // het + 2 hom
        miss += __popc(lo & ~hi);
      }
    const int64_t left = (int64_t)n - (256 * b + 128 * h);
    valid += left <= 0 ? 0 : (left >= 128 ? 128 : (int)left);
  }
  int obs = valid - miss;
  sg += __shfl_xor(sg, 32);
  obs += __shfl_xor(obs, 32);
  int c = obs > 0 ? (2 * S * sg + obs) / (2 * obs) : 0;   // round(S sg / obs)
  c = c < 2 * S ? c : 2 * S;
  if (ones_g && c == S) c = sg >= obs ? S + 1 : S - 1;
  if (!fill) c = 0;
  return ((uint32_t)(2 * S) << 24) | ((uint32_t)S << 16) | ((uint32_t)c << 8);
}

// One word = 16 codes; register v of a fragment takes the codes at bit
// 8c + 2v (one shift and mask), mapped through the lane's g' table and the
// missing-indicator table.
__device__ __forceinline__ Frag decode_fill(uint32_t w, uint32_t tbl, uint32_t& c_pop, uint32_t& c_hi) {
  c_pop += __popc(w);                  // dosage + 1 per missing call
  c_hi += __popc(w & 0xAAAAAAAAu);     // het + hom
  Frag f;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const uint32_t cv = (w >> (2 * v)) & 0x03030303u;
    f.g0[v] = (int)__builtin_amdgcn_perm(0u, tbl, cv);
    f.miss[v] = (int)__builtin_amdgcn_perm(0u, kPermMiss, cv);
  }
  return f;
}

struct SnpPipe {
  rsrc_t geno;        // this workgroup's SNP rows (exactly); rows past the last SNP read zeros
  rsrc_t opnd;        // the operand blocks
  uint32_t rowoff;    // byte offset of row lane / 8 of the wave's tile 0 in the workgroup's rows
  uint32_t tstride;   // 32 rows
  uint32_t istride;   // 8 rows
  uint32_t ub;        // byte of the lane's 16-byte unit in a 128-byte row piece (even instructions)
  uint32_t raddr[4];  // LDS addresses of the lane's pieces (b, k) of tile 0 in its genotype slot
};

// Genotype staging works on block pairs: pair q is the 128-byte piece
// 128 q .. 128 q + 127 of every row (samples 512 q .. + 511, blocks 2q and
// 2q + 1), so every DMA instruction moves whole 128-byte cache lines (rows
// start on 128-byte boundaries when they are 512 bytes or longer).  A stream
// of such pieces runs at 6.1 TB/s against 5.3 TB/s for 64-byte pieces.
//
// LDS-DMA of pair q of the wave's ST tiles into its 4 ST KB genotype slot.
// Instruction i of tile st moves rows 8 i .. 8 i + 7 of the tile: lane l
// brings unit u = (l % 8 - l / 16 - 4 i) mod 8 (16 bytes) of row 8 i + l / 8
// to slot + 1024 (4 st + i) + 16 l.  The unit rotation by row / 2 spreads
// the 16 lanes that later read one unit of 16 consecutive rows over all LDS
// banks.  Units at or past the row's bytes (rb) are addressed past the
// descriptor and land as zeros without a memory access.
template <int ST>
__device__ __forceinline__ void geno_pair_to_lds(const SnpArgs& a, const SnpPipe& pc, int q, unsigned char* dst) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t byte = (uint32_t)q * 128 + (pc.ub ^ (i & 1 ? 64u : 0u));
    const uint32_t o = byte < a.rb ? pc.rowoff + i * pc.istride + byte : 0xF0000000u;
#pragma unroll
    for (int st = 0; st < ST; ++st) buf_to_lds(pc.geno, o + (byte < a.rb ? st * pc.tstride : 0u), 0, dst + (4 * st + i) * 1024);
  }
}

// Byte offset in a genotype slot of piece (b, k) of tile 0 for lane
// (column c = lane % 32, half h = lane / 32): unit u = 4 b + 2 h + k of row
// c, i.e. samples 256 b + 128 h + 64 k .. + 63 of the pair.
__device__ __forceinline__ uint32_t geno_piece_offset(int lane, int b, int k) {
  const int c = lane & 31, h = lane >> 5, u = 4 * b + 2 * h + k;
  return 1024u * (c >> 3) + 16u * (8 * (c & 7) + ((u + (c >> 1)) & 7));
}

// The lane's 4 ST pieces of the pair in its genotype slot into registers:
// q[4 st + 2 b + k] = piece (b, k) of tile st (tile st at +4096 st).  The
// wait is inside the asm block, so no load is in flight when it ends.
template <int ST>
__device__ __forceinline__ void lds_read_pair(const SnpPipe& pc, v4i (&q)[4 * ST]) {
  static_assert(ST == 1 || ST == 2, "tile count");
  if constexpr (ST == 1) {
    asm volatile("ds_read_b128 %0, %4\n\tds_read_b128 %1, %5\n\tds_read_b128 %2, %6\n\tds_read_b128 %3, %7\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(q[0]), "=&v"(q[1]), "=&v"(q[2]), "=&v"(q[3])
                 : "v"(pc.raddr[0]), "v"(pc.raddr[1]), "v"(pc.raddr[2]), "v"(pc.raddr[3])
                 : "memory");
  } else {
    asm volatile("ds_read_b128 %0, %8\n\tds_read_b128 %1, %9\n\tds_read_b128 %2, %10\n\tds_read_b128 %3, %11\n\t"
                 "ds_read_b128 %4, %8 offset:4096\n\tds_read_b128 %5, %9 offset:4096\n\t"
                 "ds_read_b128 %6, %10 offset:4096\n\tds_read_b128 %7, %11 offset:4096\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(q[0]), "=&v"(q[1]), "=&v"(q[2]), "=&v"(q[3]), "=&v"(q[4]), "=&v"(q[5]), "=&v"(q[6]),
                   "=&v"(q[7])
                 : "v"(pc.raddr[0]), "v"(pc.raddr[1]), "v"(pc.raddr[2]), "v"(pc.raddr[3])
                 : "memory");
  }
}

// One block (b = kb % 2 of pair kb / 2).  Operand block kb + 1 goes to the
// other slot of the 2-slot operand ring.  On the first block of a pair the
// wave reads the pair's pieces from its genotype slot into registers (gc)
// and then reuses the slot for the next pair (the reads have completed
// before the DMA is issued).  A wave issues the operand DMA first and the
// genotype DMA (4 ST instructions, first blocks only) last, and loads
// complete in order, so at the top of the second block of a pair only the
// next pair's genotype DMA may stay in flight; at the top of a first block
// nothing may.
template <int NT0, int NT, int ST, int B>
__device__ __forceinline__ void snp_block(int kb, const SnpArgs& a, const SnpPipe& pc, int wave, int lane,
                                          const unsigned char* a_rd, unsigned char* a_wr, unsigned char* g_slot,
                                          v4i (&gc)[4 * ST], const uint32_t (&tbl)[ST], v16i (&acc)[ST][NT],
                                          uint32_t (&c_pop)[ST], uint32_t (&c_hi)[ST]) {
  constexpr int KS = 8, W = kSnpWaves, CH = KS * NT, b = B;   // b = kb % 2
  if (b == 1 && kb + 1 < a.nblk)
    wait_vmcnt<4 * ST>();
  else
    wait_vmcnt<0>();
  __builtin_amdgcn_s_barrier();   // block kb is in LDS; every wave is done with block kb - 1's operand slot
  if (b == 0) lds_read_pair<ST>(pc, gc);   // waits for its reads
  if (kb + 1 < a.nblk) {
    const uint32_t src = (uint32_t)(kb + 1) * CH * 1024;
#pragma unroll
    for (int i = 0; i < (CH + W - 1) / W; ++i) {
      const int c = wave + i * W;
      if (c < CH) buf_to_lds(pc.opnd, 16 * lane, src + c * 1024, a_wr + c * 1024);
    }
  }
  if (b == 0 && kb + 2 < a.nblk) geno_pair_to_lds<ST>(a, pc, kb / 2 + 1, g_slot);
  __builtin_amdgcn_sched_barrier(0);   // keep the prefetches at the top of the block
  // operand fragment c = s NT + rt is read kOperandPrefetch fragments ahead;
  // the genotype fragments of k-step s + 1 are decoded after the MFMAs of s
  constexpr int D = kOperandPrefetch < CH ? kOperandPrefetch : CH;
  const v4i* Ab = reinterpret_cast<const v4i*>(a_rd);
  v4i af[CH];
#pragma unroll
  for (int c = 0; c < D; ++c) af[c] = Ab[c * 64 + lane];
  Frag fc[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) fc[st] = decode_fill((uint32_t)gc[4 * st + 2 * b][0], tbl[st], c_pop[st], c_hi[st]);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
#pragma unroll
    for (int rt = 0; rt < NT; ++rt) {
      const int c = s * NT + rt;
      if (c + D < CH) af[c + D] = Ab[(c + D) * 64 + lane];
#pragma unroll
      for (int st = 0; st < ST; ++st)
        acc[st][rt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[c], rt < NT0 ? fc[st].g0 : fc[st].miss, acc[st][rt], 0,
                                                            0, 0);
    }
    if (s + 1 < KS) {
#pragma unroll
      for (int st = 0; st < ST; ++st)
        fc[st] = decode_fill((uint32_t)gc[4 * st + 2 * b + ((s + 1) >> 2)][(s + 1) & 3], tbl[st], c_pop[st],
                             c_hi[st]);
    }
  }
}

// Functionals of lane half h at positions q = 0 .. NQ-1 (i = h + 2 q) from
// the lane's own accumulators (slot layout of RowMap); wt = the weight table.
template <int NQ, int NT0, int NT1>
__device__ __forceinline__ void recombine(const v16i (&acc)[NT0 + NT1], int h, int k3, int km, double mu, double dmu,
                                          const double* wt, double (&fq)[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int i = h + 2 * q;
    const bool x = i >= k3, xm = i >= km;
    // X and Y are integers below 2^53, so these FP64 Horner steps are exact
    double X = fma((double)acc[(2 * q) / 16][(2 * q) % 16], 254.0, (double)acc[(2 * q + 1) / 16][(2 * q + 1) % 16]);
    // slots of term 2 (g' side) and term 1 (miss side) of columns >= k3; the
    // loop is unrolled, so the conditions and indices fold to constants
    const int s2 = 3 * NQ - 1 - q, s1 = 2 * NQ - 1 - q;
    if (s2 < 16 * NT0 && x) {
      const int s = s2 < 16 * NT0 ? s2 : 0;
      X = fma(X, 254.0, (double)acc[s / 16][s % 16]);
    }
    double Y = (double)acc[NT0 + q / 16][q % 16];
    if (s1 < 16 * NT1 && xm) {
      const int s = s1 < 16 * NT1 ? s1 : 0;
      Y = fma(Y, 254.0, (double)acc[NT0 + s / 16][s % 16]);
    }
    fq[q] = wt[i] * X - mu * wt[64 + i] + dmu * (wt[32 + i] * Y);   // zero weights past m
  }
}

// Bytes of the split plan for m <= 2 NQ functionals and d <= min(kMaxD, 2 NQ) traits.
__host__ __device__ constexpr int snp_plan_bytes(int NQ) {
  return 8 * (3 * (((2 * NQ < kMaxD ? 2 * NQ : kMaxD) + 3) & ~3) * 32 + (((2 * NQ < kMaxD ? 2 * NQ : kMaxD) + 3) & ~3) +
              4 * NQ * NQ + 1);
}

// NMF backward pass for the SNP of this lane's column (src/nmf.rs:171-221):
// with W's columns on the dosage side and the products W_r W_s on the
// missing side (NmfMap), the filled sum is
//   num_r = sum_k W_kr y_kj + sum_{k missing} W_kr (W_k . H_j)
//         = A_r + sum_s G_rs H_js,  A_r = sum_{k obs} W_kr g_kj,
//                                    G_rs = sum_{k missing} W_kr W_ks,
// so no per-call fill is formed.  One exchange per register gives both lane
// halves every row of the column; A and G are recombined in FP64 from the
// exact int32 sums of their terms, and the reference's update follows:
// den_r = (W^T W H_j)_r, H_jr *= num_r / den_r where den_r > eps, clamped to
// eps.  Half 0 stores.  ls: column scales at [0, K + P), W^T W at [64, ...).
template <int K, int NT0, int NT1>
__device__ __forceinline__ void nmf_bwd_epilogue(const SnpArgs& a, const v16i (&acc)[NT0 + NT1], const double* ls,
                                                 uint64_t j, int h) {
  constexpr int P = K * (K + 1) / 2;
  double val[K + P];
#pragma unroll
  for (int c = 0; c < K + P; ++c) val[c] = 0.0;
#pragma unroll
  for (int rt = 0; rt < NT0 + NT1; ++rt)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      uint32_t x = (uint32_t)acc[rt][v], y = x;
      const auto sw = __builtin_amdgcn_permlane32_swap(x, y, false, false);
      const int r0 = 32 * rt + 8 * (v >> 2) + (v & 3);   // lower half's row; the upper half's is r0 + 4
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int row = r0 + 4 * hh;
        const double iv = (double)(int)(hh ? sw[1] : sw[0]);
        if (row < 32 * NT0) {
          if (row < 3 * K) {
            const int c = row / 3, tm = row % 3;
            val[c] += (tm == 0 ? 1.0 : (tm == 1 ? 1.0 / 254.0 : 1.0 / 64516.0)) * iv;
          }
        } else {
          const int rr = row - 32 * NT0;
          if (rr < 2 * P) val[K + rr / 2] += (rr % 2 == 0 ? 1.0 : 1.0 / 254.0) * iv;
        }
      }
    }
  const bool valid = j < a.p && h == 0;
  double hj[K];
#pragma unroll
  for (int r = 0; r < K; ++r) hj[r] = valid ? a.Ht[j * K + r] : 0.0;
  const double* wtw = ls + 64;
#pragma unroll
  for (int r = 0; r < K; ++r) {
    double num = ls[r] * val[r];
#pragma unroll
    for (int s2 = 0; s2 < K; ++s2) {
      const int lo = r < s2 ? r : s2, hi = r < s2 ? s2 : r;
      const int pidx = lo * K - lo * (lo - 1) / 2 + (hi - lo);
      num += (ls[K + pidx] * val[K + pidx]) * hj[s2];
    }
    double den = 0.0;
#pragma unroll
    for (int s2 = 0; s2 < K; ++s2) den += wtw[r * K + s2] * hj[s2];
    double v = hj[r];
    if (den > a.eps) v *= num / den;
    if (valid) a.Ht[j * K + r] = v > a.eps ? v : a.eps;
  }
}

template <int NQ, int NT0, int NT1, int ST, int MODE>
__global__ void __launch_bounds__(64 * kSnpWaves, 2) k_snp_i8(SnpArgs a) {
  constexpr int NT = NT0 + NT1, KS = 8, W = kSnpWaves, CH = KS * NT;
  constexpr int A_SLOT = CH * 1024;
  constexpr int A0_BYTES = MODE == 0 && snp_plan_bytes(NQ) > A_SLOT ? snp_plan_bytes(NQ) : A_SLOT;
  static_assert(kWtab * 8 <= A_SLOT, "weight table must fit an operand slot");
  constexpr int G_SLOT = W * ST * 4096;
  __shared__ __align__(16) unsigned char A0[A0_BYTES];
  __shared__ __align__(16) unsigned char A1[A_SLOT];
  __shared__ __align__(16) unsigned char G0[G_SLOT];
  const int t = threadIdx.x, lane = t & 63, h = lane >> 5, col = lane & 31;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);   // wave-uniform: scalar DMA offsets
  const uint64_t wgbase = (uint64_t)blockIdx.x * (W * ST * 32);
  const uint64_t base = wgbase + (uint64_t)wave * (ST * 32);
  unsigned char* gslot = G0 + wave * (ST * 4096);   // the wave's genotype slot (one block pair)
  SnpPipe pc;
  const uint64_t wrows = a.p - wgbase < (uint64_t)(W * ST * 32) ? a.p - wgbase : (uint64_t)(W * ST * 32);
  pc.geno = make_rsrc(a.packed + wgbase * a.ld, (uint32_t)(wrows * a.ld));
  pc.opnd = make_rsrc(a.A, (uint32_t)a.nblk * CH * 1024);
  pc.rowoff = (uint32_t)((base - wgbase + (lane >> 3)) * a.ld);
  pc.tstride = (uint32_t)(32 * a.ld);
  pc.istride = (uint32_t)(8 * a.ld);
  pc.ub = 16u * (uint32_t)(((lane & 7) - (lane >> 4)) & 7);
#pragma unroll
  for (int e = 0; e < 4; ++e)
    pc.raddr[e] = (uint32_t)(uintptr_t)(lds_ptr_t)gslot + geno_piece_offset(lane, e >> 1, e & 1);
  const uint32_t coloff = (uint32_t)((base - wgbase + col) * a.ld);   // this lane's own column
  v16i acc[ST][NT];
#pragma unroll
  for (int st = 0; st < ST; ++st)
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[st][rt][e] = 0;
  uint32_t c_pop[ST], c_hi[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) c_pop[st] = c_hi[st] = 0;
  // prologue in the order the blocks issue them: operand 0, genotype pair 0
  for (int c = wave; c < CH; c += W) buf_to_lds(pc.opnd, 16 * lane, c * 1024, A0 + c * 1024);
  geno_pair_to_lds<ST>(a, pc, 0, gslot);
  // blocks 2 .. kFillBlocks - 1 of the lane's rows for the fill code (plain
  // loads; bytes past the row's bytes read zeros through the descriptor)
  static_assert(kFillBlocks >= 2, "the fill code reads block pair 0 from the genotype slot");
  v4u fw[ST][kFillBlocks][2];
#pragma unroll
  for (int st = 0; st < ST; ++st)
#pragma unroll
    for (int b = 2; b < kFillBlocks; ++b)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const uint32_t byte = (uint32_t)b * 64 + 32 * h + 16 * i;
        const uint32_t o = byte < a.rb ? coloff + byte + st * pc.tstride : 0xF0000000u;
        fw[st][b][i] = __builtin_bit_cast(v4u, __builtin_amdgcn_raw_buffer_load_b128(pc.geno, o, 0, 0));
      }
  wait_vmcnt<0>();
  uint32_t tbl[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) {
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        fw[st][b][i] = *reinterpret_cast<const v4u*>(gslot + 4096 * st + geno_piece_offset(lane, b, i));
    tbl[st] = fill_table(fw[st], a.n, h, a.S, a.fill, a.ones_g);
  }
  v4i gc[4 * ST];   // the current pair's pieces
  for (int kb = 0; kb < a.nblk; kb += 2) {
    snp_block<NT0, NT, ST, 0>(kb, a, pc, wave, lane, A0, A1, gslot, gc, tbl, acc, c_pop, c_hi);
    if (kb + 1 < a.nblk) snp_block<NT0, NT, ST, 1>(kb + 1, a, pc, wave, lane, A1, A0, gslot, gc, tbl, acc, c_pop, c_hi);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if constexpr (MODE == 3) {
    // register e of tile rt holds row 32 rt + 8 (e >> 2) + 4 h + (e & 3) of SNP base + col
    const uint64_t j = base + col;
    if (j < a.p)
#pragma unroll
      for (int rt = 0; rt < NT; ++rt)
#pragma unroll
        for (int e = 0; e < 16; ++e)
          a.raw[(uint64_t)(32 * rt + 8 * (e >> 2) + 4 * h + (e & 3)) * a.ldraw + j] = acc[0][rt][e];
    return;
  }
  if constexpr (MODE == 2) {
    double* ls = reinterpret_cast<double*>(A1);   // column scales [0, K + P), W^T W at [64, 64 + K^2)
    const int mq = NQ + NQ * (NQ + 1) / 2;
    for (int e = t; e < mq; e += 64 * W) ls[e] = a.nmf_sc[e];
    for (int e = t; e < NQ * NQ; e += 64 * W) ls[64 + e] = a.WtW[e];
    __syncthreads();
    nmf_bwd_epilogue<NQ, NT0, NT1>(a, acc[0], ls, base + col, h);
    return;
  }

  // ---- epilogue: the plan goes to A0, the weight table to A1 ----------------
  const double* plan = nullptr;
  if constexpr (MODE == 0) {
    double* sp = reinterpret_cast<double*>(A0);
    stage_split_plan(a.ea.plan, a.ea.d, a.ea.m, sp, t, 64 * W);
    plan = sp;
  }
  double* wt = reinterpret_cast<double*>(A1);
  for (int e = t; e < kWtab; e += 64 * W) wt[e] = a.wtab[e];
  __syncthreads();
  double fq[ST][NQ];
  int64_t nmiss[ST], sg[ST], nhom[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) {
    uint32_t cp = c_pop[st], ch = c_hi[st];
    cp += __shfl_xor(cp, 32);
    ch += __shfl_xor(ch, 32);
    // the all-ones row: the missing count (miss tile) or sum g' = S sg + c n_miss (g' tile)
    const int ones_m = __shfl(acc[st][NT - 1][15], 32 * a.g1 + col);
    const int ones_g = __shfl(acc[st][NT0 - 1][15], 32 * a.g1 + col);
    const int cfill = (int)((tbl[st] >> 8) & 0xFFu);
    nmiss[st] = a.ones_g ? ((int64_t)a.S * cp - ones_g) / (a.S - cfill) : (int64_t)ones_m;
    sg[st] = (int64_t)cp - nmiss[st];
    nhom[st] = sg[st] - (int64_t)ch;
    const int64_t nobs = (int64_t)a.n - nmiss[st];
    const double mu = nobs > 0 ? (double)sg[st] / (double)nobs : 0.0;
    const double dmu = mu - (double)cfill / (double)a.S;
    recombine<NQ, NT0, NT1>(acc[st], h, a.k3, a.km, mu, dmu, wt, fq[st]);
    if constexpr (MODE == 1) {
      const uint64_t j = base + 32 * st + col;
      const double scale = epi_scale(mu, a.norm);
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (j < a.p && h + 2 * q < a.m) a.Z[j * a.ldZ + h + 2 * q] = scale * fq[st][q];
    }
  }
  if constexpr (MODE == 0) {
    // One SNP per lane: functional i = h + 2 q of tile st sits in lane half h,
    // so one exchange between the halves per position gives lane half h every
    // functional of its column of tile h (two tiles), or of tile 0 in both
    // halves (one tile; half 1 then stores nothing).  The per-SNP arithmetic
    // is the same either way.
    double f[2 * NQ];
    int tile = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double x = fq[0][q], y = fq[ST - 1][q];
      swap_halves(x, y);
      f[2 * q] = x;       // even functionals: lower half's values
      f[2 * q + 1] = y;   // odd functionals: upper half's values
    }
    if constexpr (ST == 2) tile = h;
    const uint64_t j = base + 32 * tile + col;
    const bool valid = j < a.p && (ST == 2 || h == 0);
    const int64_t nm = tile ? nmiss[ST - 1] : nmiss[0], s = tile ? sg[ST - 1] : sg[0],
                  nh = tile ? nhom[ST - 1] : nhom[0];
    snp_epilogue_lane<2 * NQ>(a.ea, plan, j, valid, f, nm, s, nh);
  }
}

// (NQ, NT0, NT1) of every RowMap the passes use with m <= 32: association
// RowMap::make(K + d, K, K) for d <= 16, K >= 1; RSVD RowMap::make(l, l, l).
#define SLFMM_SNP_ASSOC(X)                                                                                     \
  X(1, 1, 1) X(2, 1, 1) X(3, 1, 1) X(4, 1, 1) X(5, 1, 1) X(6, 1, 1) X(6, 2, 1) X(7, 1, 1) X(7, 2, 1) X(8, 2, 1) \
  X(9, 2, 1) X(9, 2, 2) X(10, 2, 1) X(10, 2, 2) X(11, 2, 1) X(11, 2, 2) X(12, 2, 1) X(12, 2, 2) X(12, 3, 2)     \
  X(13, 2, 1) X(13, 2, 2) X(13, 3, 2) X(14, 2, 1) X(14, 2, 2) X(14, 3, 2) X(15, 2, 1) X(15, 2, 2) X(15, 3, 2)   \
  X(16, 3, 2)
#define SLFMM_SNP_RSVD(X)                                                                                      \
  X(1, 1, 1) X(2, 1, 1) X(3, 1, 1) X(4, 1, 1) X(5, 1, 1) X(6, 1, 1) X(7, 1, 1) X(8, 1, 1) X(9, 2, 1) X(10, 2, 1) \
  X(11, 2, 1) X(12, 2, 1) X(13, 2, 1) X(14, 2, 1) X(15, 2, 1) X(16, 2, 1) X(16, 2, 2)

// 32-SNP tiles per wave: at most 2, and 2 only with at most 3 row tiles
// (accumulator registers)
template <int NT>
constexpr int snp_tiles_max() { return NT <= 3 ? 2 : 1; }

// Layouts of up to 3 row tiles run 2 tiles per wave when that grid holds at
// least tuning.snp_two_tiles_min_blocks workgroup blocks (workgroups x
// 256-sample blocks), else 1 (twice the workgroups, for short panels or
// short rows); the per-SNP arithmetic is the same.
static void launch_snp(slfmm_ctx* ctx, const Operand& op, SnpArgs& a, int mode) {
  if (a.p == 0) return;
  hipStream_t s = ctx->stream;
  const RowMap& r = op.rmap;
  const int NT = r.NT0 + r.NT1;
  const uint64_t blocks2 = cdiv(a.p, 2 * 32 * kSnpWaves) * (uint64_t)a.nblk;
  const int st = NT <= 3 && blocks2 >= (uint64_t)ctx->tuning.snp_two_tiles_min_blocks ? 2 : 1;
  const uint64_t rows = 32 * (uint64_t)st * kSnpWaves;
  // buffer descriptors address a workgroup's rows and the operand with 32-bit byte offsets
  SLFMM_CHECK((rows + 2 * 32) * a.ld < 0xF0000000ull && (uint64_t)a.nblk * 8 * NT * 1024 < 0xFFFFFFFFull,
              SLFMM_E_ARG, "panel rows too long for the int8 kernels (use SLFMM_PREC_FP64)");
  a.m = r.m; a.k3 = r.k3; a.km = r.km; a.g1 = r.g1; a.ones_g = r.ones_g; a.S = op.S; a.fill = op.fill;
  a.rb = (uint32_t)std::min<uint64_t>(rup(cdiv(a.n, 4), 16), a.ld);
  const unsigned grid = (unsigned)cdiv(a.p, rows);
  bool done = false;
#define X(Q, A0_, A1_)                                                                                        \
  if (!done && r.NQ == Q && r.NT0 == A0_ && r.NT1 == A1_) {                                                   \
    if (st == 2)                                                                                              \
      k_snp_i8<Q, A0_, A1_, snp_tiles_max<A0_ + A1_>(), MODE_><<<grid, 64 * kSnpWaves, 0, s>>>(a);            \
    else                                                                                                      \
      k_snp_i8<Q, A0_, A1_, 1, MODE_><<<grid, 64 * kSnpWaves, 0, s>>>(a);                                     \
    done = true;                                                                                              \
  }
#define MODE_ 0
  if (mode == 0) { SLFMM_SNP_ASSOC(X) }
#undef MODE_
#define MODE_ 1
  if (mode == 1) { SLFMM_SNP_RSVD(X) }
#undef MODE_
#undef X
  SLFMM_CHECK(done, SLFMM_E_ARG, "no int8 kernel for this operand layout");
  SLFMM_HIP(hipGetLastError());
}

// Layout diagnostic (host only): dims = {NQ, NT0, NT1, g1, ones_g}; rows
// (2 x 32 (NT0 + NT1) ints, if non-null) = (column, term) of every packed row.
void i8_layout(int m, int k3, int km, int32_t* dims, int32_t* rows, uint64_t cap) {
  SLFMM_CHECK(m >= 1 && m <= 32 && k3 >= 0 && k3 <= m && km >= 0 && km <= m, SLFMM_E_ARG, "operand shape");
  const RowMap r = RowMap::make(m, k3, km);
  dims[0] = r.NQ; dims[1] = r.NT0; dims[2] = r.NT1; dims[3] = r.g1; dims[4] = r.ones_g;
  if (!rows) return;
  const int nrow = 32 * (r.NT0 + r.NT1);
  SLFMM_CHECK(cap >= (uint64_t)(2 * nrow), SLFMM_E_ARG, "rows buffer too small");
  for (int row = 0; row < nrow; ++row) {
    const int2 v = r.info(row);
    rows[2 * row] = v.x;
    rows[2 * row + 1] = v.y;
  }
}

// Association pass over every SNP of g (fused epilogue).
void launch_func_i8(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_Q, uint64_t m, const EpiArgs& ea,
                    bool /*fused_epilogue*/) {
  // Latent directions: 2 int8 terms on the g' side, 1 on the miss side;
  // covariate directions (the t-statistic numerators, last d columns): 3 on
  // the g' side, 2 on the miss side.  The miss side meets only the weight
  // |mu - c/S|, and the fill code c comes from the SNP's first 1024 samples
  // on a scale S = 32 (n <= 262144), which keeps that weight small enough
  // for one latent miss term (a second one costs a row tile: 7 instead of
  // 6 at K + d = 24).
  SLFMM_CHECK(m <= 32, SLFMM_E_ARG, "the int8 association path supports K + d <= 32 (use SLFMM_PREC_FP64)");
  Operand op = prepare_operand(ctx, "as_i8", d_Q, m, g->n, (int)m, ea.K, ea.K, true);
  SLFMM_CHECK(8 * split_plan_len(ea.d, (int)m) <= snp_plan_bytes(op.rmap.NQ), SLFMM_E_ARG,
              "association plan does not fit the int8 kernel's LDS (use SLFMM_PREC_FP64)");
  SnpArgs a{};
  a.packed = g->d; a.ld = g->ld; a.p = g->p; a.n = g->n; a.nblk = (int)op.nblk; a.norm = ea.norm;
  a.A = op.pack; a.wtab = op.wtab; a.ea = ea;
  slfmm_timed tm(ctx, "assoc_stream");
  launch_snp(ctx, op, a, 0);
}

// RSVD sketch / backward pass: Z[j][c] = s_j * (B^T y_j)_c for every SNP row.
void rsvd_project_i8(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p, const double* d_B,
                     uint64_t l, int norm, double* Z, uint64_t ldZ) {
  SLFMM_CHECK(l >= 1 && l <= 32, SLFMM_E_ARG, "the int8 RSVD path supports K + oversampling <= 32");
  Operand op = prepare_operand(ctx, "rs_i8", d_B, l, n, (int)l, (int)l, (int)l, true);
  SnpArgs a{};
  a.packed = packed; a.ld = ld; a.p = p; a.n = n; a.nblk = (int)op.nblk; a.norm = norm;
  a.A = op.pack; a.wtab = op.wtab; a.Z = Z; a.ldZ = ldZ;
  launch_snp(ctx, op, a, 1);
}

// ---- NMF backward pass on the int8 matrix cores --------------------------------
// [W | W_r W_s] (n x (K + P)), one thread per sample (pairs r <= s in NmfMap order)
__global__ void k_nmf_wprod(const double* __restrict__ W, uint64_t n, int K, double* __restrict__ Qn, int mq) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* w = W + k * K;
  double* q = Qn + k * mq;
  int p = K;
  for (int r = 0; r < K; ++r) {
    q[r] = w[r];
    for (int s2 = r; s2 < K; ++s2) q[p++] = w[r] * w[s2];
  }
}

// H^T (pe x K) <- the reference's H update from W (n x K) and W^T W on the
// rows Y (pe SNPs, row stride ld), K <= kNmfMaxK8.
void nmf_backward_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t pe, const double* W, int K,
                     const double* WtW, double eps, double* Ht) {
  SLFMM_CHECK(K >= 1 && K <= kNmfMaxK8, SLFMM_E_ARG, "the int8 NMF pass supports rank 1..8");
  SLFMM_CHECK(n <= 8000000, SLFMM_E_ARG, "too many samples for the int8 kernels");
  if (pe == 0) return;
  hipStream_t s = ctx->stream;
  const NmfMap map = NmfMap::make(K);
  const int mq = K + map.P, NT = map.NT0 + map.NT1;
  double* Qn = ctx->buf<double>("nmf8_q", n * mq);
  k_nmf_wprod<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(W, n, K, Qn, mq);
  double* sc = ctx->buf<double>("nmf8_sc", mq);
  colabsmax_scale(ctx, Qn, mq, n, mq, nullptr, nullptr, sc);
  const uint64_t nblk = cdiv(n, 256);
  const uint64_t total = nblk * 8 * NT * 1024;
  int8_t* pack = ctx->buf<int8_t>("nmf8_pack", total);
  k_pack_rows<<<(unsigned)std::min<uint64_t>(cdiv(total, 256), 16384), 256, 0, s>>>(Qn, mq, n, map, sc, NT, total,
                                                                                  pack);
  SnpArgs a{};
  a.packed = Y; a.ld = ld; a.p = pe; a.n = n; a.nblk = (int)nblk;
  a.rb = (uint32_t)std::min<uint64_t>(rup(cdiv(n, 4), 16), ld);
  a.m = mq; a.S = 1; a.fill = 0; a.A = pack;
  a.Ht = Ht; a.nmf_sc = sc; a.WtW = WtW; a.eps = eps;
  const uint64_t rows = 32 * kSnpWaves;   // one tile per wave
  SLFMM_CHECK((rows + 2 * 32) * a.ld < 0xF0000000ull && total < 0xFFFFFFFFull, SLFMM_E_ARG,
              "panel rows too long for the int8 kernels");
  const unsigned grid = (unsigned)cdiv(pe, rows);
  switch (K) {
#define SLFMM_NMF8(KK)                                                                                        \
  case KK:                                                                                                    \
    k_snp_i8<KK, (3 * KK + 31) / 32, (KK * (KK + 1) + 31) / 32, 1, 2><<<grid, 64 * kSnpWaves, 0, s>>>(a);     \
    break;
    SLFMM_NMF8(1) SLFMM_NMF8(2) SLFMM_NMF8(3) SLFMM_NMF8(4) SLFMM_NMF8(5) SLFMM_NMF8(6) SLFMM_NMF8(7)
    SLFMM_NMF8(8)
#undef SLFMM_NMF8
  }
  SLFMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// RSVD forward pass on the sample-major copy: W = Y Q_z (n x l).
//   W[k][c] = sum_j g0_jk Qt_jc + sum_j miss_jk (mu Qt)_jc - sum_j mu_j Qt_jc
// with Qt = diag(s) Q_z.  A side: genotype fragments of 16 samples x 16*KS
// SNPs; B side: Qt and mu*Qt as 2 int8 terms per column, pre-permuted.
// Workgroup: 4 waves x 4 sample tiles (256 samples) x one SNP chunk.
// ---------------------------------------------------------------------------
// (fragment counts fit 32 bits and NCB, KS are compile-time, so the index
// arithmetic needs no 64-bit division)
template <int NCB, int KS>
__global__ void k_pack_cols(const double* __restrict__ Qz, uint64_t ldq, uint64_t pe, int l,
                            const double* __restrict__ ssnp, const double* __restrict__ mu,
                            const double* __restrict__ csc, uint32_t nsnp, int8_t* __restrict__ out) {
  // One thread per (SNP, column) of the padded block grid (nsnp = nblk 64 KS rows,
  // 16 NCB columns): both parts (Qt, mu Qt) and both terms are written from
  // one load of Q_z.  Byte (part, term) goes to fragment ct = part 2 NCB +
  // term NCB + col / 16 of k-step (blk, s), lane g 16 + col % 16, byte j.
  constexpr int NTB = 4 * NCB, NC = 16 * NCB;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < nsnp * NC; e += gridDim.x * blockDim.x) {
    const int col = (int)(e % NC);
    const uint32_t snp = e / NC;
    const uint32_t blk = snp / (64 * KS), r = snp % (64 * KS);
    const int g = (int)(r / (16 * KS)), s = (int)((r / 16) % KS), r3 = (int)(r % 16);
    const int j = (r3 & 3) * 4 + (r3 >> 2);
    int q[2][2] = {{0, 0}, {0, 0}};
    if (col < l && snp < pe) {
      const double x0 = Qz[(uint64_t)snp * ldq + col] * ssnp[snp];
      const double x1 = x0 * mu[snp];
      const double y0 = x0 / csc[col], y1 = x1 / csc[l + col];
      q[0][0] = quant_term(y0, 0); q[0][1] = quant_term(y0, 1);
      q[1][0] = quant_term(y1, 0); q[1][1] = quant_term(y1, 1);
    }
    const uint64_t base = (uint64_t)(blk * KS + s) * NTB * 1024 + (uint64_t)(g * 16 + col % 16) * 16 + j;
#pragma unroll
    for (int part = 0; part < 2; ++part)
#pragma unroll
      for (int term = 0; term < 2; ++term)
        out[base + (uint64_t)(part * 2 * NCB + term * NCB + col / 16) * 1024] = (int8_t)q[part][term];
  }
}

// vhat_c = sum_j mu_j * Qt^_jc (the 2-term value of Qt), column reduction.
struct FnVhat {
  const double* Qz; uint64_t ldq; const double* ssnp; const double* mu; const double* csc;
  __device__ void operator()(uint64_t r, int c, double (&v)[1]) const {
    const double q = Qz[r * ldq + c] * ssnp[r] / csc[c];
    v[0] = mu[r] * (csc[c] * ((double)quant_term(q, 0) + (double)quant_term(q, 1) / 254.0));
  }
};
struct GStore1 {
  double* out;
  __device__ void operator()(int c, const double (&v)[1]) const { out[c] = v[0]; }
};

// Epi::WV waves x 4 tiles of 16 samples per workgroup share each staged B
// block: 8 for the RSVD pass, 4 for the NMF pass (two workgroups per CU,
// whose barriers interleave; measured 1.09 vs 1.19 ms per NMF forward pass
// at n = 5000, p = 2M, and 0.90 vs 0.87 ms per RSVD fit the other way).
// Epilogue of the RSVD forward pass: the 2 terms of Qt and of mu Qt of
// column 16 cb + l16 sit in tiles cb, NCB + cb, 2 NCB + cb, 3 NCB + cb.
template <int NCB>
struct FwdRsvdEpi {
  static constexpr int NTB = 4 * NCB, NG = 2 * NCB;   // tiles; the first NG meet the dosage fragment
  static constexpr int WV = 8;                        // waves per workgroup
  const double* csc; int l; double* part; uint64_t ldp;
  template <int ST>
  __device__ void operator()(const v4i (&acc)[ST][NTB], uint64_t sbase, uint64_t n, int g, int l16,
                             uint64_t chunk) const {
    // lane-local combination: output (sample 4g + r of tile st, column 16 cb + l16)
#pragma unroll
    for (int st = 0; st < ST; ++st)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const int col = 16 * cb + l16;
        if (col >= l) continue;
        const double w0 = csc[col], w1 = csc[col] / 254.0, m0 = csc[l + col], m1 = csc[l + col] / 254.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint64_t k = sbase + 16 * st + 4 * g + r;
          if (k >= n) continue;
          double v = w0 * (double)acc[st][cb][r] + w1 * (double)acc[st][NCB + cb][r] +
                     m0 * (double)acc[st][2 * NCB + cb][r] + m1 * (double)acc[st][3 * NCB + cb][r];
          part[(chunk * n + k) * ldp + col] = v;
        }
      }
  }
};

// Forward pass on the sample-major copy with a tile plan: tiles [0, NG)
// meet the dosage fragment, [NG, NTB) the missing-call fragment; Epi
// combines the int32 sums of a workgroup's SNP chunk into its partial row.
// Grid: (sample blocks of 64 Epi::WV, SNP chunks of nblk_chunk blocks).
template <class Epi>
__global__ void __launch_bounds__(64 * Epi::WV) k_fwd_i8(const uint8_t* __restrict__ YT, uint64_t ldt, uint64_t n,
                                                           int nblk_chunk, int nblk_total, const v4i* __restrict__ Bp,
                                                           Epi epi) {
  constexpr int NTB = Epi::NTB, NG = Epi::NG, WV = Epi::WV;
  constexpr int KS = 4;
  constexpr int ST = 4;
  constexpr int CH = KS * NTB;   // 1 KB B chunks (64 lanes x 16 B) per block
  constexpr int BLK = CH * 64;
  __shared__ v4i Bs[2 * BLK];
  const int bx = (int)blockIdx.x, by = (int)blockIdx.y;
  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, l16 = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);   // wave-uniform: scalar DMA offsets
  const uint64_t sbase = (uint64_t)bx * (ST * 16 * WV) + wave * (ST * 16);
  const int b0 = by * nblk_chunk;
  const int b1 = min(nblk_total, b0 + nblk_chunk);
  const uint8_t* rowp[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) {
    uint64_t k = sbase + 16 * st + l16;
    if (k >= n) k = n - 1;
    rowp[st] = YT + k * ldt + 4 * KS * g;
  }
  v4i acc[ST][NTB];
#pragma unroll
  for (int st = 0; st < ST; ++st)
#pragma unroll
    for (int ct = 0; ct < NTB; ++ct) acc[st][ct] = v4i{0, 0, 0, 0};
  uint32_t dummy0 = 0, dummy1 = 0, dummy2 = 0;
  // B blocks go straight from global memory into the LDS ring by DMA: the
  // waves take the block's chunks in turn (byte offsets < 2^32, checked by
  // the launchers)
  const rsrc_t brs = make_rsrc(Bp, (uint32_t)((uint64_t)nblk_total * BLK * 16));
  auto stage_b = [&](int kb, int slot) {
    unsigned char* dst = reinterpret_cast<unsigned char*>(Bs + slot * BLK);
    for (int c = wave; c < CH; c += WV)
      buf_to_lds(brs, 16 * lane, (uint32_t)(((uint64_t)kb * CH + c) * 1024), dst + c * 1024);
  };
  uint32_t gw[ST][KS], gn[ST][KS];
  if (b0 < b1) {
    stage_b(b0, 0);
#pragma unroll
    for (int st = 0; st < ST; ++st) load_words<KS>(rowp[st] + (uint64_t)b0 * 16 * KS, gw[st]);
  }
  wait_vmcnt<0>();
  __syncthreads();
  for (int kb = b0; kb < b1; ++kb) {
    const bool more = kb + 1 < b1;
    if (more) {
      stage_b(kb + 1, (kb + 1 - b0) & 1);
#pragma unroll
      for (int st = 0; st < ST; ++st) load_words<KS>(rowp[st] + (uint64_t)(kb + 1) * 16 * KS, gn[st]);
    }
    if (kb == nblk_total - 1) {
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int64_t lim = (int64_t)ldt - ((int64_t)kb * 16 * KS + 4 * KS * g + 4 * s);
        const uint32_t mask = lim >= 4 ? 0xFFFFFFFFu : (lim <= 0 ? 0u : ((1u << (8 * lim)) - 1u));
#pragma unroll
        for (int st = 0; st < ST; ++st) gw[st][s] &= mask;
      }
    }
    const v4i* Bb = Bs + ((kb - b0) & 1) * BLK;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      v4i bf[NTB];
#pragma unroll
      for (int ct = 0; ct < NTB; ++ct) bf[ct] = Bb[(s * NTB + ct) * 64 + lane];
#pragma unroll
      for (int st = 0; st < ST; ++st) {
        Frag f = decode_word(gw[st][s], dummy0, dummy1);
#pragma unroll
        for (int ct = 0; ct < NG; ++ct)
          acc[st][ct] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f.g0, bf[ct], acc[st][ct], 0, 0, 0);
#pragma unroll
        for (int ct = NG; ct < NTB; ++ct)
          acc[st][ct] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f.miss, bf[ct], acc[st][ct], 0, 0, 0);
      }
    }
    if (more) {
#pragma unroll
      for (int st = 0; st < ST; ++st)
#pragma unroll
        for (int s = 0; s < KS; ++s) gw[st][s] = gn[st][s];
    }
    wait_vmcnt<0>();   // this wave's B DMA has landed
    __syncthreads();
  }
  (void)dummy0; (void)dummy1; (void)dummy2;
  epi(acc, sbase, n, g, l16, (uint64_t)by);
}

// SNP blocks per chunk of a k_fwd_i8 launch: about fwd_target_wg
// workgroups, rounded to whole rounds of the workgroups the device holds at
// once (CUs x the kernel's occupancy), so no round runs a few stragglers;
// at least min_chunks chunks.
template <class Kern>
static uint64_t fwd_blocks_per_chunk(slfmm_ctx* ctx, Kern kern, int threads, uint64_t nblk, uint64_t sblocks,
                                     uint64_t min_chunks) {
  uint64_t& resident = ctx->resident_wgs[reinterpret_cast<const void*>(kern)];
  if (resident == 0) {
    int cus = 0, occ = 0;
    SLFMM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    SLFMM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, 0));
    resident = (uint64_t)std::max(cus, 1) * (uint64_t)std::max(occ, 1);
  }
  const uint64_t target = std::max<uint32_t>(1u, ctx->tuning.fwd_target_wg);
  const uint64_t rounds = std::max<uint64_t>(1, (target + resident / 2) / resident);
  uint64_t nchunk = std::max<uint64_t>(1, rounds * resident / sblocks);
  nchunk = std::min(std::max(nchunk, min_chunks), nblk);
  return cdiv(nblk, nchunk);
}

__global__ void k_fwd_reduce(const double* __restrict__ part, uint64_t n, int l, uint64_t ldp, int nchunk,
                             const double* __restrict__ vhat, double* W, uint64_t ldw) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * (uint64_t)l) return;
  uint64_t k = e / l;
  int c = (int)(e % l);
  double s = 0.0;
  for (int q = 0; q < nchunk; ++q) s += part[((uint64_t)q * n + k) * ldp + c];
  W[k * ldw + c] = s - vhat[c];
}

void rsvd_accumulate_i8(slfmm_ctx* ctx, const uint8_t* YT, uint64_t ldt, uint64_t n, uint64_t pe, const double* Qz,
                        uint64_t ldq, uint64_t l, const double* mu, const double* ssnp, double* W, uint64_t ldw) {
  hipStream_t s = ctx->stream;
  const int NCB = (int)cdiv(l, 16);
  SLFMM_CHECK(NCB >= 1 && NCB <= 2, SLFMM_E_ARG, "int8 forward pass supports K + oversampling <= 32");
  const int KS = 4, NTB = 4 * NCB;
  double* csc = ctx->buf<double>("fw_csc", 2 * l);
  colabsmax_scale(ctx, Qz, ldq, pe, l, ssnp, mu, csc);   // csc[c]: Qt, csc[l + c]: mu Qt
  const uint64_t nblk = cdiv(pe, 64 * KS);
  const uint64_t total = nblk * KS * NTB * 1024;
  int8_t* Bp = ctx->buf<int8_t>("fw_pack", total);
  SLFMM_CHECK(total < 0xFFFFFFFFull, SLFMM_E_ARG, "estimation subset too large for the int8 forward pass");
  const uint64_t nsnp = nblk * 64 * KS;   // every (SNP, column) slot of the padded blocks, 16 NCB columns
  const unsigned pgrid = (unsigned)std::min<uint64_t>(cdiv(nsnp * 16 * NCB, 256), 16384);
  if (NCB == 1)
    k_pack_cols<1, 4><<<pgrid, 256, 0, s>>>(Qz, ldq, pe, (int)l, ssnp, mu, csc, (uint32_t)nsnp, Bp);
  else
    k_pack_cols<2, 4><<<pgrid, 256, 0, s>>>(Qz, ldq, pe, (int)l, ssnp, mu, csc, (uint32_t)nsnp, Bp);
  double* vhat = ctx->buf<double>("fw_vhat", l);
  colreduce<1, kColSum>(ctx, FnVhat{Qz, ldq, ssnp, mu, csc}, GStore1{vhat}, pe, (int)l);
  constexpr int WV = FwdRsvdEpi<1>::WV;
  const uint64_t sblocks = cdiv(n, 64 * WV);
  // a chunk's int32 sums stay exact below 2^31 / 254 SNPs (|q| <= 127, dosage <= 2)
  const uint64_t per = NCB == 1
                           ? fwd_blocks_per_chunk(ctx, k_fwd_i8<FwdRsvdEpi<1>>, 64 * WV, nblk, sblocks, cdiv(pe, 8u << 20))
                           : fwd_blocks_per_chunk(ctx, k_fwd_i8<FwdRsvdEpi<2>>, 64 * WV, nblk, sblocks, cdiv(pe, 8u << 20));
  const uint64_t nchunk = cdiv(nblk, per);
  const uint64_t ldp = l;
  double* part = ctx->buf<double>("fw_part", nchunk * n * ldp);
  const dim3 grid((unsigned)sblocks, (unsigned)nchunk);
  if (NCB == 1)
    k_fwd_i8<<<grid, 64 * WV, 0, s>>>(YT, ldt, n, (int)per, (int)nblk, reinterpret_cast<const v4i*>(Bp),
                                     FwdRsvdEpi<1>{csc, (int)l, part, ldp});
  else
    k_fwd_i8<<<grid, 64 * WV, 0, s>>>(YT, ldt, n, (int)per, (int)nblk, reinterpret_cast<const v4i*>(Bp),
                                     FwdRsvdEpi<2>{csc, (int)l, part, ldp});
  k_fwd_reduce<<<(unsigned)cdiv(n * l, 256), 256, 0, s>>>(part, n, (int)l, ldp, (int)nchunk, vhat, W, ldw);
  SLFMM_HIP(hipGetLastError());
}

// ---- NMF-filled functionals on the int8 matrix cores -------------------------------
// With the missing calls of SNP j filled by W_k . H_j (NMF imputation,
// src/parallel.rs:252-321), the centred functional of operand column c is
//   F_jc = sum_{k obs} A_kc g_kj + sum_r H_jr M_jcr - mu_j sum_k A_kc,
//   M_jcr = sum_{k missing at j} A_kc W_kr,
// so the dosage fragment meets A (3 int8 terms) and the missing-call
// fragment the products A_c W_r (2 terms): one SNP-side pass (mode 3) per
// group of columns writes the exact int32 sums of every packed row, and
// k_fill_combine forms F in FP64 with the SNP's own H_j.  Operand of a group
// of g columns (NmfMap with K -> g, P -> g K): [A_c | A_c W_r (c, r)].
__global__ void k_fill_prod(const double* __restrict__ A, uint64_t lda, uint64_t n, int c0, int g,
                            const double* __restrict__ W, int K, double* __restrict__ Qn) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int mq = g + g * K;
  double* q = Qn + k * mq;
  for (int c = 0; c < g; ++c) {
    const double a = A[k * lda + c0 + c];
    q[c] = a;
    for (int r = 0; r < K; ++r) q[g + c * K + r] = a * W[k * K + r];
  }
}

// F[j][c0 + c] (times outscale[j] if given) from the raw sums of one group;
// one thread per SNP (its H_j read once), consecutive threads on consecutive
// SNPs for every raw row.
__global__ void k_fill_combine(const int32_t* __restrict__ raw, uint64_t ldraw, uint64_t p, NmfMap map, int c0,
                               const double* __restrict__ sc, const double* __restrict__ Ht,
                               const double* __restrict__ mu, const double* __restrict__ colsum,
                               const double* __restrict__ outscale, double* __restrict__ F, uint64_t ldF) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p) return;
  const int g = map.K, K = map.P / map.K;
  double h[kNmfMaxK8];
  for (int r = 0; r < K; ++r) h[r] = Ht[j * K + r];
  const double mj = mu[j], os = outscale ? outscale[j] : 1.0;
  auto R = [&](int row) { return (double)raw[(uint64_t)row * ldraw + j]; };
  const int m0 = 32 * map.NT0;
  for (int c = 0; c < g; ++c) {
    double f = sc[c] * (R(3 * c) + R(3 * c + 1) / 254.0 + R(3 * c + 2) / 64516.0);
    for (int r = 0; r < K; ++r) {
      const int q = c * K + r;
      f += h[r] * (sc[g + q] * (R(m0 + 2 * q) + R(m0 + 2 * q + 1) / 254.0));
    }
    f -= mj * colsum[c0 + c];
    F[j * ldF + c0 + c] = f * os;
  }
}

// Tile plans of the mode-3 pass: (NT0, NT1) with NT0 + NT1 <= 6.
#define SLFMM_SNP_FILL(X) X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 5) X(2, 1) X(2, 2) X(2, 3) X(2, 4)

// Largest column group of a mode-3 pass at fill rank K.
static int fill_group(int m, int K) {
  for (int g = m; g >= 1; --g) {
    const int nt0 = (3 * g + 31) / 32, nt1 = (2 * g * K + 31) / 32;
    if (nt0 <= 2 && nt0 + nt1 <= 6) return g;
  }
  return 0;
}

struct FnColA {
  const double* A; uint64_t lda;
  __device__ void operator()(uint64_t r, int c, double (&v)[1]) const { v[0] = A[r * lda + c]; }
};

bool fill_functionals_i8_supported(int m, int K) { return m >= 1 && K >= 1 && K <= kNmfMaxK8 && fill_group(m, K) >= 1; }

void fill_functionals_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t p, const double* A,
                         uint64_t lda, int m, const double* W, const double* Ht, int K, const double* mu,
                         const double* outscale, double* F, uint64_t ldF) {
  SLFMM_CHECK(fill_functionals_i8_supported(m, K), SLFMM_E_ARG, "no int8 plan for these filled functionals");
  SLFMM_CHECK(n <= 8000000, SLFMM_E_ARG, "too many samples for the int8 kernels");
  if (p == 0) return;
  hipStream_t s = ctx->stream;
  const int gmax = fill_group(m, K);
  double* colsum = ctx->buf<double>("fi8_colsum", m);
  colreduce<1, kColSum>(ctx, FnColA{A, lda}, GStore1{colsum}, n, m);
  const uint64_t nblk = cdiv(n, 256);
  for (int c0 = 0; c0 < m; c0 += gmax) {
    const int g = std::min(gmax, m - c0);
    const NmfMap map{g, g * K, (3 * g + 31) / 32, (2 * g * K + 31) / 32};
    const int mq = g + g * K, NT = map.NT0 + map.NT1;
    double* Qn = ctx->buf<double>("fi8_q", n * mq);
    k_fill_prod<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(A, lda, n, c0, g, W, K, Qn);
    double* sc = ctx->buf<double>("fi8_sc", mq);
    for (int q0 = 0; q0 < mq; q0 += 64)   // column reductions take up to 64 columns
      colabsmax_scale(ctx, Qn + q0, mq, n, std::min(64, mq - q0), nullptr, nullptr, sc + q0);
    const uint64_t total = nblk * 8 * NT * 1024;
    int8_t* pack = ctx->buf<int8_t>("fi8_pack", total);
    k_pack_rows<<<(unsigned)std::min<uint64_t>(cdiv(total, 256), 16384), 256, 0, s>>>(Qn, mq, n, map, sc, NT, total,
                                                                                    pack);
    const uint64_t ldraw = rup(p, 64);
    int32_t* raw = ctx->buf<int32_t>("fi8_raw", (uint64_t)32 * NT * ldraw);
    SnpArgs a{};
    a.packed = Y; a.ld = ld; a.p = p; a.n = n; a.nblk = (int)nblk;
    a.rb = (uint32_t)std::min<uint64_t>(rup(cdiv(n, 4), 16), ld);
    a.m = mq; a.S = 1; a.fill = 0; a.A = pack;
    a.raw = raw; a.ldraw = ldraw;
    const uint64_t rows = 32 * kSnpWaves;   // one tile per wave
    SLFMM_CHECK((rows + 2 * 32) * a.ld < 0xF0000000ull && total < 0xFFFFFFFFull, SLFMM_E_ARG,
                "panel rows too long for the int8 kernels");
    const unsigned grid = (unsigned)cdiv(p, rows);
    bool done = false;
#define X(A0_, A1_)                                                                               \
    if (!done && map.NT0 == A0_ && map.NT1 == A1_) {                                              \
      k_snp_i8<1, A0_, A1_, 1, 3><<<grid, 64 * kSnpWaves, 0, s>>>(a);                             \
      done = true;                                                                                \
    }
    SLFMM_SNP_FILL(X)
#undef X
    SLFMM_CHECK(done, SLFMM_E_ARG, "no int8 kernel for this filled-functional layout");
    k_fill_combine<<<(unsigned)cdiv(p, 256), 256, 0, s>>>(raw, ldraw, p, map, c0, sc, Ht, mu, colsum, outscale, F,
                                                         ldF);
    SLFMM_HIP(hipGetLastError());
  }
}

// ---- moments of NMF-filled rows on the int8 matrix cores ---------------------------
// S1_j = sum_{k missing} W_k . H_j = H_j . a_j,  S2_j = H_j^T G_j H_j and
// SWp_j = sum_{k missing} Wp_k, with a_j = sum_{k missing} W_k and G_j =
// sum_{k missing} W_k W_k^T: one mode-3 pass whose missing-call side meets
// [W_r | W_r W_s (r <= s) | Wp_r] in 2 int8 terms (a zero column on the
// dosage side), then an FP64 combine with the SNP's own H_j.
__global__ void k_fstat_cols(const double* __restrict__ W, const double* __restrict__ Wp, uint64_t n, int K,
                             double* __restrict__ Qn, int mq) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* w = W + k * K;
  double* q = Qn + k * mq;
  q[0] = 0.0;
  int c = 1;
  for (int r = 0; r < K; ++r) q[c++] = w[r];
  for (int r = 0; r < K; ++r)
    for (int s2 = r; s2 < K; ++s2) q[c++] = w[r] * w[s2];
  if (Wp)
    for (int r = 0; r < K; ++r) q[c++] = Wp[k * K + r];
}

__global__ void k_fstat_combine(const int32_t* __restrict__ raw, uint64_t ldraw, uint64_t p, NmfMap map, int K,
                                const double* __restrict__ sc, const double* __restrict__ Ht,
                                double* __restrict__ S1, double* __restrict__ S2, double* __restrict__ SWp) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p) return;
  const int m0 = 32 * map.NT0;
  // miss-side column q (operand column 1 + q): scale sc[1 + q], rows m0 + 2 q, + 1
  auto M = [&](int q) {
    return sc[1 + q] * ((double)raw[(uint64_t)(m0 + 2 * q) * ldraw + j] +
                        (double)raw[(uint64_t)(m0 + 2 * q + 1) * ldraw + j] / 254.0);
  };
  double h[kNmfMaxK8];
  for (int r = 0; r < K; ++r) h[r] = Ht[j * K + r];
  double s1 = 0.0, s2 = 0.0;
  for (int r = 0; r < K; ++r) s1 += h[r] * M(r);
  int q = K;
  for (int r = 0; r < K; ++r)
    for (int s3 = r; s3 < K; ++s3, ++q) s2 += (r == s3 ? 1.0 : 2.0) * h[r] * h[s3] * M(q);
  S1[j] = s1;
  S2[j] = s2;
  if (SWp)
    for (int r = 0; r < K; ++r) SWp[j * K + r] = M(q + r);
}

void fill_stats_i8(slfmm_ctx* ctx, const uint8_t* Y, uint64_t ld, uint64_t n, uint64_t p, const double* W,
                   const double* Ht, int K, const double* Wp, double* S1, double* S2, double* SWp) {
  SLFMM_CHECK(K >= 1 && K <= kNmfMaxK8, SLFMM_E_ARG, "the int8 fill statistics support rank 1..8");
  SLFMM_CHECK(n <= 8000000, SLFMM_E_ARG, "too many samples for the int8 kernels");
  if (p == 0) return;
  hipStream_t s = ctx->stream;
  const int nm = K + K * (K + 1) / 2 + (Wp ? K : 0), mq = 1 + nm;
  const NmfMap map{1, nm, 1, (2 * nm + 31) / 32};
  const int NT = map.NT0 + map.NT1;
  double* Qn = ctx->buf<double>("fs8_q", n * mq);
  k_fstat_cols<<<(unsigned)cdiv(n, 256), 256, 0, s>>>(W, Wp, n, K, Qn, mq);
  double* sc = ctx->buf<double>("fs8_sc", mq);
  for (int q0 = 0; q0 < mq; q0 += 64)   // column reductions take up to 64 columns
    colabsmax_scale(ctx, Qn + q0, mq, n, std::min(64, mq - q0), nullptr, nullptr, sc + q0);
  const uint64_t nblk = cdiv(n, 256);
  const uint64_t total = nblk * 8 * NT * 1024;
  int8_t* pack = ctx->buf<int8_t>("fs8_pack", total);
  k_pack_rows<<<(unsigned)std::min<uint64_t>(cdiv(total, 256), 16384), 256, 0, s>>>(Qn, mq, n, map, sc, NT, total,
                                                                                  pack);
  const uint64_t ldraw = rup(p, 64);
  int32_t* raw = ctx->buf<int32_t>("fs8_raw", (uint64_t)32 * NT * ldraw);
  SnpArgs a{};
  a.packed = Y; a.ld = ld; a.p = p; a.n = n; a.nblk = (int)nblk;
  a.rb = (uint32_t)std::min<uint64_t>(rup(cdiv(n, 4), 16), ld);
  a.m = mq; a.S = 1; a.fill = 0; a.A = pack;
  a.raw = raw; a.ldraw = ldraw;
  const uint64_t rows = 32 * kSnpWaves;
  SLFMM_CHECK((rows + 2 * 32) * a.ld < 0xF0000000ull && total < 0xFFFFFFFFull, SLFMM_E_ARG,
              "panel rows too long for the int8 kernels");
  const unsigned grid = (unsigned)cdiv(p, rows);
  bool done = false;
#define X(A0_, A1_)                                                                               \
  if (!done && map.NT0 == A0_ && map.NT1 == A1_) {                                                \
    k_snp_i8<1, A0_, A1_, 1, 3><<<grid, 64 * kSnpWaves, 0, s>>>(a);                               \
    done = true;                                                                                  \
  }
  SLFMM_SNP_FILL(X)
#undef X
  SLFMM_CHECK(done, SLFMM_E_ARG, "no int8 kernel for the fill statistics layout");
  k_fstat_combine<<<(unsigned)cdiv(p, 256), 256, 0, s>>>(raw, ldraw, p, map, K, sc, Ht, S1, S2, SWp);
  SLFMM_HIP(hipGetLastError());
}

// ---- NMF forward pass on the int8 matrix cores ---------------------------------
// With the missing calls of sample k filled by W_k . H_j (src/nmf.rs:131-147),
//   num_W[k][r] = sum_j y'_kj H_jr + sum_s W_ks G_k[r][s],
//   G_k[r][s] = sum_{j missing in k} H_jr H_js,
// where y' is the dosage with missing calls as 0.  Both sums run on the
// sample-major copy: the dosage fragment meets H in 3 int8 terms per column,
// the missing-call fragment meets the pairwise products H_r H_s (r <= s) in
// 2.  Slots of the 16-column tiles: g tiles hold term t of column r at
// 3 r + t, miss tiles term t of pair p at 16 NG + 2 p + t.
struct NmfFwdMap {
  int K, P, NG, NM;
  static NmfFwdMap make(int K) {
    const int P = K * (K + 1) / 2;
    return NmfFwdMap{K, P, (3 * K + 15) / 16, (2 * P + 15) / 16};
  }
  // (column of [H | H_r H_s], term) of a slot; column -1: zero slot
  __host__ __device__ int2 info(int slot) const {
    if (slot < 16 * NG) return slot < 3 * K ? make_int2(slot / 3, slot % 3) : make_int2(-1, 0);
    const int r = slot - 16 * NG;
    return r < 2 * P ? make_int2(K + r / 2, r % 2) : make_int2(-1, 0);
  }
};

// Pair p (NmfMap order: r = 0.., s = r..K-1) -> (r, s)
__host__ __device__ __forceinline__ int2 nmf_pair(int p, int K) {
  int r = 0;
  while (p >= K - r) { p -= K - r; ++r; }
  return make_int2(r, r + p);
}

// Column maxima of |[H | H_r H_s]| per workgroup: each thread keeps the K + P
// running maxima of its rows (one row = one SNP's K loadings), then the
// workgroup combines them in LDS; k_colfin finishes across workgroups.
template <int K>
__global__ void __launch_bounds__(256) k_nmf_hmax(const double* __restrict__ Ht, uint64_t pe, uint64_t rpb,
                                                  double* __restrict__ part) {
  constexpr int M = K + K * (K + 1) / 2;
  __shared__ double red[M][257];
  const int t = threadIdx.x;
  double mx[M];
#pragma unroll
  for (int c = 0; c < M; ++c) mx[c] = 0.0;
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb, r1 = min(pe, r0 + rpb);
  for (uint64_t r = r0 + t; r < r1; r += 256) {
    double h[K];
#pragma unroll
    for (int a = 0; a < K; ++a) h[a] = Ht[r * K + a];
#pragma unroll
    for (int a = 0; a < K; ++a) mx[a] = fmax(mx[a], fabs(h[a]));
    int q = K;
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int b = a; b < K; ++b, ++q) mx[q] = fmax(mx[q], fabs(h[a] * h[b]));
  }
#pragma unroll
  for (int c = 0; c < M; ++c) red[c][t] = mx[c];
  __syncthreads();
  if (t < M) {
    double v = 0.0;
    for (int i = 0; i < 256; ++i) v = fmax(v, red[t][i]);
    part[(uint64_t)blockIdx.x * M + t] = v;
  }
}

struct GScaleMax {
  double* out;
  __device__ void operator()(int c, const double (&v)[1]) const { out[c] = v[0] > 0.0 ? v[0] / 127.0 : 1.0; }
};

// B operand of k_fwd_i8 from Ht (pe x K); one workgroup per block of 64 KS
// SNPs, whose loadings are staged in LDS.  Lane chunk (s, ct, lane) of the
// block: lane g 16 + l16 holds slot 16 ct + l16 of the 16 SNPs
// 16 KS g + 16 s + r3 at byte (r3 & 3) 4 + (r3 >> 2).
template <int KS>
__global__ void __launch_bounds__(256) k_pack_nmf_cols(const double* __restrict__ Ht, uint64_t pe, NmfFwdMap map,
                                                       const double* __restrict__ sc, v4i* __restrict__ out) {
  __shared__ double hs[64 * KS * kNmfMaxK8];
  const int K = map.K, NTB = map.NG + map.NM;
  const uint64_t snp0 = (uint64_t)blockIdx.x * 64 * KS;
  for (int i = threadIdx.x; i < 64 * KS * K; i += 256) hs[i] = snp0 + i / K < pe ? Ht[snp0 * K + i] : 0.0;
  __syncthreads();
  const int nch = KS * NTB * 64;
  for (int c = threadIdx.x; c < nch; c += 256) {
    const int lane = c & 63, frag = c >> 6, ct = frag % NTB, s = frag / NTB, g = lane >> 4;
    const int2 ci = map.info(16 * ct + (lane & 15));
    uint32_t word[4] = {0u, 0u, 0u, 0u};
    if (ci.x >= 0) {
      int a = ci.x, b = -1;
      if (a >= K) {
        const int2 rs = nmf_pair(a - K, K);
        a = rs.x; b = rs.y;
      }
      const double inv = 1.0 / sc[ci.x];
      const double* h = hs + (16 * KS * g + 16 * s) * K;
#pragma unroll 4
      for (int r3 = 0; r3 < 16; ++r3) {
        const double v = b >= 0 ? h[r3 * K + a] * h[r3 * K + b] : h[r3 * K + a];
        const int q = quant_term(v * inv, ci.y);
        word[r3 & 3] |= (uint32_t)(q & 0xFF) << (8 * (r3 >> 2));
      }
    }
    out[(uint64_t)blockIdx.x * nch + c] = v4i{(int)word[0], (int)word[1], (int)word[2], (int)word[3]};
  }
}

// Epilogue: the int32 sums of every slot go out as they are (the chunk sums
// are exact in 64 bits; the terms are recombined after them).
template <int NTB_, int NG_>
struct FwdRawEpi {
  static constexpr int NTB = NTB_, NG = NG_;
  static constexpr int WV = 4;   // waves per workgroup
  int32_t* part;
  template <int ST>
  __device__ void operator()(const v4i (&acc)[ST][NTB], uint64_t sbase, uint64_t n, int g, int l16,
                             uint64_t chunk) const {
#pragma unroll
    for (int st = 0; st < ST; ++st)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint64_t k = sbase + 16 * st + 4 * g + r;
        if (k >= n) continue;
        int32_t* row = part + (chunk * n + k) * (16 * NTB);
#pragma unroll
        for (int ct = 0; ct < NTB; ++ct) row[16 * ct + l16] = acc[st][ct][r];
      }
  }
};

// Chunk sums of every slot (chunk order, exact in 64 bits): one thread per
// (sample, slot), so a wave reads whole partial rows.
__global__ void k_sum_slots(const int32_t* __restrict__ part, uint64_t count, int nchunk, double* __restrict__ out) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  int64_t t = 0;
  for (int c = 0; c < nchunk; ++c) t += part[(uint64_t)c * count + e];
  out[e] = (double)t;
}

// num_W[k][r] = A_r + sum_s W_ks G[r][s] from the slot sums of sample k and
// the W of the current iterate
__global__ void k_nmf_fwd_combine(const double* __restrict__ ssum, uint64_t n, NmfFwdMap map,
                                  const double* __restrict__ sc, const double* __restrict__ W,
                                  double* __restrict__ num) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int K = map.K;
  if (e >= n * (uint64_t)K) return;
  const uint64_t k = e / K;
  const int r = (int)(e % K);
  const double* S = ssum + k * 16 * (uint64_t)(map.NG + map.NM);
  double v = sc[r] * (S[3 * r] + S[3 * r + 1] / 254.0 + S[3 * r + 2] / 64516.0);
  int p = 0;
  for (int a = 0; a < K; ++a)
    for (int b = a; b < K; ++b, ++p) {
      if (a != r && b != r) continue;
      const int slot = 16 * map.NG + 2 * p;
      v += W[k * K + (a == r ? b : a)] * (sc[K + p] * (S[slot] + S[slot + 1] / 254.0));
    }
  num[e] = v;
}

// num_W (n x K) of the reference's forward pass from the sample-major copy
// YT (n x ldt bytes) of the estimation rows, Ht (pe x K) and the current W.
void nmf_forward_i8(slfmm_ctx* ctx, const uint8_t* YT, uint64_t ldt, uint64_t n, uint64_t pe, const double* Ht,
                    const double* W, int K, double* num) {
  SLFMM_CHECK(K >= 1 && K <= kNmfMaxK8, SLFMM_E_ARG, "the int8 NMF pass supports rank 1..8");
  hipStream_t s = ctx->stream;
  const NmfFwdMap map = NmfFwdMap::make(K);
  const int KS = 4, NTB = map.NG + map.NM, mq = K + map.P;
  double* sc = ctx->buf<double>("nmf8f_sc", mq);
  {
    uint64_t nb = std::min<uint64_t>(std::max<uint64_t>(cdiv(pe, 2048), 1), 1024);
    const uint64_t rpb = cdiv(pe, nb);
    nb = cdiv(pe, rpb);
    double* hpart = ctx->buf<double>("nmf8f_hpart", nb * mq);
    switch (K) {
#define SLFMM_HMAX(KK) case KK: k_nmf_hmax<KK><<<(unsigned)nb, 256, 0, s>>>(Ht, pe, rpb, hpart); break;
      SLFMM_HMAX(1) SLFMM_HMAX(2) SLFMM_HMAX(3) SLFMM_HMAX(4) SLFMM_HMAX(5) SLFMM_HMAX(6) SLFMM_HMAX(7) SLFMM_HMAX(8)
#undef SLFMM_HMAX
    }
    k_colfin<1, kColMax, GScaleMax><<<(unsigned)cdiv(mq, 4), 256, 0, s>>>(hpart, (int)nb, mq, GScaleMax{sc});
  }
  const uint64_t nblk = cdiv(pe, 64 * KS);
  const uint64_t n16 = nblk * KS * NTB * 64;   // 16-byte lane chunks
  SLFMM_CHECK(n16 * 16 < 0xFFFFFFFFull, SLFMM_E_ARG, "estimation panel too large for the int8 NMF forward pass");
  v4i* Bp = ctx->buf<v4i>("nmf8f_pack", n16);
  k_pack_nmf_cols<4><<<(unsigned)nblk, 256, 0, s>>>(Ht, pe, map, sc, Bp);
  constexpr int WV = FwdRawEpi<1, 1>::WV;
  const uint64_t sblocks = cdiv(n, 64 * WV);
  // a chunk's int32 sums stay exact below 2^31 / 254 SNPs (|q| <= 127, dosage <= 2)
  uint64_t per = 0;
  switch (K) {
#define SLFMM_NMF8F(KK)                                                                                       \
  case KK:                                                                                                    \
    per = fwd_blocks_per_chunk(                                                                               \
        ctx, k_fwd_i8<FwdRawEpi<(3 * KK + 15) / 16 + (KK * (KK + 1) + 15) / 16, (3 * KK + 15) / 16>>, 64 * WV, \
        nblk, sblocks, cdiv(pe, 8u << 20));                                                                   \
    break;
    SLFMM_NMF8F(1) SLFMM_NMF8F(2) SLFMM_NMF8F(3) SLFMM_NMF8F(4) SLFMM_NMF8F(5) SLFMM_NMF8F(6) SLFMM_NMF8F(7)
    SLFMM_NMF8F(8)
#undef SLFMM_NMF8F
  }
  const uint64_t nchunk = cdiv(nblk, per);
  int32_t* part = ctx->buf<int32_t>("nmf8f_part", nchunk * n * 16 * NTB);
  const dim3 grid((unsigned)sblocks, (unsigned)nchunk);
  switch (K) {
#define SLFMM_NMF8F(KK)                                                                                       \
  case KK:                                                                                                    \
    k_fwd_i8<<<grid, 64 * WV, 0, s>>>(                                                                        \
        YT, ldt, n, (int)per, (int)nblk, Bp,                                                                  \
        FwdRawEpi<(3 * KK + 15) / 16 + (KK * (KK + 1) + 15) / 16, (3 * KK + 15) / 16>{part});                 \
    break;
    SLFMM_NMF8F(1) SLFMM_NMF8F(2) SLFMM_NMF8F(3) SLFMM_NMF8F(4) SLFMM_NMF8F(5) SLFMM_NMF8F(6) SLFMM_NMF8F(7)
    SLFMM_NMF8F(8)
#undef SLFMM_NMF8F
  }
  const uint64_t nslot = n * 16 * NTB;
  double* ssum = ctx->buf<double>("nmf8f_ssum", nslot);
  k_sum_slots<<<(unsigned)cdiv(nslot, 256), 256, 0, s>>>(part, nslot, (int)nchunk, ssum);
  k_nmf_fwd_combine<<<(unsigned)cdiv(n * K, 256), 256, 0, s>>>(ssum, n, map, sc, W, num);
  SLFMM_HIP(hipGetLastError());
}

// ---- NMF-filled RSVD forward pass on the int8 matrix cores ------------------------
// With the missing calls of sample k filled by W_k . H_j (NmfInRam,
// src/parallel.rs:284-321), the centred forward product is
//   W_kc = sum_j g'_kj Qt_jc + sum_r W_kr sum_{j missing in k} H_jr Qt_jc
//          - sum_j mu_j Qt_jc,     Qt = diag(sc) Q_z,
// with g' the dosage (missing calls 0).  Per group of g columns, one pass of
// k_fwd_i8 over the sample-major copy: the dosage fragment meets Qt_c in 3
// int8 terms, the missing-call fragment the products H_r Qt_c in 2
// (NmfFwdMap with K -> g, P -> g K); the exact int32 slot sums are combined
// in FP64 with the sample's own W_k.  The last term (vhat) is an FP64 column
// sum.
struct FnFillB {
  const double* Qz; uint64_t ldq; const double* sc; const double* Ht; int K, c0, g;
  __device__ double val(uint64_t j, int x) const {
    if (x < g) return sc[j] * Qz[j * ldq + c0 + x];
    const int q = x - g;
    return Ht[j * K + q % K] * (sc[j] * Qz[j * ldq + c0 + q / K]);
  }
  __device__ void operator()(uint64_t j, int x, double (&v)[1]) const { v[0] = fabs(val(j, x)); }
};
struct FnVhatF {
  const double* Qz; uint64_t ldq; const double* sc; const double* mu;
  __device__ void operator()(uint64_t j, int c, double (&v)[1]) const { v[0] = mu[j] * (sc[j] * Qz[j * ldq + c]); }
};

// Largest column group of the filled forward pass at rank K (at most 7
// tiles of 16 per pass, at most 16 columns).
constexpr int fill_fwd_gmax(int K) { return 112 / (3 + 2 * K) < 16 ? 112 / (3 + 2 * K) : 16; }

// Column maxima of |B| of one group per workgroup (rows of SNPs, Qt and H
// values in registers; K and the group bound are template parameters).
template <int K>
__global__ void __launch_bounds__(256) k_fill_bmax(FnFillB fb, uint64_t pe, uint64_t rpb, double* __restrict__ part) {
  constexpr int G = fill_fwd_gmax(K), M = G + G * K;
  __shared__ double red[M][257];
  const int t = threadIdx.x, g = fb.g, mq = g + g * K;
  double mx[M];
#pragma unroll
  for (int x = 0; x < M; ++x) mx[x] = 0.0;
  const uint64_t r0 = (uint64_t)blockIdx.x * rpb, r1 = min(pe, r0 + rpb);
  for (uint64_t j = r0 + t; j < r1; j += 256) {
    double qt[G], h[K];
#pragma unroll
    for (int c = 0; c < G; ++c) qt[c] = c < g ? fb.sc[j] * fb.Qz[j * fb.ldq + fb.c0 + c] : 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) h[r] = fb.Ht[j * K + r];
#pragma unroll
    for (int c = 0; c < G; ++c) {
      mx[c] = fmax(mx[c], fabs(qt[c]));
#pragma unroll
      for (int r = 0; r < K; ++r) mx[G + c * K + r] = fmax(mx[G + c * K + r], fabs(h[r] * qt[c]));
    }
  }
#pragma unroll
  for (int x = 0; x < M; ++x) red[x][t] = mx[x];
  __syncthreads();
  for (int x = t; x < mq; x += 256) {   // operand column x of the group
    const int i = x < g ? x : G + (x - g);   // (x - g) = c K + r
    double v = 0.0;
    for (int q = 0; q < 256; ++q) v = fmax(v, red[i][q]);
    part[(uint64_t)blockIdx.x * mq + x] = v;
  }
}

// B operand of one group (layout of k_pack_nmf_cols): one workgroup per
// block of 64 KS SNPs, whose Qt and H values are staged in LDS.
template <int KS>
__global__ void __launch_bounds__(256) k_pack_fill_cols(FnFillB fb, uint64_t pe, NmfFwdMap map,
                                                        const double* __restrict__ bsc, v4i* __restrict__ out) {
  __shared__ double qs[64 * KS][16];
  __shared__ double hs[64 * KS][kNmfMaxK8];
  const int NTB = map.NG + map.NM, g = fb.g, K = fb.K;
  const uint64_t snp0 = (uint64_t)blockIdx.x * 64 * KS;
  for (int e = threadIdx.x; e < 64 * KS * 16; e += 256) {
    const int jj = e >> 4, c = e & 15;
    const uint64_t j = snp0 + jj;
    qs[jj][c] = j < pe && c < g ? fb.sc[j] * fb.Qz[j * fb.ldq + fb.c0 + c] : 0.0;
  }
  for (int e = threadIdx.x; e < 64 * KS * K; e += 256) {
    const int jj = e / K, r = e % K;
    const uint64_t j = snp0 + jj;
    hs[jj][r] = j < pe ? fb.Ht[j * K + r] : 0.0;
  }
  __syncthreads();
  const int nch = KS * NTB * 64;
  for (int c = threadIdx.x; c < nch; c += 256) {
    const int lane = c & 63, frag = c >> 6, ct = frag % NTB, s = frag / NTB, gg = lane >> 4;
    const int2 ci = map.info(16 * ct + (lane & 15));
    uint32_t word[4] = {0u, 0u, 0u, 0u};
    if (ci.x >= 0) {
      const double inv = 1.0 / bsc[ci.x];
      const int col = ci.x < g ? ci.x : (ci.x - g) / K, r = ci.x < g ? -1 : (ci.x - g) % K;
      const int jb = 16 * KS * gg + 16 * s;
#pragma unroll 4
      for (int r3 = 0; r3 < 16; ++r3) {
        const double qt = qs[jb + r3][col];
        const double v = r < 0 ? qt : hs[jb + r3][r] * qt;
        const int q = quant_term(v * inv, ci.y);
        word[r3 & 3] |= (uint32_t)(q & 0xFF) << (8 * (r3 >> 2));
      }
    }
    out[(uint64_t)blockIdx.x * nch + c] = v4i{(int)word[0], (int)word[1], (int)word[2], (int)word[3]};
  }
}

// W[k][c0 + c] from the slot sums of sample k
__global__ void k_fill_fwd_combine(const double* __restrict__ ssum, uint64_t n, NmfFwdMap map, int c0,
                                   const double* __restrict__ bsc, const double* __restrict__ Wf,
                                   const double* __restrict__ vhat, double* __restrict__ W, uint64_t ldw) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = map.K, K = map.P / map.K;
  if (e >= n * (uint64_t)g) return;
  const uint64_t k = e / g;
  const int c = (int)(e % g);
  const double* S = ssum + k * 16 * (uint64_t)(map.NG + map.NM);
  double v = bsc[c] * (S[3 * c] + S[3 * c + 1] / 254.0 + S[3 * c + 2] / 64516.0);
  for (int r = 0; r < K; ++r) {
    const int q = c * K + r, slot = 16 * map.NG + 2 * q;
    v += Wf[k * K + r] * (bsc[g + q] * (S[slot] + S[slot + 1] / 254.0));
  }
  W[k * ldw + c0 + c] = v - vhat[c0 + c];
}

// Tile plans (NG, NM) of the filled forward pass (NG + NM <= 7, fill rank <= 8).
#define SLFMM_FWD_FILL(X) \
  X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 5) X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(3, 2) X(3, 3) X(3, 4)

static int fill_fwd_group(int l, int K) {
  for (int g = std::min(l, fill_fwd_gmax(K)); g >= 1; --g)
    if ((3 * g + 15) / 16 + (2 * g * K + 15) / 16 <= 7) return g;
  return 0;
}

bool fill_forward_i8_supported(int l, int K) { return l >= 1 && K >= 1 && K <= kNmfMaxK8 && fill_fwd_group(l, K) >= 1; }

void fill_forward_i8(slfmm_ctx* ctx, const uint8_t* YT, uint64_t ldt, uint64_t n, uint64_t pe, const double* Qz,
                     uint64_t ldq, int l, const double* mu, const double* sc, const double* Wf, const double* Ht, int K,
                     double* W, uint64_t ldw) {
  SLFMM_CHECK(fill_forward_i8_supported(l, K), SLFMM_E_ARG, "no int8 plan for this filled forward pass");
  hipStream_t s = ctx->stream;
  constexpr int KS = 4, WV = FwdRawEpi<1, 1>::WV;
  double* vhat = ctx->buf<double>("ff8_vhat", l);
  colreduce<1, kColSum>(ctx, FnVhatF{Qz, ldq, sc, mu}, GStore1{vhat}, pe, l);
  const uint64_t nblk = cdiv(pe, 64 * KS), sblocks = cdiv(n, 64 * WV);
  const int gmax = fill_fwd_group(l, K);
  for (int c0 = 0; c0 < l; c0 += gmax) {
    const int g = std::min(gmax, l - c0);
    const NmfFwdMap map{g, g * K, (3 * g + 15) / 16, (2 * g * K + 15) / 16};
    const int NTB = map.NG + map.NM, mq = g + g * K;
    const FnFillB fb{Qz, ldq, sc, Ht, K, c0, g};
    double* bsc = ctx->buf<double>("ff8_sc", mq);
    {
      uint64_t nb = std::min<uint64_t>(std::max<uint64_t>(cdiv(pe, 2048), 1), 1024);
      const uint64_t rpb = cdiv(pe, nb);
      nb = cdiv(pe, rpb);
      double* bpart = ctx->buf<double>("ff8_bpart", nb * mq);
      switch (K) {
#define SLFMM_BMAX(KK) case KK: k_fill_bmax<KK><<<(unsigned)nb, 256, 0, s>>>(fb, pe, rpb, bpart); break;
        SLFMM_BMAX(1) SLFMM_BMAX(2) SLFMM_BMAX(3) SLFMM_BMAX(4) SLFMM_BMAX(5) SLFMM_BMAX(6) SLFMM_BMAX(7)
        SLFMM_BMAX(8)
#undef SLFMM_BMAX
      }
      k_colfin<1, kColMax, GScaleMax><<<(unsigned)cdiv(mq, 4), 256, 0, s>>>(bpart, (int)nb, mq, GScaleMax{bsc});
    }
    const uint64_t n16 = nblk * KS * NTB * 64;
    SLFMM_CHECK(n16 * 16 < 0xFFFFFFFFull, SLFMM_E_ARG, "estimation panel too large for the int8 filled forward pass");
    v4i* Bp = ctx->buf<v4i>("ff8_pack", n16);
    k_pack_fill_cols<KS><<<(unsigned)nblk, 256, 0, s>>>(fb, pe, map, bsc, Bp);
    uint64_t per = 0;
    bool done = false;
#define X(NG_, NM_)                                                                                         \
    if (!done && map.NG == NG_ && map.NM == NM_) {                                                          \
      per = fwd_blocks_per_chunk(ctx, k_fwd_i8<FwdRawEpi<NG_ + NM_, NG_>>, 64 * WV, nblk, sblocks,         \
                                 cdiv(pe, 8u << 20));                                                       \
      const uint64_t nchunk = cdiv(nblk, per);                                                              \
      int32_t* part = ctx->buf<int32_t>("ff8_part", nchunk * n * 16 * NTB);                                 \
      k_fwd_i8<<<dim3((unsigned)sblocks, (unsigned)nchunk), 64 * WV, 0, s>>>(                               \
          YT, ldt, n, (int)per, (int)nblk, Bp, FwdRawEpi<NG_ + NM_, NG_>{part});                            \
      const uint64_t nslot = n * 16 * NTB;                                                                  \
      double* ssum = ctx->buf<double>("ff8_ssum", nslot);                                                   \
      k_sum_slots<<<(unsigned)cdiv(nslot, 256), 256, 0, s>>>(part, nslot, (int)nchunk, ssum);              \
      k_fill_fwd_combine<<<(unsigned)cdiv(n * g, 256), 256, 0, s>>>(ssum, n, map, c0, bsc, Wf, vhat, W, ldw); \
      done = true;                                                                                          \
    }
    SLFMM_FWD_FILL(X)
#undef X
    SLFMM_CHECK(done, SLFMM_E_ARG, "no int8 kernel for this filled forward layout");
    SLFMM_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// Layout probe (test-only): one v_mfma_i32_16x16x64_i8 with the fragment
// mapping the kernels above assume.  A is 16 x 64, B is 64 x 16, row-major.
// ---------------------------------------------------------------------------
__global__ void k_probe_mfma_i8(const int8_t* __restrict__ Afrag, const int8_t* __restrict__ B, int* D) {
  __shared__ __align__(16) unsigned char s_a[1024];
  const int lane = threadIdx.x, g = lane >> 4, l16 = lane & 15;
  chunk_to_lds(reinterpret_cast<const unsigned char*>(Afrag), s_a, lane);   // the kernels' LDS-DMA path
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  v4i a = reinterpret_cast<const v4i*>(s_a)[lane];
  v4i b;
  for (int v = 0; v < 4; ++v) {
    uint32_t wb = 0;
    for (int c = 0; c < 4; ++c) wb |= (uint32_t)(uint8_t)B[(16 * g + 4 * c + v) * 16 + l16] << (8 * c);
    b[v] = (int)wb;
  }
  v4i d = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, v4i{0, 0, 0, 0}, 0, 0, 0);
  for (int r = 0; r < 4; ++r) D[(4 * g + r) * 16 + l16] = d[r];
}

void probe_mfma_i8(slfmm_ctx* ctx, const int8_t* hA, const int8_t* hB, int32_t* hD) {
  hipStream_t s = ctx->stream;
  // A (16 x 64, row-major) -> one fragment chunk: lane l, byte 4v+c = A[l%16][16(l/16) + 4c + v]
  std::vector<int8_t> frag(1024);
  for (int l = 0; l < 64; ++l)
    for (int v = 0; v < 4; ++v)
      for (int c = 0; c < 4; ++c) frag[l * 16 + 4 * v + c] = hA[(l & 15) * 64 + 16 * (l >> 4) + 4 * c + v];
  int8_t* dA = ctx->buf<int8_t>("probe_A", 1024);
  int8_t* dB = ctx->buf<int8_t>("probe_B", 1024);
  int* dD = ctx->buf<int>("probe_D", 256);
  SLFMM_HIP(hipMemcpyAsync(dA, frag.data(), 1024, hipMemcpyHostToDevice, s));
  SLFMM_HIP(hipMemcpyAsync(dB, hB, 1024, hipMemcpyHostToDevice, s));
  k_probe_mfma_i8<<<1, 64, 0, s>>>(dA, dB, dD);
  SLFMM_HIP(hipGetLastError());
  SLFMM_HIP(hipMemcpyAsync(hD, dD, 1024, hipMemcpyDeviceToHost, s));
  SLFMM_HIP(hipStreamSynchronize(s));   // frag is a host buffer read by the async copy above
}

}  // namespace slfmm
