// Mixed-precision genotype-stream kernels on gfx950 int8 MFMA.
//
// Genotype codes are exact small integers, so they enter the matrix cores
// as int8 with no rounding: per 16 samples of one SNP the lane decodes one
// 32-bit word of packed codes into two int8 operand fragments,
//   g0   = dosage with missing -> 0   (00->0, 01->0, 10->1, 11->2)
//   miss = missing indicator          (01->1, else 0)
// and v_mfma_i32_16x16x64_i8 accumulates exactly in int32.  The dense
// operand (an orthonormal basis Q of [U_c|X] for the association, the
// sketch B for the RSVD) is carried as 2 or 3 int8 terms per entry with a
// per-column scale, Q ~ sc (t0 + t1/254 + t2/254^2) (about 15 / 22
// significant bits).  Per SNP and functional i, with G_t, M_t, O_t the int32
// sums of term t against g0, miss and 1:
//   f_i = sum_t w_t (G_t - mu O_t) + mu sum_t w_t M_t,   w_t = sc_i 254^-t
// is the functional of the centred column (g - mu on observed calls, 0 on
// missing calls; src/bed.rs:339-343), and the FP64 per-SNP epilogue of the
// association pass runs on it unchanged.
//
// Operand fragments: for a block of 64*KS samples, lane group g = lane/16
// of a wave loads KS consecutive words of its SNP row, word s serving k-step
// s; byte c of decoded register v holds sample
//   64*KS*blk + 16*KS*g + 16*s + 4*c + v,
// and the dense operand is pre-permuted in memory into exactly this order
// (one 16-byte fragment per lane), so the same K index meets on both sides.
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

// Row layout of a packed operand with m functionals: columns i < k3 carry
// 2 int8 terms, columns i >= k3 carry 3, on the g0 side (rows 0 .. 16 NT0);
// every column carries 2 terms on the miss side (rows 16 NT0 ..).
// The g0 block ends with one all-ones row, whose accumulator is the dosage
// sum of the SNP over its observed calls.
struct RowMap {
  int m, k3, NT0, NT1;
  __host__ __device__ static RowMap make(int m, int k3) {
    RowMap r{m, k3, 0, 0};
    r.NT0 = (2 * k3 + 3 * (m - k3) + 1 + 15) / 16;
    r.NT1 = (2 * m + 15) / 16;
    return r;
  }
  __host__ __device__ int ones() const { return 2 * k3 + 3 * (m - k3); }
  // (column, term) of packed row `row`: column -1 for padding, -2 for the ones row
  __device__ int2 info(int row) const {
    if (row < 16 * NT0) {
      if (row < 2 * k3) return make_int2(row / 2, row % 2);
      const int r = row - 2 * k3;
      if (r < 3 * (m - k3)) return make_int2(k3 + r / 3, r % 3);
      return row == ones() ? make_int2(-2, 0) : make_int2(-1, 0);
    }
    const int r = row - 16 * NT0;
    return r < 2 * m ? make_int2(r / 2, r % 2) : make_int2(-1, 0);
  }
  __host__ __device__ int ng(int i) const { return i < k3 ? 2 : 3; }
  __host__ __device__ int rg(int i, int t) const { return i < k3 ? 2 * i + t : 2 * k3 + 3 * (i - k3) + t; }
  __host__ __device__ int rm(int i, int t) const { return 16 * NT0 + 2 * i + t; }
};

// Apack for the SNP-side kernels: rows = functional terms, K = samples.
__global__ void k_pack_rows(const double* __restrict__ Q, uint64_t ldq, uint64_t n, RowMap rmap,
                            const double* __restrict__ sc, int NT, int KS, uint64_t total, int8_t* __restrict__ out) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e & 15), lane = (int)((e >> 4) & 63);
    const uint64_t frag = e >> 10;                 // (blk*KS + s)*NT + rt
    const int rt = (int)(frag % NT);
    const uint64_t ks = frag / NT;
    const int s = (int)(ks % KS);
    const uint64_t blk = ks / KS;
    const int v = j >> 2, c = j & 3, g = lane >> 4;
    const uint64_t sample = blk * 64 * KS + 16 * KS * g + 16 * s + 4 * c + v;
    const int2 ri = rmap.info(16 * rt + (lane & 15));
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

// Row recombination table of one packed operand (device copy).
struct Recomb {
  int m;
  int ones;                // row of the all-ones operand row (dosage sum)
  int ng[kMaxM];           // terms on the g0 side (2 or 3)
  int rg[kMaxM][3];        // their rows
  int rm[kMaxM][2];        // rows of the two terms on the miss side
  // With X = sum_t G_t 254^(ng-1-t), Y = M_0 254 + M_1 (exact integers):
  //   f_i = wg_i X + mu (wm_i Y - C_i),   C_i = sum_t sc_i 254^-t O_t
  double wg[kMaxM];        // sc_i / 254^(ng-1)
  double wm[kMaxM];        // sc_i / 254
  double C[kMaxM];
};

__global__ void k_recomb(RowMap rmap, const double* __restrict__ sc, const double* __restrict__ O, Recomb* rc) {
  const int i = threadIdx.x;
  if (i == 0) {
    rc->m = rmap.m;
    rc->ones = rmap.ones();
  }
  if (i >= kMaxM) return;
  if (i >= rmap.m) {   // padding functionals: zero weights on valid rows, so f_i = 0
    rc->ng[i] = 2;
    for (int t = 0; t < 3; ++t) rc->rg[i][t] = 0;
    rc->rm[i][0] = rc->rm[i][1] = 0;
    rc->wg[i] = rc->wm[i] = rc->C[i] = 0.0;
    return;
  }
  const int ng = rmap.ng(i);
  rc->ng[i] = ng;
  for (int t = 0; t < 3; ++t) rc->rg[i][t] = t < ng ? rmap.rg(i, t) : -1;
  rc->rm[i][0] = rmap.rm(i, 0);
  rc->rm[i][1] = rmap.rm(i, 1);
  rc->wg[i] = sc[i] * (ng == 3 ? 1.0 / 64516.0 : 1.0 / 254.0);
  rc->wm[i] = sc[i] / 254.0;
  double c = O[i * 3 + 0] + O[i * 3 + 1] / 254.0;
  if (ng == 3) c += O[i * 3 + 2] / 64516.0;
  rc->C[i] = sc[i] * c;
}

struct Operand {
  int8_t* pack = nullptr;
  int NT0 = 0, NT1 = 0, KS = 4;
  uint64_t nblk = 0;
  Recomb* rc = nullptr;
};

static int ks_for(int NT) { return NT <= 8 ? 4 : 2; }   // k-steps per staged block of the SNP-side kernels

// Quantise Q (n x m, ldq) into the packed fragment layout; columns from k3
// on carry 3 terms on the g0 side.  Device work only (no host round trip).
static Operand prepare_operand(slfmm_ctx* ctx, const char* tag, const double* Q, uint64_t ldq, uint64_t n, int m,
                               int k3) {
  hipStream_t s = ctx->stream;
  SLFMM_CHECK(m >= 1 && m <= kMaxM && k3 >= 0 && k3 <= m, SLFMM_E_ARG, "operand shape");
  Operand op;
  std::string T(tag);
  const RowMap rmap = RowMap::make(m, k3);
  op.NT0 = rmap.NT0;
  op.NT1 = rmap.NT1;
  const int NT = op.NT0 + op.NT1;
  op.KS = ks_for(NT);
  op.nblk = cdiv(n, 64 * op.KS);
  double* sc = ctx->buf<double>(T + "_sc", m);
  colabsmax_scale(ctx, Q, ldq, n, m, nullptr, nullptr, sc);
  double* O = ctx->buf<double>(T + "_O", m * 3);
  colreduce<3, kColSum>(ctx, FnTerms{Q, ldq, sc}, GStore3{O}, n, m);
  op.rc = ctx->buf<Recomb>(T + "_rc", 1);
  k_recomb<<<1, kMaxM, 0, s>>>(rmap, sc, O, op.rc);
  const uint64_t total = op.nblk * op.KS * NT * 1024;
  op.pack = ctx->buf<int8_t>(T + "_pack", total);
  k_pack_rows<<<(unsigned)std::min<uint64_t>(cdiv(total, 256), 16384), 256, 0, s>>>(Q, ldq, n, rmap, sc, NT, op.KS,
                                                                                  total, op.pack);
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
// bytes 0..3 pick bytes of the second operand).  Counts: code 11 (hom-alt)
// and code 01 (missing); the dosage sum comes from the all-ones operand row.
constexpr uint32_t kPermG0 = 0x02010000u;     // 00->0, 01->0, 10->1, 11->2
constexpr uint32_t kPermMiss = 0x00000100u;   // 01->1, else 0

// Per-word counts for the epilogue, three instructions per 16 codes:
// popc(w) sums b0 + b1 over the codes (00->0, 01->1, 10->1, 11->2), which
// exceeds the dosage (0, 0, 1, 2) by one exactly at missing calls; popc of the
// high bits counts 10 and 11.  With the exact dosage sum sg of the all-ones
// operand row, n_miss = pop - sg and n_hom = sg - hi.
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

// ---------------------------------------------------------------------------
// SNP-side kernel: functionals of every SNP row against a packed operand.
// MODE 0: fused association epilogue;  MODE 1: store s_c * f (RSVD sketch).
// Workgroup: 4 waves x 4 SNP tiles of 16 SNPs (256 SNPs).
// ---------------------------------------------------------------------------
struct GtArgs {
  const uint8_t* packed;
  uint64_t ld, p, n;
  int nblk, m, norm;
  const v4i* A;
  const Recomb* rc;
  EpiArgs ea;
  double* Z;
  uint64_t ldZ;
};

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Asynchronous copy of one 1 KB fragment chunk (64 lanes x 16 B) from global
// memory straight into LDS (global_load_lds_dwordx4): lane l's 16 bytes land
// at dst + 16 l.  The destination is uniform across the wave.
__device__ __forceinline__ void chunk_to_lds(const unsigned char* src_chunk, unsigned char* dst_chunk, int lane) {
  __builtin_amdgcn_global_load_lds((const void*)(src_chunk + 16 * lane), (lds_ptr_t)(dst_chunk), 16, 0, 0);
}

// ---------------------------------------------------------------------------
// Epilogue of the SNP-side kernels.  After the block loop, lane (g, l16) of
// a wave holds rows 16 rt + 4 g + r of SNP column l16 of every tile st.  Per
// tile the wave stages its int32 row sums in its own `red` area ([16 NT
// rows][16 SNPs]); then every lane recombines the functionals i = g + 4 q of
// SNP j = base + 16 st + l16 (all 64 lanes busy), and either stores them
// scaled (MODE 1, RSVD sketch) or feeds the split per-SNP epilogue (MODE 0).
// `red` is private to the wave and a wave's LDS accesses complete in order,
// so the tiles need no workgroup barrier.  The recombination table is read
// from its LDS copy.
// ---------------------------------------------------------------------------
// functionals per lane group: MAXM / 4 for the association (MAXM = 24 or 32),
// 8 for the RSVD sketch (l <= 32)
template <int MAXM, int MODE>
constexpr int epi_nq() { return MODE == 1 ? 8 : MAXM / 4; }

// Copies the recombination table to LDS (all threads of the workgroup take
// part; the caller synchronises before reading it).
__device__ __forceinline__ const Recomb* stage_recomb(const Recomb* src, unsigned char* dst, int t, int nthreads) {
  static_assert(sizeof(Recomb) % 4 == 0, "table is copied as words");
  const int* s = reinterpret_cast<const int*>(src);
  int* d = reinterpret_cast<int*>(dst);
  for (int e = t; e < (int)(sizeof(Recomb) / 4); e += nthreads) d[e] = s[e];
  return reinterpret_cast<const Recomb*>(dst);
}

template <int ST, int NT, int MAXM, int MODE>
__device__ __forceinline__ void gt_epilogue(const GtArgs& a, const v4i (&acc)[ST][NT], uint32_t (&c_pop)[ST],
                                            uint32_t (&c_hi)[ST], int* red, const double* plan, const Recomb* rc,
                                            uint64_t base, int g, int l16) {
  constexpr int NQ = epi_nq<MAXM, MODE>();
#pragma unroll
  for (int st = 0; st < ST; ++st) {
    c_pop[st] += __shfl_xor(c_pop[st], 16); c_pop[st] += __shfl_xor(c_pop[st], 32);
    c_hi[st] += __shfl_xor(c_hi[st], 16); c_hi[st] += __shfl_xor(c_hi[st], 32);
  }
  const int ones = rc->ones;
#pragma unroll
  for (int st = 0; st < ST; ++st) {
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[(16 * rt + 4 * g + r) * 16 + l16] = acc[st][rt][r];
    asm volatile("" ::: "memory");   // the wave's sums are written before any lane reads them
    const uint64_t j = base + 16 * st + l16;
    const bool valid = j < a.p;
    const int64_t sg = red[ones * 16 + l16];   // dosage sum (all-ones operand row)
    const int64_t nmiss = (int64_t)c_pop[st] - sg, nhom = sg - (int64_t)c_hi[st];
    const int64_t nobs = (int64_t)a.n - nmiss;
    const double mu = nobs > 0 ? (double)sg / (double)nobs : 0.0;
    double fq[NQ];   // functionals past m have zero weights in the table: fq = 0
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int i = g + 4 * q;
      // X and Y are integers below 2^47, so these FP64 Horner steps are exact
      const int ng = rc->ng[i];
      double X = fma((double)red[rc->rg[i][0] * 16 + l16], 254.0, (double)red[rc->rg[i][1] * 16 + l16]);
      if (ng == 3) X = fma(X, 254.0, (double)red[rc->rg[i][2] * 16 + l16]);
      const double Y = fma((double)red[rc->rm[i][0] * 16 + l16], 254.0, (double)red[rc->rm[i][1] * 16 + l16]);
      fq[q] = rc->wg[i] * X + mu * (rc->wm[i] * Y - rc->C[i]);
    }
    asm volatile("" ::: "memory");   // this tile's reads precede the next tile's writes
    if constexpr (MODE == 1) {
      const double scale = epi_scale(mu, a.norm);
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (valid && g + 4 * q < a.m) a.Z[j * a.ldZ + g + 4 * q] = scale * fq[q];
    } else {
      snp_epilogue_split<NQ>(a.ea, plan, j, valid, g, fq, nmiss, sg, nhom);
    }
  }
}

template <int NT0, int NT1, int MAXM, int MODE>
__global__ void __launch_bounds__(256, 2) k_gt_i8(GtArgs a) {
  constexpr int NT = NT0 + NT1;
  constexpr int KS = 2;                                    // k-steps (64 samples) per staged block
  constexpr int ST = 4;                                    // SNP tiles of 16 per wave
  constexpr int CH = KS * NT;                              // 1 KB operand chunks per block
  constexpr int LDS_A = 2 * CH * 1024;
  constexpr int LDS_RED = 4 * NT * 1024;                   // one SNP tile of int32 sums per wave
  constexpr int PLAN_MAX = MODE == 0 ? 3 * kMaxD * 32 + kMaxD + MAXM * MAXM + 1 : 0;   // split_plan_len bound
  constexpr int LDS_E = LDS_RED + PLAN_MAX * 8 + (int)sizeof(Recomb);
  constexpr int LDS = LDS_A > LDS_E ? LDS_A : LDS_E;
  __shared__ __align__(16) unsigned char lds[LDS];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = lane >> 4, l16 = lane & 15;
  const uint64_t base = ((uint64_t)blockIdx.x * 4 + wave) * (ST * 16);
  const unsigned char* Ag = reinterpret_cast<const unsigned char*>(a.A);
  const uint8_t* rowp[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) {
    uint64_t j = base + 16 * st + l16;
    if (j >= a.p) j = a.p - 1;
    rowp[st] = a.packed + j * a.ld + 4 * KS * g;
  }
  for (int c = wave; c < CH; c += 4) chunk_to_lds(Ag + (uint64_t)c * 1024, lds + c * 1024, lane);
  uint32_t gw[ST][KS], gn[ST][KS];
#pragma unroll
  for (int st = 0; st < ST; ++st) load_words<KS>(rowp[st], gw[st]);
  v4i acc[ST][NT];
#pragma unroll
  for (int st = 0; st < ST; ++st)
#pragma unroll
    for (int rt = 0; rt < NT; ++rt) acc[st][rt] = v4i{0, 0, 0, 0};
  uint32_t c_pop[ST], c_hi[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) c_pop[st] = c_hi[st] = 0;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  for (int kb = 0; kb < a.nblk; ++kb) {
    const bool more = kb + 1 < a.nblk;
    if (more) {
      unsigned char* dst = lds + ((kb + 1) & 1) * CH * 1024;
      const unsigned char* src = Ag + (uint64_t)(kb + 1) * CH * 1024;
      for (int c = wave; c < CH; c += 4) chunk_to_lds(src + (uint64_t)c * 1024, dst + c * 1024, lane);
#pragma unroll
      for (int st = 0; st < ST; ++st) load_words<KS>(rowp[st] + (uint64_t)(kb + 1) * 16 * KS, gn[st]);
    }
    if (kb == a.nblk - 1) {   // bytes at or past the row stride belong to the next row
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int64_t lim = (int64_t)a.ld - ((int64_t)kb * 16 * KS + 4 * KS * g + 4 * s);
        const uint32_t mask = lim >= 4 ? 0xFFFFFFFFu : (lim <= 0 ? 0u : ((1u << (8 * lim)) - 1u));
#pragma unroll
        for (int st = 0; st < ST; ++st) gw[st][s] &= mask;
      }
    }
    const v4i* Ab = reinterpret_cast<const v4i*>(lds + (kb & 1) * CH * 1024);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      Frag f[ST];
#pragma unroll
      for (int st = 0; st < ST; ++st) f[st] = decode_word(gw[st][s], c_pop[st], c_hi[st]);
#pragma unroll
      for (int rt = 0; rt < NT; ++rt) {
        const v4i af = Ab[(s * NT + rt) * 64 + lane];
#pragma unroll
        for (int st = 0; st < ST; ++st)
          acc[st][rt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, rt < NT0 ? f[st].g0 : f[st].miss, acc[st][rt], 0, 0, 0);
      }
    }
    if (more) {
#pragma unroll
      for (int st = 0; st < ST; ++st)
#pragma unroll
        for (int s = 0; s < KS; ++s) gw[st][s] = gn[st][s];
    }
    __builtin_amdgcn_s_waitcnt(0);   // the LDS copy of block kb+1 has landed
    __syncthreads();
  }

  // ---- epilogue --------------------------------------------------------------
  const double* plan = nullptr;
  if constexpr (MODE == 0) {
    double* sp = reinterpret_cast<double*>(lds + LDS_RED);
    stage_split_plan(a.ea.plan, a.ea.d, a.ea.m, sp, t, 256);
    plan = sp;
  }
  const Recomb* rc = stage_recomb(a.rc, lds + LDS_RED + PLAN_MAX * 8, t, 256);
  __syncthreads();
  gt_epilogue<ST, NT, MAXM, MODE>(a, acc, c_pop, c_hi, reinterpret_cast<int*>(lds) + wave * (NT * 256), plan, rc,
                                  base, g, l16);
}

// ---------------------------------------------------------------------------
// Pipelined kernel for NT <= 8: 8 waves x 4 tiles of 16 SNPs per workgroup,
// blocks of 4 k-steps (256 samples, 64 bytes per SNP row).  Both streams go
// to LDS by buffer-descriptor DMA (buffer_load_dwordx4 ... lds): the
// genotype segments into a 3-slot ring (blocks kb+1 and kb+2 in flight while
// kb is computed), the operand blocks into a 2-slot ring.  The block loop is
// unrolled over the ring period (6) so every slot is a distinct LDS object.
// ---------------------------------------------------------------------------
constexpr int kPipeWaves = 8;
constexpr int kPipeTiles = 4;   // SNP tiles of 16 per wave
constexpr int kOperandPrefetch = 6;   // operand fragments read ahead of their MFMAs

template <int ST>
__device__ __forceinline__ void lds_read_tiles(const unsigned char* p, v4i (&q)[ST]) {
  const uint32_t addr = (uint32_t)(uintptr_t)(lds_ptr_t)p;
  static_assert(ST >= 1 && ST <= 4, "tile count");
  if constexpr (ST == 1) {
    asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=&v"(q[0]) : "v"(addr) : "memory");
  } else if constexpr (ST == 2) {
    asm volatile("ds_read_b128 %0, %2\n\tds_read_b128 %1, %2 offset:1024\n\ts_waitcnt lgkmcnt(0)"
                 : "=&v"(q[0]), "=&v"(q[1]) : "v"(addr) : "memory");
  } else if constexpr (ST == 3) {
    asm volatile("ds_read_b128 %0, %3\n\tds_read_b128 %1, %3 offset:1024\n\t"
                 "ds_read_b128 %2, %3 offset:2048\n\ts_waitcnt lgkmcnt(0)"
                 : "=&v"(q[0]), "=&v"(q[1]), "=&v"(q[2]) : "v"(addr) : "memory");
  } else {
    asm volatile("ds_read_b128 %0, %4\n\tds_read_b128 %1, %4 offset:1024\n\t"
                 "ds_read_b128 %2, %4 offset:2048\n\tds_read_b128 %3, %4 offset:3072\n\ts_waitcnt lgkmcnt(0)"
                 : "=&v"(q[0]), "=&v"(q[1]), "=&v"(q[2]), "=&v"(q[3]) : "v"(addr) : "memory");
  }
}

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
// LDS-DMA through a buffer descriptor: lane l's 16 bytes at byte offset voff
// land at dst + 16 l.  (The buffer form keeps the compiler's LDS-read waits
// partial: after a global_load_lds it waits for every outstanding LDS read.)
__device__ __forceinline__ void buf_to_lds(rsrc_t rs, uint32_t voff, unsigned char* dst) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_ptr_t)dst, 16, voff, 0, 0, 0);
}

struct PipeCtx {
  rsrc_t geno;        // this workgroup's SNP rows; rows past the last SNP read zeros
  rsrc_t opnd;        // the operand blocks
  uint32_t gvoff;     // byte offset of this lane's 16-byte piece of its tile-0 row
  uint32_t tstride;   // 16 rows
};

// One block of the pipelined kernel.  Operand block kb+1 and genotype block
// kb+2 are issued while kb is computed, in that order, so the wait at the top
// of block kb leaves only the genotypes of kb+1 in flight.
template <int NT0, int NT, int ST>
__device__ __forceinline__ void pipe_block(int kb, const GtArgs& a, const PipeCtx& pc, int wave, int lane, int g,
                                           const unsigned char* a_rd, unsigned char* a_wr,
                                           const unsigned char* g_rd, unsigned char* g_wr,
                                           v4i (&acc)[ST][NT], uint32_t (&c_pop)[ST], uint32_t (&c_hi)[ST]) {
  constexpr int KS = 4;
  constexpr int W = kPipeWaves;
  constexpr int CH = KS * NT;
  wait_vmcnt<ST>();
  __builtin_amdgcn_s_barrier();
  const v4i* Ab = reinterpret_cast<const v4i*>(a_rd);
  uint32_t gw[ST][KS];
  {
    v4i q[ST];
    lds_read_tiles<ST>(g_rd + wave * (ST * 1024) + 16 * lane, q);
#pragma unroll
    for (int st = 0; st < ST; ++st)
#pragma unroll
      for (int s = 0; s < KS; ++s) gw[st][s] = (uint32_t)q[st][s];
  }
  if (kb == a.nblk - 1) {   // bytes at or past the row stride belong to the next row
    asm volatile("" ::: "memory");   // keep this a branch: only the last block pays for the mask
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int64_t lim = (int64_t)a.ld - ((int64_t)kb * 64 + 16 * g + 4 * s);
      const uint32_t mask = lim >= 4 ? 0xFFFFFFFFu : (lim <= 0 ? 0u : ((1u << (8 * lim)) - 1u));
#pragma unroll
      for (int st = 0; st < ST; ++st) gw[st][s] &= mask;
    }
  }
  // the ring slots written below are not the ones read in this block
  if (kb + 1 < a.nblk) {
    const uint32_t src = (uint32_t)(kb + 1) * CH * 1024 + 16 * lane;
#pragma unroll
    for (int i = 0; i < (CH + W - 1) / W; ++i) {   // fragments wave, wave + W, ... (compile-time trip count)
      const int c = wave + i * W;
      if (c < CH) buf_to_lds(pc.opnd, src + c * 1024, a_wr + c * 1024);
    }
  }
  {
    const uint32_t src = (uint32_t)(kb + 2 < a.nblk ? kb + 2 : 0) * 64;
    unsigned char* dst = g_wr + wave * (ST * 1024);
#pragma unroll
    for (int st = 0; st < ST; ++st) buf_to_lds(pc.geno, pc.gvoff + st * pc.tstride + src, dst + st * 1024);
  }
  // operand fragment c = s NT + rt is read kOperandPrefetch fragments ahead;
  // the genotype fragments of k-step s + 1 are decoded after the MFMAs of s
  constexpr int D = kOperandPrefetch < CH ? kOperandPrefetch : CH;
  v4i af[CH];
#pragma unroll
  for (int c = 0; c < D; ++c) af[c] = Ab[c * 64 + lane];
  Frag fc[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) fc[st] = decode_word(gw[st][0], c_pop[st], c_hi[st]);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
#pragma unroll
    for (int rt = 0; rt < NT; ++rt) {
      const int c = s * NT + rt;
      if (c + D < CH) af[c + D] = Ab[(c + D) * 64 + lane];
#pragma unroll
      for (int st = 0; st < ST; ++st)
        acc[st][rt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[c], rt < NT0 ? fc[st].g0 : fc[st].miss, acc[st][rt], 0,
                                                            0, 0);
    }
    if (s + 1 < KS) {
#pragma unroll
      for (int st = 0; st < ST; ++st) fc[st] = decode_word(gw[st][s + 1], c_pop[st], c_hi[st]);
    }
  }
}

template <int NT0, int NT1, int MAXM, int MODE>
__global__ void __launch_bounds__(512, 1) k_gt_i8p(GtArgs a) {
  constexpr int NT = NT0 + NT1;
  constexpr int KS = 4;
  constexpr int ST = kPipeTiles;
  constexpr int W = kPipeWaves;
  constexpr int CH = KS * NT;
  constexpr int A_SLOT = CH * 1024;
  constexpr int G_SLOT = W * ST * 1024;
  static_assert(NT * 1024 * W <= 3 * G_SLOT, "epilogue sums of all waves must fit the genotype ring");
  static_assert(sizeof(Recomb) <= A_SLOT, "recombination table must fit an operand slot");
  __shared__ __align__(16) unsigned char A0[A_SLOT];
  __shared__ __align__(16) unsigned char A1[A_SLOT];
  __shared__ __align__(16) unsigned char G0[G_SLOT];
  __shared__ __align__(16) unsigned char G1[G_SLOT];
  __shared__ __align__(16) unsigned char G2[G_SLOT];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = lane >> 4, l16 = lane & 15;
  const uint64_t wgbase = (uint64_t)blockIdx.x * W * ST * 16;
  const uint64_t base = wgbase + (uint64_t)wave * (ST * 16);
  PipeCtx pc;
  const uint64_t gbytes = (a.p - wgbase) * a.ld;
  pc.geno = make_rsrc(a.packed + wgbase * a.ld, (uint32_t)(gbytes < 0xFFFFFFFFull ? gbytes : 0xFFFFFFFFull));
  pc.opnd = make_rsrc(a.A, (uint32_t)a.nblk * CH * 1024);
  pc.gvoff = (uint32_t)((base - wgbase + l16) * a.ld + 16 * g);
  pc.tstride = (uint32_t)(16 * a.ld);
  v4i acc[ST][NT];
#pragma unroll
  for (int st = 0; st < ST; ++st)
#pragma unroll
    for (int rt = 0; rt < NT; ++rt) acc[st][rt] = v4i{0, 0, 0, 0};
  uint32_t c_pop[ST], c_hi[ST];
#pragma unroll
  for (int st = 0; st < ST; ++st) c_pop[st] = c_hi[st] = 0;
  // Prologue in the order the blocks issue them: operand 0 and genotypes 0,
  // then genotypes 1.  The first block's wait leaves the genotypes of block 1 in flight.
  for (int c = wave; c < CH; c += W) buf_to_lds(pc.opnd, c * 1024 + 16 * lane, A0 + c * 1024);
#pragma unroll
  for (int st = 0; st < ST; ++st) buf_to_lds(pc.geno, pc.gvoff + st * pc.tstride, G0 + wave * (ST * 1024) + st * 1024);
  {
    const uint32_t src = a.nblk > 1 ? 64 : 0;
#pragma unroll
    for (int st = 0; st < ST; ++st)
      buf_to_lds(pc.geno, pc.gvoff + st * pc.tstride + src, G1 + wave * (ST * 1024) + st * 1024);
  }
#define SLFMM_PIPE(PH, AR_, AW_, GR_, GW_) \
  if (kb + PH < a.nblk) pipe_block<NT0, NT, ST>(kb + PH, a, pc, wave, lane, g, AR_, AW_, GR_, GW_, acc, c_pop, c_hi);
  for (int kb = 0; kb < a.nblk; kb += 6) {
    SLFMM_PIPE(0, A0, A1, G0, G2)
    SLFMM_PIPE(1, A1, A0, G1, G0)
    SLFMM_PIPE(2, A0, A1, G2, G1)
    SLFMM_PIPE(3, A1, A0, G0, G2)
    SLFMM_PIPE(4, A0, A1, G1, G0)
    SLFMM_PIPE(5, A1, A0, G2, G1)
  }
#undef SLFMM_PIPE
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  // ---- epilogue: per-wave int32 sums of one SNP tile; waves spread over the
  // three genotype slots, the plan goes to A0, the recombination table to A1
  constexpr int WPS = G_SLOT / (NT * 1024);   // waves per slot
  unsigned char* gslot = wave < WPS ? G0 : (wave < 2 * WPS ? G1 : G2);
  int* red = reinterpret_cast<int*>(gslot) + (wave % WPS) * (NT * 256);
  const double* plan = nullptr;
  if constexpr (MODE == 0) {
    double* sp = reinterpret_cast<double*>(A0);
    stage_split_plan(a.ea.plan, a.ea.d, a.ea.m, sp, t, 512);
    plan = sp;
  }
  const Recomb* rc = stage_recomb(a.rc, A1, t, 512);
  __syncthreads();
  gt_epilogue<ST, NT, MAXM, MODE>(a, acc, c_pop, c_hi, red, plan, rc, base, g, l16);
}

#define SLFMM_GT_CASES(X) \
  X(1, 1) X(2, 1) X(3, 1) X(2, 2) X(3, 2) X(4, 2) X(3, 3) X(4, 3) X(5, 3) X(4, 4) X(5, 4) X(6, 4)

template <int A0, int A1>
static bool launch_pipe(hipStream_t s, unsigned grid, GtArgs& a, int mode, int maxm) {
  if constexpr (A0 + A1 <= 8) {
    if (mode == 1) k_gt_i8p<A0, A1, 1, 1><<<grid, 512, 0, s>>>(a);
    else if (maxm <= 24) k_gt_i8p<A0, A1, 24, 0><<<grid, 512, 0, s>>>(a);
    else if (maxm <= 32) k_gt_i8p<A0, A1, 32, 0><<<grid, 512, 0, s>>>(a);
    else return false;
    return true;
  }
  return false;
}

static void launch_gt(hipStream_t s, const Operand& op, GtArgs& a, int mode, int maxm) {
  if (a.p == 0) return;
  const int key = op.NT0 * 100 + op.NT1;
  const int NT = op.NT0 + op.NT1;
  bool done = false;
  if (NT <= 8) {
    // buffer descriptors address a workgroup's rows and the operand with 32-bit byte offsets
    const uint64_t rows = 16 * kPipeTiles * kPipeWaves;
    SLFMM_CHECK(rows * a.ld + 64 * (uint64_t)a.nblk < 0xFFFFFFFFull &&
                    (uint64_t)a.nblk * 4 * NT * 1024 < 0xFFFFFFFFull,
                SLFMM_E_ARG, "panel rows too long for the int8 association kernel (use SLFMM_PREC_FP64)");
    const unsigned grid = (unsigned)cdiv(a.p, rows);
#define X(A0, A1)                                                                      \
    if (!done && key == A0 * 100 + A1) done = launch_pipe<A0, A1>(s, grid, a, mode, maxm);
    SLFMM_GT_CASES(X)
#undef X
  } else {
    const unsigned grid = (unsigned)cdiv(a.p, 256);
#define X(A0, A1)                                                                      \
    if (!done && key == A0 * 100 + A1 && A0 + A1 > 8) {                                \
      if (mode == 1) k_gt_i8<A0, A1, 1, 1><<<grid, 256, 0, s>>>(a);                    \
      else if (maxm <= 24) k_gt_i8<A0, A1, 24, 0><<<grid, 256, 0, s>>>(a);             \
      else if (maxm <= 32) k_gt_i8<A0, A1, 32, 0><<<grid, 256, 0, s>>>(a);             \
      done = true;                                                                     \
    }
    SLFMM_GT_CASES(X)
#undef X
  }
  SLFMM_CHECK(done, SLFMM_E_ARG, "no int8 kernel for this tile shape (K + d too large)");
  SLFMM_HIP(hipGetLastError());
}

// Association pass over every SNP of g (fused epilogue).
void launch_func_i8(slfmm_ctx* ctx, const slfmm_geno* g, const double* d_Q, uint64_t m, const EpiArgs& ea,
                    bool /*fused_epilogue*/) {
  // Latent directions: 2 int8 terms; covariate directions (the t-statistic
  // numerators, last d columns of the basis): 3 terms while K + d <= 32.
  SLFMM_CHECK(m <= 32, SLFMM_E_ARG, "the int8 association path supports K + d <= 32 (use SLFMM_PREC_FP64)");
  Operand op = prepare_operand(ctx, "as_i8", d_Q, m, g->n, (int)m, ea.K);
  const uint64_t plan_bytes = 8 * (uint64_t)split_plan_len(ea.d, (int)m);
  SLFMM_CHECK(op.NT0 + op.NT1 > 8 || plan_bytes <= (uint64_t)4 * (op.NT0 + op.NT1) * 1024, SLFMM_E_ARG,
              "association plan does not fit the int8 kernel's LDS (use SLFMM_PREC_FP64)");
  GtArgs a{};
  a.packed = g->d; a.ld = g->ld; a.p = g->p; a.n = g->n; a.nblk = (int)op.nblk; a.m = (int)m; a.norm = ea.norm;
  a.A = reinterpret_cast<const v4i*>(op.pack); a.rc = op.rc; a.ea = ea;
  slfmm_timed tm(ctx, "assoc_stream");
  launch_gt(ctx->stream, op, a, 0, (int)m);
}

// RSVD sketch / backward pass: Z[j][c] = s_j * (B^T y_j)_c for every SNP row.
void rsvd_project_i8(slfmm_ctx* ctx, const uint8_t* packed, uint64_t ld, uint64_t n, uint64_t p, const double* d_B,
                     uint64_t l, int norm, double* Z, uint64_t ldZ) {
  SLFMM_CHECK(l >= 1 && l <= 32, SLFMM_E_ARG, "the int8 RSVD path supports K + oversampling <= 32");
  Operand op = prepare_operand(ctx, "rs_i8", d_B, l, n, (int)l, (int)l);
  GtArgs a{};
  a.packed = packed; a.ld = ld; a.p = p; a.n = n; a.nblk = (int)op.nblk; a.m = (int)l; a.norm = norm;
  a.A = reinterpret_cast<const v4i*>(op.pack); a.rc = op.rc; a.Z = Z; a.ldZ = ldZ;
  launch_gt(ctx->stream, op, a, 1, 1);
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

// WV waves x 4 tiles of 16 samples per workgroup share each staged B block.
constexpr int kFwdWaves = 8;
template <int NCB>
__global__ void __launch_bounds__(64 * kFwdWaves) k_fwd_i8(const uint8_t* __restrict__ YT, uint64_t ldt, uint64_t n,
                                                           int blk0, int nblk_chunk, int nblk_total,
                                                           const v4i* __restrict__ Bp, const double* __restrict__ csc,
                                                           int l, double* __restrict__ part, uint64_t ldp) {
  constexpr int NTB = 4 * NCB;
  constexpr int KS = 4;
  constexpr int ST = 4;
  constexpr int NTH = 64 * kFwdWaves;
  constexpr int BLK = KS * NTB * 64;
  __shared__ v4i Bs[2 * BLK];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = lane >> 4, l16 = lane & 15;
  const uint64_t sbase = (uint64_t)blockIdx.x * (ST * 16 * kFwdWaves) + wave * (ST * 16);
  const int b0 = blk0 + blockIdx.y * nblk_chunk;
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
  constexpr int PER = (BLK + NTH - 1) / NTH;
  if (b0 < b1) {
    for (int q = 0; q < PER; ++q) {
      int idx = t + NTH * q;
      if (idx < BLK) Bs[idx] = Bp[(uint64_t)b0 * BLK + idx];
    }
  }
  uint32_t gw[ST][KS], gn[ST][KS];
  if (b0 < b1)
#pragma unroll
    for (int st = 0; st < ST; ++st) load_words<KS>(rowp[st] + (uint64_t)b0 * 16 * KS, gw[st]);
  __syncthreads();
  for (int kb = b0; kb < b1; ++kb) {
    const bool more = kb + 1 < b1;
    v4i pre[PER];
    if (more) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        int idx = t + NTH * q;
        if (idx < BLK) pre[q] = Bp[(uint64_t)(kb + 1) * BLK + idx];
      }
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
        for (int ct = 0; ct < 2 * NCB; ++ct)
          acc[st][ct] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f.g0, bf[ct], acc[st][ct], 0, 0, 0);
#pragma unroll
        for (int ct = 2 * NCB; ct < NTB; ++ct)
          acc[st][ct] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f.miss, bf[ct], acc[st][ct], 0, 0, 0);
      }
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        int idx = t + NTH * q;
        if (idx < BLK) Bs[((kb + 1 - b0) & 1) * BLK + idx] = pre[q];
      }
#pragma unroll
      for (int st = 0; st < ST; ++st)
#pragma unroll
        for (int s = 0; s < KS; ++s) gw[st][s] = gn[st][s];
    }
    __syncthreads();
  }
  (void)dummy0; (void)dummy1; (void)dummy2;
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
        part[((uint64_t)blockIdx.y * n + k) * ldp + col] = v;
      }
    }
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
  const uint64_t sblocks = cdiv(n, 64 * kFwdWaves);
  // about 512 workgroups: fewer SNP chunks mean fewer partial sums to reduce
  // (measured at n = 10 000, 100 000 SNPs: 512 beats 256, 1024 and 2048)
  uint64_t nchunk = std::max<uint64_t>(1, std::min<uint64_t>(nblk, cdiv(512, sblocks)));
  const uint64_t per = cdiv(nblk, nchunk);
  nchunk = cdiv(nblk, per);
  const uint64_t ldp = l;
  double* part = ctx->buf<double>("fw_part", nchunk * n * ldp);
  dim3 grid((unsigned)sblocks, (unsigned)nchunk);
  if (NCB == 1)
    k_fwd_i8<1><<<grid, 64 * kFwdWaves, 0, s>>>(YT, ldt, n, 0, (int)per, (int)nblk, reinterpret_cast<const v4i*>(Bp), csc,
                                     (int)l, part, ldp);
  else
    k_fwd_i8<2><<<grid, 64 * kFwdWaves, 0, s>>>(YT, ldt, n, 0, (int)per, (int)nblk, reinterpret_cast<const v4i*>(Bp), csc,
                                     (int)l, part, ldp);
  k_fwd_reduce<<<(unsigned)cdiv(n * l, 256), 256, 0, s>>>(part, n, (int)l, ldp, (int)nchunk, vhat, W, ldw);
  SLFMM_HIP(hipGetLastError());
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
