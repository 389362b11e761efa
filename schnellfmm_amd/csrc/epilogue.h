// Per-SNP association epilogue shared by the FP64 and the mixed-precision
// genotype-stream kernels (included by assoc.hip and assoc_i8.hip).
#pragma once
#include "internal.h"

namespace slfmm {

// Plan layout (doubles): P (d x m) | Vd (d) | Rm (d x m) | Fm (d x m) |
// L (m x m) | fallback flag (1.0: RSS by the quadratic form L L^T).
__host__ __device__ inline int plan_len(int d, int m) { return d * m * 3 + d + m * m + 1; }

// LDS layout of the plan for snp_epilogue_split: P, Rm and Fm as
// [dp][32] with dp = 4 ceil(d / 4) (zero rows past d, zero columns past m),
// Vd [dp] (1 past d), then L [m][m] and the fallback flag as in the plan.
__host__ __device__ inline int split_plan_len(int d, int m) {
  const int dp = (d + 3) & ~3;
  return 3 * dp * 32 + dp + m * m + 1;
}

// Copies the plan into the split layout (all threads of the workgroup take
// part; the caller synchronises before reading it).
__device__ inline void stage_split_plan(const double* __restrict__ plan, int d, int m, double* sp, int t, int nt) {
  const int dp = (d + 3) & ~3;
  const double* P = plan;
  const double* Vd = P + d * m;
  const double* Rm = Vd + d;
  const double* Fm = Rm + d * m;
  const double* L = Fm + d * m;
  for (int e = t; e < dp * 32; e += nt) {
    const int u = e >> 5, i = e & 31;
    const bool in = u < d && i < m;
    sp[e] = in ? P[u * m + i] : 0.0;
    sp[dp * 32 + e] = in ? Rm[u * m + i] : 0.0;
    sp[2 * dp * 32 + e] = in ? Fm[u * m + i] : 0.0;
  }
  for (int e = t; e < dp; e += nt) sp[3 * dp * 32 + e] = e < d ? Vd[e] : 1.0;
  for (int e = t; e < m * m + 1; e += nt) sp[3 * dp * 32 + dp + e] = L[e];
}

struct EpiArgs {
  uint64_t n, p;
  int d, K, m, norm;
  double df;
  const double* plan;
  const int32_t* n_miss;
  const int32_t* sum_g;
  const int32_t* n_hom;
  double* beta;   // p x d
  double* tstat;  // p x d
  double* r2;     // p x 3
  // NMF-filled columns: per-SNP scale and TSS replace the integer-count ones
  const double* fill_sc = nullptr;
  const double* fill_tss = nullptr;
};

// compute_scale (src/bed.rs:382-395); contraction-free in this file too.
__device__ __forceinline__ double epi_scale(double mean, int norm) {
  double pp = __dmul_rn(mean, 0.5);
  double twopq = __dmul_rn(__dmul_rn(2.0, pp), __dsub_rn(1.0, pp));
  if (norm == SLFMM_NORM_CENTER_ONLY) return twopq < 1e-20 ? 0.0 : 1.0;
  double den = sqrt(twopq);
  return den > 1e-10 ? 1.0 / den : 0.0;
}

template <int MAXM>
__device__ void snp_epilogue(const EpiArgs& a, const double* sPlan, uint64_t j, const double (&fu)[MAXM],
                             int64_t nmiss, int64_t sg, int64_t nh) {
  const int d = a.d, m = a.m;
  const int64_t nobs = (int64_t)a.n - nmiss;
  const double mean = nobs > 0 ? (double)sg / (double)nobs : 0.0;
  double sc = epi_scale(mean, a.norm);
  const int64_t sg2 = sg + 2 * nh;
  double tss = nobs > 0 ? sc * sc * ((double)(nobs * sg2 - sg * sg) / (double)nobs) : 0.0;
  if (a.fill_sc) {
    sc = a.fill_sc[j];
    tss = a.fill_tss[j];
  }
  double fy[MAXM];
#pragma unroll
  for (int r = 0; r < MAXM; ++r) fy[r] = r < m ? sc * fu[r] : 0.0;
  const double* P = sPlan;
  const double* Vd = P + d * m;
  const double* Rm = Vd + d;
  const double* Fm = Rm + d * m;
  const double* L = Fm + d * m;
  const bool fallback = L[m * m] != 0.0;
  double proj = 0.0;
  if (!fallback) {
#pragma unroll
    for (int r = 0; r < MAXM; ++r) proj += fy[r] * fy[r];
  } else {
    for (int c = 0; c < m; ++c) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < MAXM; ++r) if (r < m) s += L[r * m + c] * fy[r];
      proj += s * s;
    }
  }
  double rss = tss - proj;
  if (rss < 0.0) rss = 0.0;
  const double sigma2 = rss / a.df;
  double sscov = 0.0;
  for (int jj = 0; jj < d; ++jj) {
    double coef = 0.0, b = 0.0, fc = 0.0;
#pragma unroll
    for (int r = 0; r < MAXM; ++r)
      if (r < m) {
        coef += P[jj * m + r] * fy[r];
        b += Rm[jj * m + r] * fy[r];
        fc += Fm[jj * m + r] * fy[r];
      }
    sscov += fc * fc;
    // t_test (src/testing.rs:1036-1048)
    double se = sqrt(sigma2 * Vd[jj]);
    double tv = 0.0;
    if (se >= 1e-300 && isfinite(se)) {
      tv = coef / se;
      if (!isfinite(tv)) tv = 0.0;
    }
    a.tstat[j * d + jj] = tv;
    a.beta[j * d + jj] = b;
  }
  double r0 = 0.0, r1 = 0.0, r2v = 0.0;
  if (tss > 1e-300) {                    // src/testing.rs:394-404
    double sslat = tss - rss - sscov;
    if (sslat < 0.0) sslat = 0.0;
    r0 = sscov / tss;
    r1 = sslat / tss;
    r2v = rss / tss;
  }
  a.r2[j * 3 + 0] = r0;
  a.r2[j * 3 + 1] = r1;
  a.r2[j * 3 + 2] = r2v;
}


// The same outputs as snp_epilogue for one SNP per lane: all functionals of
// the SNP in the lane's registers (f[i] = f_i, zero past m), the plan in the
// split layout (stage_split_plan).  No cross-lane traffic; `valid` masks the
// stores.
template <int M>
__device__ void snp_epilogue_lane(const EpiArgs& a, const double* sPlan, uint64_t j, bool valid, const double (&f)[M],
                                  int64_t nmiss, int64_t sg, int64_t nh) {
  static_assert(M <= 32, "split plan rows hold 32 functionals");
  const int d = a.d, m = a.m;
  const int64_t nobs = (int64_t)a.n - nmiss;
  const double mean = nobs > 0 ? (double)sg / (double)nobs : 0.0;
  const double sc = epi_scale(mean, a.norm);
  const int64_t sg2 = sg + 2 * nh;
  const double tss = nobs > 0 ? sc * sc * ((double)(nobs * sg2 - sg * sg) / (double)nobs) : 0.0;
  double fy[M];   // zero for i >= m (f is)
#pragma unroll
  for (int i = 0; i < M; ++i) fy[i] = sc * f[i];
  const int dp = (d + 3) & ~3;
  const double* P = sPlan;   // split layout: rows of 32, zero past m
  const double* Rm = P + dp * 32;
  const double* Fm = Rm + dp * 32;
  const double* Vd = Fm + dp * 32;
  const double* L = Vd + dp;
  const bool fallback = L[m * m] != 0.0;
  double proj = 0.0;
  if (!fallback) {
#pragma unroll
    for (int i = 0; i < M; ++i) proj += fy[i] * fy[i];
  } else {
    for (int c = 0; c < m; ++c) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r < m) s += L[r * m + c] * fy[r];
      proj += s * s;
    }
  }
  double rss = tss - proj;
  if (rss < 0.0) rss = 0.0;
  const double sigma2 = rss / a.df;
  double sscov = 0.0;
  for (int jj = 0; jj < d; ++jj) {
    double coef = 0.0, b = 0.0, fc = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      coef += P[jj * 32 + i] * fy[i];
      b += Rm[jj * 32 + i] * fy[i];
      fc += Fm[jj * 32 + i] * fy[i];
    }
    sscov += fc * fc;
    if (valid) {
      // t_test (src/testing.rs:1036-1048)
      const double se = sqrt(sigma2 * Vd[jj]);
      double tv = 0.0;
      if (se >= 1e-300 && isfinite(se)) {
        tv = coef / se;
        if (!isfinite(tv)) tv = 0.0;
      }
      a.tstat[j * d + jj] = tv;
      a.beta[j * d + jj] = b;
    }
  }
  if (valid) {   // src/testing.rs:394-404
    double r0 = 0.0, r1 = 0.0, r2v = 0.0;
    if (tss > 1e-300) {
      double sslat = tss - rss - sscov;
      if (sslat < 0.0) sslat = 0.0;
      r0 = sscov / tss;
      r1 = sslat / tss;
      r2v = rss / tss;
    }
    a.r2[j * 3 + 0] = r0;
    a.r2[j * 3 + 1] = r1;
    a.r2[j * 3 + 2] = r2v;
  }
}

// Exchange of two doubles between the lane halves (v_permlane32_swap on both
// words): x's upper 32 lanes take y's lower 32 lanes and y's lower 32 lanes
// take x's upper 32 lanes.  With x = y every lane gets the lower partner's
// value in x and the upper partner's in y.
__device__ __forceinline__ void swap_halves(double& x, double& y) {
  const uint64_t ux = (uint64_t)__double_as_longlong(x), uy = (uint64_t)__double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ux, (uint32_t)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(ux >> 32), (uint32_t)(uy >> 32), false, false);
  x = __longlong_as_double((long long)(((uint64_t)hi[0] << 32) | lo[0]));
  y = __longlong_as_double((long long)(((uint64_t)hi[1] << 32) | lo[1]));
}

// x_lower + x_upper of a lane-swap pair: v_permlane16_swap (row r of 16 lanes
// with row r ^ 1) or v_permlane32_swap (lanes l with l ^ 32), applied to both
// words of the double.  Swapping a register with itself leaves the lower
// partner's value in element 0 and the upper partner's in element 1, so every
// lane adds the same two values in the same order.
template <int HALF>
__device__ __forceinline__ double add_swapped(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)u, hi = (uint32_t)(u >> 32);
  uint32_t l0, l1, h0, h1;
  if constexpr (HALF == 16) {
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    l0 = a[0]; l1 = a[1]; h0 = b[0]; h1 = b[1];
  } else {
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    l0 = a[0]; l1 = a[1]; h0 = b[0]; h1 = b[1];
  }
  return __longlong_as_double((long long)(((uint64_t)h0 << 32) | l0)) +
         __longlong_as_double((long long)(((uint64_t)h1 << 32) | l1));
}

// Sum over the four lane groups of a wave (lanes l, l^16, l^32, l^48):
// (x0 + x1) + (x2 + x3), the same bits in every lane.  Every lane of the wave
// must take part.
__device__ __forceinline__ double sum_lane_groups(double x) { return add_swapped<32>(add_swapped<16>(x)); }

// The same outputs as snp_epilogue, with the functionals of SNP j spread
// over the four lane groups: lane group g holds fq[q] = f_i for
// i = g + 4 q.  Dot products are formed per lane and summed over the lane
// groups; trait jj is finished and stored by lane group jj % 4, r2 fraction
// r by lane group r.  Every lane
// of the wave must call this (it shuffles); `valid` masks the stores.
template <int NQ>
__device__ void snp_epilogue_split(const EpiArgs& a, const double* sPlan, uint64_t j, bool valid, int g,
                                   const double (&fq)[NQ], int64_t nmiss, int64_t sg, int64_t nh) {
  const int d = a.d, m = a.m;
  const int64_t nobs = (int64_t)a.n - nmiss;
  const double mean = nobs > 0 ? (double)sg / (double)nobs : 0.0;
  const double sc = epi_scale(mean, a.norm);
  const int64_t sg2 = sg + 2 * nh;
  const double tss = nobs > 0 ? sc * sc * ((double)(nobs * sg2 - sg * sg) / (double)nobs) : 0.0;
  double fy[NQ];   // zero for i >= m (fq is)
#pragma unroll
  for (int q = 0; q < NQ; ++q) fy[q] = sc * fq[q];
  const int dp = (d + 3) & ~3;
  const double* P = sPlan;   // split layout (stage_split_plan)
  const double* Rm = P + dp * 32;
  const double* Fm = Rm + dp * 32;
  const double* Vd = Fm + dp * 32;
  const double* L = Vd + dp;
  const bool fallback = L[m * m] != 0.0;
  double proj = 0.0;
  if (!fallback) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) proj += fy[q] * fy[q];
    proj = sum_lane_groups(proj);
  } else {
    for (int c = 0; c < m; ++c) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (g + 4 * q < m) s += L[(g + 4 * q) * m + c] * fy[q];
      s = sum_lane_groups(s);
      proj += s * s;
    }
  }
  double rss = tss - proj;
  if (rss < 0.0) rss = 0.0;
  const double sigma2 = rss / a.df;
  double sscov = 0.0;
  for (int jb = 0; jb < d; jb += 4) {
    // the sums of traits jb .. jb + 3 reach every lane; lane group g finishes trait jb + g
    double cf[4], bt[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {   // traits past d have zero rows: they add exact zeros
      const int jj = jb + u;
      double coef = 0.0, b = 0.0, fc = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int i = g + 4 * q;
        coef += P[jj * 32 + i] * fy[q];
        b += Rm[jj * 32 + i] * fy[q];
        fc += Fm[jj * 32 + i] * fy[q];
      }
      coef = sum_lane_groups(coef);
      b = sum_lane_groups(b);
      fc = sum_lane_groups(fc);
      sscov += fc * fc;
      cf[u] = coef;
      bt[u] = b;
    }
    const int jj = jb + g;
    const double coef = g == 0 ? cf[0] : (g == 1 ? cf[1] : (g == 2 ? cf[2] : cf[3]));
    const double b = g == 0 ? bt[0] : (g == 1 ? bt[1] : (g == 2 ? bt[2] : bt[3]));
    if (valid && jj < d) {
      // t_test (src/testing.rs:1036-1048)
      const double se = sqrt(sigma2 * Vd[jj]);
      double tv = 0.0;
      if (se >= 1e-300 && isfinite(se)) {
        tv = coef / se;
        if (!isfinite(tv)) tv = 0.0;
      }
      a.tstat[j * d + jj] = tv;
      a.beta[j * d + jj] = b;
    }
  }
  if (valid && g < 3) {   // lane group g writes r2 fraction g (src/testing.rs:394-404)
    double v = 0.0;
    if (tss > 1e-300) {
      double sslat = tss - rss - sscov;
      if (sslat < 0.0) sslat = 0.0;
      v = (g == 0 ? sscov : (g == 1 ? sslat : rss)) / tss;
    }
    a.r2[j * 3 + g] = v;
  }
}

// snp_epilogue_split with the functionals of SNP j spread over the two
// lane halves of a wave (lanes l and l ^ 32): half h holds fq[q] = f_i for
// i = h + 2 q.  Dot products are formed per lane and summed over the two
// halves (v_permlane32_swap: the same bits in both lanes); trait jj is
// finished and stored by half jj % 2, r2 fractions 0 and 2 by half 0 and
// fraction 1 by half 1.  Every lane of the wave must call this; `valid`
// masks the stores.
template <int NQ>
__device__ void snp_epilogue_split2(const EpiArgs& a, const double* sPlan, uint64_t j, bool valid, int h,
                                    const double (&fq)[NQ], int64_t nmiss, int64_t sg, int64_t nh) {
  const int d = a.d, m = a.m;
  const int64_t nobs = (int64_t)a.n - nmiss;
  const double mean = nobs > 0 ? (double)sg / (double)nobs : 0.0;
  const double sc = epi_scale(mean, a.norm);
  const int64_t sg2 = sg + 2 * nh;
  const double tss = nobs > 0 ? sc * sc * ((double)(nobs * sg2 - sg * sg) / (double)nobs) : 0.0;
  double fy[NQ];   // zero for i >= m (fq is)
#pragma unroll
  for (int q = 0; q < NQ; ++q) fy[q] = sc * fq[q];
  const int dp = (d + 3) & ~3;
  const double* P = sPlan;   // split layout (stage_split_plan)
  const double* Rm = P + dp * 32;
  const double* Fm = Rm + dp * 32;
  const double* Vd = Fm + dp * 32;
  const double* L = Vd + dp;
  const bool fallback = L[m * m] != 0.0;
  double proj = 0.0;
  if (!fallback) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) proj += fy[q] * fy[q];
    proj = add_swapped<32>(proj);
  } else {
    for (int c = 0; c < m; ++c) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (h + 2 * q < m) s += L[(h + 2 * q) * m + c] * fy[q];
      s = add_swapped<32>(s);
      proj += s * s;
    }
  }
  double rss = tss - proj;
  if (rss < 0.0) rss = 0.0;
  const double sigma2 = rss / a.df;
  double sscov = 0.0;
  for (int jb = 0; jb < d; jb += 2) {
    // the sums of traits jb, jb + 1 reach both halves; half h finishes trait jb + h
    double cf[2], bt[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {   // traits past d have zero rows: they add exact zeros
      const int jj = jb + u;
      double coef = 0.0, b = 0.0, fc = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int i = h + 2 * q;
        coef += P[jj * 32 + i] * fy[q];
        b += Rm[jj * 32 + i] * fy[q];
        fc += Fm[jj * 32 + i] * fy[q];
      }
      coef = add_swapped<32>(coef);
      b = add_swapped<32>(b);
      fc = add_swapped<32>(fc);
      sscov += fc * fc;
      cf[u] = coef;
      bt[u] = b;
    }
    const int jj = jb + h;
    const double coef = h == 0 ? cf[0] : cf[1];
    const double b = h == 0 ? bt[0] : bt[1];
    if (valid && jj < d) {
      // t_test (src/testing.rs:1036-1048)
      const double se = sqrt(sigma2 * Vd[jj]);
      double tv = 0.0;
      if (se >= 1e-300 && isfinite(se)) {
        tv = coef / se;
        if (!isfinite(tv)) tv = 0.0;
      }
      a.tstat[j * d + jj] = tv;
      a.beta[j * d + jj] = b;
    }
  }
  if (valid) {   // r2 fractions (src/testing.rs:394-404): 0 and 2 by half 0, 1 by half 1
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (tss > 1e-300) {
      double sslat = tss - rss - sscov;
      if (sslat < 0.0) sslat = 0.0;
      v0 = sscov / tss;
      v1 = sslat / tss;
      v2 = rss / tss;
    }
    if (h == 0) {
      a.r2[j * 3 + 0] = v0;
      a.r2[j * 3 + 2] = v2;
    } else {
      a.r2[j * 3 + 1] = v1;
    }
  }
}

}  // namespace slfmm
