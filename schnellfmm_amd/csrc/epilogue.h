// Per-SNP association epilogue shared by the FP64 and the mixed-precision
// genotype-stream kernels (included by assoc.hip and assoc_i8.hip).
#pragma once
#include "internal.h"

namespace slfmm {

// Plan layout (doubles): P (d x m) | Vd (d) | Rm (d x m) | Fm (d x m) |
// L (m x m) | fallback flag (1.0: RSS by the quadratic form L L^T).
__host__ __device__ inline int plan_len(int d, int m) { return d * m * 3 + d + m * m + 1; }

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


// Sum over the four lane groups of a wave (lanes l, l^16, l^32, l^48).  The
// two butterfly steps give every lane the same bits: (x0 + x1) + (x2 + x3)
// up to commutation of each addition.
__device__ __forceinline__ double sum_lane_groups(double x) {
  x += __shfl_xor(x, 16, 64);
  return x + __shfl_xor(x, 32, 64);
}

// The same outputs as snp_epilogue, with the functionals of SNP j spread
// over the four lane groups: lane group g holds fq[q] = f_i for
// i = g + 4 q.  Dot products are formed per lane and summed over the lane
// groups; trait jj is finished and stored by lane group jj % 4.  Every lane
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
  double fy[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) fy[q] = g + 4 * q < m ? sc * fq[q] : 0.0;
  const double* P = sPlan;
  const double* Vd = P + d * m;
  const double* Rm = Vd + d;
  const double* Fm = Rm + d * m;
  const double* L = Fm + d * m;
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
  for (int jj = 0; jj < d; ++jj) {
    double coef = 0.0, b = 0.0, fc = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int i = g + 4 * q;
      if (i < m) {
        coef += P[jj * m + i] * fy[q];
        b += Rm[jj * m + i] * fy[q];
        fc += Fm[jj * m + i] * fy[q];
      }
    }
    coef = sum_lane_groups(coef);
    b = sum_lane_groups(b);
    fc = sum_lane_groups(fc);
    sscov += fc * fc;
    if (valid && (jj & 3) == g) {
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
  if (valid && g == 0) {
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
}

}  // namespace slfmm
