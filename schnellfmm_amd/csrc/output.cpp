// Output stage: the calibrated results table written after the association
// pass (coalesce_output_tsv, src/testing.rs:740-801), with the .bim fields it
// carries (parse_bim, src/bed.rs:480-503).  Host code: several threads format
// blocks of rows, one thread writes them in order.
//
// Row layout (src/testing.rs:757-761, :777-796):
//   chr \t pos \t snp_id { \t p_c \t beta_c \t t_c } for each covariate c
//   \t r2_cov \t r2_latent \t r2_resid \n
// every number in Rust's `{:.6e}` form.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <future>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace slfmm {
namespace {

// ---- Rust `{:.6e}` ----------------------------------------------------------
// Seven significant digits of the exact decimal value of the double, rounded to
// nearest with ties to even (core::fmt's exact mode), as d.dddddd, then 'e' and
// the decimal exponent with no '+' and no leading zeros: 1.5 -> "1.500000e0",
// 1e-7 -> "1.000000e-7".  NaN -> "NaN" (no sign), infinities "inf" / "-inf",
// negative zero "-0.000000e0".
//
// Fast path: scale |x| by 10^(6-e) in 80-bit long double (64-bit significand;
// relative error below 2^-62, i.e. < 3e-12 absolute on a value below 1e7) and
// round.  A value whose discarded fraction lies within 1e-7 of one half could
// round either way under that error and takes the C library's exact
// conversion instead.
struct Pow10Table {
  static constexpr int kLo = -340, kHi = 360;
  long double v[kHi - kLo];
  Pow10Table() {
    for (int k = kLo; k < kHi; ++k) v[k - kLo] = powl(10.0L, (long double)k);
  }
  long double operator()(int k) const { return v[k - kLo]; }
};
const Pow10Table kP10;

char* put_exponent(char* s, int e) {
  *s++ = 'e';
  if (e < 0) {
    *s++ = '-';
    e = -e;
  }
  char b[4];
  int n = 0;
  do {
    b[n++] = (char)('0' + e % 10);
    e /= 10;
  } while (e);
  while (n) *s++ = b[--n];
  return s;
}

char* fmt_e6_exact(double x, char* s) {   // |x| finite, nonzero; sign already written
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.6e", std::fabs(x));
  const char* ep = std::strchr(buf, 'e');
  std::memcpy(s, buf, (size_t)(ep - buf));
  s += ep - buf;
  return put_exponent(s, std::atoi(ep + 1));
}

char* fmt_e6(double x, char* s) {
  if (std::isnan(x)) {
    std::memcpy(s, "NaN", 3);
    return s + 3;
  }
  if (std::signbit(x)) *s++ = '-';
  const double a = std::fabs(x);
  if (std::isinf(a)) {
    std::memcpy(s, "inf", 3);
    return s + 3;
  }
  if (a == 0.0) {
    std::memcpy(s, "0.000000e0", 10);
    return s + 10;
  }
  int be = 0;
  std::frexp(a, &be);
  int e = (int)std::floor((be - 1) * 0.30102999566398120);   // floor(log10 a) or one below
  long double sc = (long double)a * kP10(6 - e);
  for (int g = 0; g < 3 && sc >= 1e7L; ++g) sc = (long double)a * kP10(6 - ++e);
  for (int g = 0; g < 3 && sc < 1e6L; ++g) sc = (long double)a * kP10(6 - --e);
  if (!(sc >= 1e6L && sc < 1e7L)) return fmt_e6_exact(x, s);
  uint64_t q = (uint64_t)sc;
  const long double fr = sc - (long double)q;
  if (std::fabs((double)(fr - 0.5L)) < 1e-7) return fmt_e6_exact(x, s);
  if (fr > 0.5L) ++q;
  if (q == 10000000) {
    q = 1000000;
    ++e;
  }
  char dg[7];
  for (int i = 6; i >= 0; --i) {
    dg[i] = (char)('0' + q % 10);
    q /= 10;
  }
  *s++ = dg[0];
  *s++ = '.';
  std::memcpy(s, dg + 1, 6);
  s += 6;
  return put_exponent(s, e);
}

char* put_u64(char* s, uint64_t v) {
  char b[24];
  int n = 0;
  do {
    b[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) *s++ = b[--n];
  return s;
}

// ---- .bim fields -------------------------------------------------------------
// Whitespace-separated fields chrom, snp_id, cM, pos, allele1, allele2; the same
// validation and messages as parse_bim (src/bed.rs:480-503).
struct BimTable {
  std::string text;                   // whole file
  std::vector<uint64_t> chrom_at;     // offset, length packed as (off << 16 | len)
  std::vector<uint64_t> id_at;
  std::vector<uint64_t> pos;
};

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

void read_bim(const std::string& path, BimTable& t) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error(SLFMM_E_IO, "Failed to parse " + path + ": cannot open");
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  t.text.resize(len > 0 ? (size_t)len : 0);
  const size_t got = t.text.empty() ? 0 : std::fread(&t.text[0], 1, t.text.size(), f);
  std::fclose(f);
  if (got != t.text.size()) throw Error(SLFMM_E_IO, "Failed to parse " + path + ": short read");
  const char* b = t.text.data();
  const size_t n = t.text.size();
  size_t i = 0, line = 0;
  while (i < n) {
    size_t e = i;
    while (e < n && b[e] != '\n') ++e;
    ++line;
    size_t fo[6], fl[6];
    int nf = 0;
    for (size_t k = i; k < e;) {
      while (k < e && is_space(b[k])) ++k;
      if (k >= e) break;
      const size_t s0 = k;
      while (k < e && !is_space(b[k])) ++k;
      if (nf < 6) {
        fo[nf] = s0;
        fl[nf] = k - s0;
      }
      ++nf;
    }
    if (nf < 6)
      throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": BIM line " + std::to_string(line) + " has " +
                                     std::to_string(nf) + " fields, expected 6");
    const std::string cm(b + fo[2], fl[2]);
    char* end = nullptr;
    std::strtod(cm.c_str(), &end);
    if (*end)
      throw Error(SLFMM_E_PARSE,
                  "Failed to parse " + path + ": BIM line " + std::to_string(line) + ": invalid cM '" + cm + "'");
    // u64::from_str: optional '+', then decimal digits, no overflow
    uint64_t v = 0;
    const size_t k0 = (fl[3] > 0 && b[fo[3]] == '+') ? 1 : 0;
    bool ok = fl[3] > k0;
    for (size_t k = k0; ok && k < fl[3]; ++k) {
      const char c = b[fo[3] + k];
      const uint64_t dgt = (uint64_t)(c - '0');
      ok = c >= '0' && c <= '9' && v <= (UINT64_MAX - dgt) / 10;
      v = v * 10 + dgt;
    }
    if (!ok)
      throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": BIM line " + std::to_string(line) +
                                     ": invalid position '" + std::string(b + fo[3], fl[3]) + "'");
    if (fl[0] >= 65536 || fl[1] >= 65536)
      throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": BIM line " + std::to_string(line) + ": field too long");
    t.chrom_at.push_back((uint64_t)fo[0] << 16 | fl[0]);
    t.id_at.push_back((uint64_t)fo[1] << 16 | fl[1]);
    t.pos.push_back(v);
    i = e + 1;
  }
}

}  // namespace
}  // namespace slfmm

using namespace slfmm;

extern "C" {

int slfmm_format_e6(const double* x, uint64_t count, char* out, uint64_t stride) {
  if ((!x || !out) && count) return SLFMM_E_ARG;
  if (stride < 16) return SLFMM_E_ARG;
  for (uint64_t i = 0; i < count; ++i) {
    char* s = fmt_e6(x[i], out + i * stride);
    *s = '\0';
  }
  return SLFMM_OK;
}

int slfmm_write_tsv(slfmm_ctx* ctx, const char* path, const char* bim_path, const char* const* chrom,
                    const uint64_t* pos, const char* const* snp_id, uint64_t p, uint64_t d,
                    const char* const* cov_names, const double* pvalue, const double* beta, const double* tstat,
                    const double* r2, uint32_t flags, int n_threads) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(path && (bim_path || p == 0 || (chrom && pos && snp_id)), SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(p == 0 || (pvalue && beta && tstat && r2), SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(d >= 1, SLFMM_E_ARG, "need at least one covariate");
    SLFMM_CHECK(!(flags & SLFMM_OUT_DEVICE) || ctx, SLFMM_E_ARG, "device arrays need a context");
    BimTable bim;
    if (bim_path) {
      read_bim(bim_path, bim);
      SLFMM_CHECK(bim.pos.size() == p, SLFMM_E_ARG,
                  std::string(bim_path) + " has " + std::to_string(bim.pos.size()) + " records, expected " +
                      std::to_string(p));
    }
    // per-SNP arrays: device pointers are copied to host memory first
    std::vector<double> hp, hb, ht, hr;
    if (flags & SLFMM_OUT_DEVICE) {
      hp.resize(p * d); hb.resize(p * d); ht.resize(p * d); hr.resize(p * 3);
      const std::pair<std::vector<double>*, const double*> cp[4] = {{&hp, pvalue}, {&hb, beta}, {&ht, tstat},
                                                                    {&hr, r2}};
      for (auto& c : cp)
        if (!c.first->empty())
          SLFMM_HIP(hipMemcpyAsync(c.first->data(), c.second, c.first->size() * 8, hipMemcpyDeviceToHost,
                                   ctx->stream));
      SLFMM_HIP(hipStreamSynchronize(ctx->stream));
      pvalue = hp.data(); beta = hb.data(); tstat = ht.data(); r2 = hr.data();
    }
    FILE* f = std::fopen(path, "wb");
    SLFMM_CHECK(f, SLFMM_E_IO, std::string("Failed to create ") + path);
    struct Closer {
      FILE* f;
      ~Closer() { if (f) std::fclose(f); }
    } closer{f};
    std::string head = "chr\tpos\tsnp_id";
    for (uint64_t c = 0; c < d; ++c) {
      const std::string nm = cov_names ? std::string(cov_names[c]) : "cov_" + std::to_string(c);
      head += "\tp_" + nm + "\tbeta_" + nm + "\tt_" + nm;
    }
    head += "\tr2_cov\tr2_latent\tr2_resid\n";
    SLFMM_CHECK(std::fwrite(head.data(), 1, head.size(), f) == head.size(), SLFMM_E_IO,
                std::string("write failed: ") + path);

    const uint64_t rows_per = 32768;
    const uint64_t nblk = (p + rows_per - 1) / rows_per;
    int nt = n_threads > 0 ? n_threads : (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    nt = std::max(1, std::min<int>(nt, 64));
    auto format_block = [&](uint64_t blk) {
      const uint64_t j0 = blk * rows_per, j1 = std::min(p, j0 + rows_per);
      size_t meta = 0;
      for (uint64_t j = j0; j < j1; ++j)
        meta += bim_path ? (bim.chrom_at[j] & 0xFFFF) + (bim.id_at[j] & 0xFFFF)
                         : std::strlen(chrom[j]) + std::strlen(snp_id[j]);
      std::string out;
      out.resize(meta + (j1 - j0) * (24 + 3 + (3 * d + 3) * 16));
      char* s = &out[0];
      for (uint64_t j = j0; j < j1; ++j) {
        if (bim_path) {
          const uint64_t c = bim.chrom_at[j], i = bim.id_at[j];
          std::memcpy(s, bim.text.data() + (c >> 16), c & 0xFFFF); s += c & 0xFFFF;
          *s++ = '\t';
          s = put_u64(s, bim.pos[j]);
          *s++ = '\t';
          std::memcpy(s, bim.text.data() + (i >> 16), i & 0xFFFF); s += i & 0xFFFF;
        } else {
          const size_t lc = std::strlen(chrom[j]), li = std::strlen(snp_id[j]);
          std::memcpy(s, chrom[j], lc); s += lc;
          *s++ = '\t';
          s = put_u64(s, pos[j]);
          *s++ = '\t';
          std::memcpy(s, snp_id[j], li); s += li;
        }
        for (uint64_t c = 0; c < d; ++c) {
          *s++ = '\t'; s = fmt_e6(pvalue[j * d + c], s);
          *s++ = '\t'; s = fmt_e6(beta[j * d + c], s);
          *s++ = '\t'; s = fmt_e6(tstat[j * d + c], s);
        }
        for (int c = 0; c < 3; ++c) {
          *s++ = '\t'; s = fmt_e6(r2[j * 3 + c], s);
        }
        *s++ = '\n';
      }
      out.resize((size_t)(s - out.data()));
      return out;
    };
    std::deque<std::future<std::string>> q;
    uint64_t next = 0;
    const size_t window = (size_t)nt;   // blocks formatted at once
    auto launch = [&](uint64_t b) { return std::async(std::launch::async, format_block, b); };
    for (; next < nblk && q.size() < window; ++next) q.push_back(launch(next));
    while (!q.empty()) {
      const std::string blk = q.front().get();
      q.pop_front();
      if (next < nblk) q.push_back(launch(next++));
      SLFMM_CHECK(std::fwrite(blk.data(), 1, blk.size(), f) == blk.size(), SLFMM_E_IO,
                  std::string("write failed: ") + path);
    }
    const int rc = std::fclose(f);
    closer.f = nullptr;
    SLFMM_CHECK(rc == 0, SLFMM_E_IO, std::string("write failed: ") + path);
  });
}

}  // extern "C"
