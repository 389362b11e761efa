// C ABI of the slfmm library and the host-side orchestration that mirrors
// the reference's library entry points (src/lib.rs, src/bed.rs).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>

#include "internal.h"

using namespace slfmm;

void* slfmm_ctx::scratch(const std::string& name, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (arena) {
    auto a = arena_map.find(name);
    if (a != arena_map.end() && a->second.second >= bytes) return arena + a->second.first;
    const size_t off = (arena_top + 255) & ~(size_t)255;
    if (off > arena_cap || bytes > arena_cap - off)
      throw Error(SLFMM_E_NOMEM, "allocation-free call: scratch (" + name + ", " + std::to_string(bytes) +
                                     " bytes) exceeds the " + std::to_string(arena_cap >> 20) +
                                     " MiB static buffer");
    arena_map[name] = {off, bytes};
    arena_top = off + bytes;
    return arena + off;
  }
  auto it = pool.find(name);
  if (it != pool.end() && it->second.second >= bytes) return it->second.first;
  if (capturing) throw CaptureAbort();
  ++pool_gen;
  if (it != pool.end()) {
    SLFMM_HIP(hipStreamSynchronize(stream));
    SLFMM_HIP(hipFree(it->second.first));
    pool.erase(it);
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    throw Error(SLFMM_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes (" + name + ") failed");
  }
  pool[name] = {p, bytes};
  return p;
}

void* slfmm_ctx::pinned_buf(int i, size_t bytes) {
  if (pinned[i] && pinned_bytes[i] >= bytes) return pinned[i];
  if (pinned[i]) {
    SLFMM_HIP(hipDeviceSynchronize());
    (void)hipHostFree(pinned[i]);
    pinned[i] = nullptr;
    pinned_bytes[i] = 0;
  }
  if (hipHostMalloc(&pinned[i], bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    pinned[i] = nullptr;
    throw Error(SLFMM_E_NOMEM, "pinned host allocation of " + std::to_string(bytes) + " bytes failed");
  }
  pinned_bytes[i] = bytes;
  return pinned[i];
}

void slfmm_ctx::upload(void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  if (arena) {
    SLFMM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    return;
  }
  size_t off = (up_top + 255) & ~(size_t)255;
  if (off > up_cap || bytes > up_cap - off) {   // grow once the queued uploads are done
    SLFMM_HIP(hipStreamSynchronize(stream));
    if (up_buf) (void)hipHostFree(up_buf);
    up_buf = nullptr;
    up_cap = 0;
    // room for this call's uploads so far and this one, twice over: a call
    // that repeats the same uploads then fits without growing again
    const size_t cap = std::max<size_t>(2 * (off + bytes), 1 << 20);
    if (hipHostMalloc(reinterpret_cast<void**>(&up_buf), cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      up_buf = nullptr;
      throw Error(SLFMM_E_NOMEM, "pinned host allocation of " + std::to_string(cap) + " bytes failed");
    }
    up_cap = cap;
    off = 0;
  }
  std::memcpy(up_buf + off, src, bytes);
  SLFMM_HIP(hipMemcpyAsync(dst, up_buf + off, bytes, hipMemcpyHostToDevice, stream));
  if (!ev_up) SLFMM_HIP(hipEventCreateWithFlags(&ev_up, hipEventDisableTiming));
  SLFMM_HIP(hipEventRecord(ev_up, stream));
  up_top = off + bytes;
}

void slfmm_ctx::upload_reset() {
  if (up_top && ev_up) SLFMM_HIP(hipEventSynchronize(ev_up));
  up_top = 0;
}

void slfmm::fp64_stage(slfmm_ctx* ctx, const std::string& stage, const std::string& why) {
  if (!ctx->fp64_noted.insert(stage).second) return;
  fprintf(stderr, "schnellfmm_amd: SLFMM_PREC_I8 requested, but the %s runs in FP64 (%s)\n", stage.c_str(),
          why.c_str());
}

void slfmm_ctx::ensure_side_stream() {
  if (side_stream) return;
  SLFMM_HIP(hipStreamCreateWithFlags(&side_stream, hipStreamNonBlocking));
  SLFMM_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  SLFMM_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
}

void slfmm_ctx::ensure_copy_stream() {
  if (copy_stream) return;
  SLFMM_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b) {
    SLFMM_HIP(hipEventCreateWithFlags(&ev_copied[b], hipEventDisableTiming));
    SLFMM_HIP(hipEventCreateWithFlags(&ev_consumed[b], hipEventDisableTiming));
  }
}

std::pair<hipEvent_t, hipEvent_t> slfmm_ctx::tic(const std::string&) {
  hipEvent_t a = nullptr, b = nullptr;
  SLFMM_HIP(hipEventCreate(&a));
  SLFMM_HIP(hipEventCreate(&b));
  SLFMM_HIP(hipEventRecord(a, stream));
  return {a, b};
}

void slfmm_ctx::toc(const std::string& name, std::pair<hipEvent_t, hipEvent_t> ev) {
  (void)hipEventRecord(ev.second, stream);
  timers[name].pending.push_back(ev);
}

void slfmm_ctx::collect() {
  for (auto& kv : timers) {
    for (auto& ev : kv.second.pending) {
      float ms = 0.f;
      SLFMM_HIP(hipEventSynchronize(ev.second));
      SLFMM_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
      kv.second.ms += ms;
      kv.second.launches += 1;
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    kv.second.pending.clear();
  }
}

int* slfmm_ctx::check_slot(const std::string& warn, int code, const std::string& e) {
  SLFMM_CHECK(checks.size() < (size_t)kMaxChecks, SLFMM_E_ARG, "too many solver status slots in one call");
  if (!status) status = static_cast<int*>(scratch("__status", kMaxChecks * sizeof(int)));
  if (checks.empty()) launch_fill(stream, reinterpret_cast<double*>(status), 0.0, kMaxChecks / 2);   // a kernel: graph-safe
  checks.push_back({warn, code, check_prefix + e});
  return status + (checks.size() - 1);
}

std::vector<int> slfmm_ctx::finish_checks() {
  std::vector<Check> cs;
  cs.swap(checks);
  std::vector<int> v(cs.size(), 0);
  if (cs.empty()) return v;
  SLFMM_HIP(hipMemcpyAsync(v.data(), status, cs.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
  SLFMM_HIP(hipStreamSynchronize(stream));
  for (size_t i = 0; i < cs.size(); ++i) {
    if (v[i] == 1 && !cs[i].warn.empty()) fprintf(stderr, "%s\n", cs[i].warn.c_str());
    if (v[i] >= 2 && cs[i].code != SLFMM_OK) throw Error(cs[i].code, cs[i].err);
  }
  return v;
}

slfmm_ctx::~slfmm_ctx() {
  for (slfmm_ctx* w : workers) delete w;
  if (copy_stream) (void)hipStreamSynchronize(copy_stream);
  if (side_stream) (void)hipStreamSynchronize(side_stream);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  for (int b = 0; b < 2; ++b) {
    if (ev_copied[b]) (void)hipEventDestroy(ev_copied[b]);
    if (ev_consumed[b]) (void)hipEventDestroy(ev_consumed[b]);
  }
  for (auto& kv : timers)
    for (auto& ev : kv.second.pending) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& kv : graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  for (auto& kv : pool) (void)hipFree(kv.second.first);
  for (int b = 0; b < 2; ++b)
    if (pinned[b]) (void)hipHostFree(pinned[b]);
  if (up_buf) (void)hipHostFree(up_buf);
  if (ev_up) (void)hipEventDestroy(ev_up);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  if (side_stream) (void)hipStreamDestroy(side_stream);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// Device copy of a host array through a named scratch buffer.
template <class T>
T* to_dev(slfmm_ctx* ctx, const char* name, const T* h, size_t count) {
  T* d = ctx->buf<T>(name, count);
  if (count) ctx->upload(d, h, count * sizeof(T));
  return d;
}

template <class T>
void to_host(slfmm_ctx* ctx, T* h, const T* d, size_t count) {
  if (count && h) SLFMM_HIP(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
}

// ---- .bim / .fam parsing (src/bed.rs:456-503) -------------------------------
size_t count_records(const std::string& path, const char* kind, bool bim) {
  std::ifstream f(path);
  if (!f) throw Error(SLFMM_E_IO, "Failed to parse " + path + ": cannot open");
  std::string line;
  size_t i = 0;
  while (std::getline(f, line)) {
    ++i;
    std::istringstream ss(line);
    std::vector<std::string> fld;
    std::string w;
    while (ss >> w) fld.push_back(w);
    if (fld.size() < 6)
      throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": " + kind + " line " + std::to_string(i) + " has " +
                                     std::to_string(fld.size()) + " fields, expected 6");
    char* end = nullptr;
    if (bim) {
      std::strtod(fld[2].c_str(), &end);
      if (*end) throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": BIM line " + std::to_string(i) + ": invalid cM '" + fld[2] + "'");
      for (char c : fld[3])
        if (c < '0' || c > '9') throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": BIM line " + std::to_string(i) + ": invalid position '" + fld[3] + "'");
    } else {
      long v = std::strtol(fld[4].c_str(), &end, 10);
      if (*end || v < 0 || v > 255) throw Error(SLFMM_E_PARSE, "Failed to parse " + path + ": FAM line " + std::to_string(i) + ": invalid sex '" + fld[4] + "'");
    }
  }
  return i;
}

std::string with_ext(const std::string& bed, const char* ext) {
  size_t slash = bed.find_last_of('/');
  size_t dot = bed.find_last_of('.');
  std::string stem = (dot != std::string::npos && (slash == std::string::npos || dot > slash)) ? bed.substr(0, dot) : bed;
  return stem + "." + ext;
}

struct Mapped {
  const uint8_t* p = nullptr;
  size_t len = 0;
  int fd = -1;
  ~Mapped() {
    if (p && len) munmap((void*)p, len);
    if (fd >= 0) close(fd);
  }
};

void bed_open(const std::string& path, uint64_t& n, uint64_t& p, Mapped* map) {
  n = count_records(with_ext(path, "fam"), "FAM", false);
  p = count_records(with_ext(path, "bim"), "BIM", true);
  Mapped local;
  Mapped& m = map ? *map : local;
  m.fd = open(path.c_str(), O_RDONLY);
  if (m.fd < 0) throw Error(SLFMM_E_IO, "Failed to open " + path);
  struct stat st;
  fstat(m.fd, &st);
  m.len = (size_t)st.st_size;
  if (m.len) {
    void* a = mmap(nullptr, m.len, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (a == MAP_FAILED) throw Error(SLFMM_E_IO, "mmap failed for " + path);
    m.p = (const uint8_t*)a;
    madvise(a, m.len, MADV_SEQUENTIAL);
  }
  if (m.len < 3 || m.p[0] != 0x6C || m.p[1] != 0x1B || m.p[2] != 0x01)
    throw Error(SLFMM_E_BED_MAGIC, "Invalid .bed file (bad magic bytes): " + path);
  const uint64_t expected = 3 + cdiv(n, 4) * p;
  if (m.len != expected)
    throw Error(SLFMM_E_BED_SIZE, "BED file size mismatch: expected " + std::to_string(expected) + " bytes (" +
                                      std::to_string(n) + " samples × " + std::to_string(p) + " SNPs), got " +
                                      std::to_string(m.len));
}

void check_cfg(const slfmm_config* cfg) {
  SLFMM_CHECK(cfg != nullptr, SLFMM_E_ARG, "config is NULL");
  SLFMM_CHECK(cfg->precision == SLFMM_PREC_FP64 || cfg->precision == SLFMM_PREC_I8, SLFMM_E_ARG, "unknown precision");
  SLFMM_CHECK(cfg->norm == SLFMM_NORM_CENTER_ONLY || cfg->norm == SLFMM_NORM_EIGENSTRAT, SLFMM_E_ARG, "unknown norm");
  SLFMM_CHECK(cfg->impute == SLFMM_IMPUTE_MEAN, SLFMM_E_ARG, "only mean imputation is available in this entry point");
}

void center_host(const double* x, uint64_t n, uint64_t d, int scale, double* out) {
  // src/lib.rs:102-117: subtract column means; optionally divide by sd (n-1)
  for (uint64_t j = 0; j < d; ++j) {
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += x[i * d + j];
    double mean = s / (double)n;
    for (uint64_t i = 0; i < n; ++i) out[i * d + j] = x[i * d + j] - mean;
    if (scale) {
      double v = 0.0;
      for (uint64_t i = 0; i < n; ++i) v += out[i * d + j] * out[i * d + j];
      double sd = std::sqrt(v / ((double)n - 1.0));
      if (sd > 1e-14)
        for (uint64_t i = 0; i < n; ++i) out[i * d + j] /= sd;
    }
  }
}

}  // namespace

namespace slfmm {
std::string& no_ctx_error() {
  static thread_local std::string e;
  return e;
}
}  // namespace slfmm

namespace slfmm {

// memcpy over a few host threads (large copies from a mapped file into
// pinned staging memory run at host-memory speed, not one core's)
static void parallel_memcpy(uint8_t* dst, const uint8_t* src, size_t bytes) {
  const size_t piece = (size_t)32 << 20;
  const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / piece));
  if (nt == 1) {
    if (bytes) std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (bytes + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const size_t a = (size_t)t * per, b = std::min(bytes, a + per);
    if (a < b) th.emplace_back([=] { std::memcpy(dst + a, src + a, b - a); });
  }
  for (auto& x : th) x.join();
}

uint64_t host_chunk_rows(const slfmm_geno* g) {
  if (g->chunk_snps) return g->chunk_snps;
  const uint64_t target = (uint64_t)1 << 30;   // ~1 GiB of rows per chunk
  return std::max<uint64_t>(512, (target / g->ld) / 512 * 512);
}

void host_chunk_upload(slfmm_ctx* ctx, const slfmm_geno* g, uint64_t j0, uint64_t j1, int b, uint8_t* dst) {
  const uint64_t rows = j1 - j0, bytes = rows * g->bps;
  const uint8_t* pin = g->host + j0 * g->bps;   // the rows of a .bed are contiguous
  if (!g->host_pinned) {   // pageable rows: through the staging buffer (several host threads)
    auto* stage = static_cast<uint8_t*>(ctx->pinned_buf(b, std::max<uint64_t>(bytes, 16)));
    parallel_memcpy(stage, pin, bytes);
    pin = stage;
  }
  if (rows)
    SLFMM_HIP(hipMemcpy2DAsync(dst, g->ld, pin, g->bps, g->bps, rows, hipMemcpyHostToDevice, ctx->copy_stream));
  SLFMM_HIP(hipEventRecord(ctx->ev_copied[b], ctx->copy_stream));
}

void host_rows_to_device(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* rows, uint64_t m, uint8_t* dst) {
  const uint64_t piece_rows = std::max<uint64_t>(1, ((uint64_t)256 << 20) / g->bps);
  for (uint64_t r0 = 0; r0 < m; r0 += piece_rows) {
    const uint64_t nr = std::min(piece_rows, m - r0);
    auto* pin = static_cast<uint8_t*>(ctx->pinned_buf(0, nr * g->bps));
    if (!rows) {
      parallel_memcpy(pin, g->host + r0 * g->bps, nr * g->bps);
    } else {
      for (uint64_t r = 0; r < nr; ++r)
        std::memcpy(pin + r * g->bps, g->host + (uint64_t)rows[r0 + r] * g->bps, g->bps);
    }
    SLFMM_HIP(hipMemcpy2DAsync(dst + r0 * g->ld, g->ld, pin, g->bps, g->bps, nr, hipMemcpyHostToDevice, ctx->stream));
    geno_clear_tail(ctx->stream, dst + r0 * g->ld, g->ld, nr, g->n);
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));   // the staging buffer is reused by the next piece
  }
}

}  // namespace slfmm

namespace {

// Page-locked host memory (hipHostMalloc, torch pin_memory): DMA reads it directly.
bool is_pinned(const void* ptr) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, ptr) != hipSuccess) {
    (void)hipGetLastError();   // unregistered memory may report an error: clear it
    return false;
  }
  return at.type == hipMemoryTypeHost;
}

void require_device(const slfmm_geno* g) {
  SLFMM_CHECK(!g || !g->host, SLFMM_E_ARG,
              "this entry point needs a device-resident panel (this panel is streamed from host memory)");
}

}  // namespace

extern "C" {

void slfmm_config_default(slfmm_config* c) {
  if (!c) return;
  c->k = 5; c->lambda = 1e-5; c->chunk_size = 10000; c->oversampling = 10; c->n_power_iter = 2;
  c->seed = 42; c->norm = SLFMM_NORM_CENTER_ONLY; c->scale_cov = 0; c->impute = SLFMM_IMPUTE_MEAN;
  c->precision = SLFMM_PREC_FP64;
}

void slfmm_nmf_config_default(slfmm_nmf_config* c) {
  if (!c) return;
  c->k = 5; c->n_iter = 10; c->eps = 1e-16; c->cv_rate = 0.0005; c->chunk_size = 10000; c->seed = 42;
}

const char* slfmm_version(void) { return "slfmm 0.1.0 (gfx950)"; }

int slfmm_create(int device, slfmm_ctx** out) {
  if (!out) return SLFMM_E_ARG;
  *out = nullptr;
  slfmm_ctx* c = new slfmm_ctx();
  c->device = device;
  int rc = guarded(c, [&] {
    SLFMM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    noalloc_register(device);
  });
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return SLFMM_OK;
}

void slfmm_destroy(slfmm_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->live_genos > 0) {   // panels still bound: the last slfmm_geno_free deletes it
    ctx->closed = true;
    return;
  }
  delete ctx;
}

const char* slfmm_last_error(const slfmm_ctx* ctx) { return ctx ? ctx->err.c_str() : no_ctx_error().c_str(); }

int slfmm_synchronize(slfmm_ctx* ctx) {
  if (!ctx) return SLFMM_E_ARG;
  return guarded(ctx, [&] { SLFMM_HIP(hipStreamSynchronize(ctx->stream)); });
}

int slfmm_bed_open(slfmm_ctx* ctx, const char* bed_path, uint64_t* n_samples, uint64_t* n_snps) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(bed_path && n_samples && n_snps, SLFMM_E_ARG, "NULL argument");
    bed_open(bed_path, *n_samples, *n_snps, nullptr);
  });
}

int slfmm_geno_load_bed(slfmm_ctx* ctx, const char* bed_path, const int64_t* keep, uint64_t n_keep,
                        slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && bed_path && out, SLFMM_E_ARG, "NULL argument");
    Mapped m;
    uint64_t n = 0, p = 0;
    bed_open(bed_path, n, p, &m);
    auto* g = new slfmm_geno();
    try {
      geno_upload(ctx, m.p + 3, n, p, keep, n_keep, g);
      SLFMM_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      slfmm_geno_free(g);
      throw;
    }
    *out = g;
  });
}

int slfmm_geno_upload(slfmm_ctx* ctx, const uint8_t* packed, uint64_t n_phys, uint64_t n_snps,
                      const int64_t* keep, uint64_t n_keep, slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && out && (packed || n_snps == 0), SLFMM_E_ARG, "NULL argument");
    auto* g = new slfmm_geno();
    try {
      geno_upload(ctx, packed, n_phys, n_snps, keep, n_keep, g);
      SLFMM_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      slfmm_geno_free(g);
      throw;
    }
    *out = g;
  });
}

int slfmm_geno_wrap_device(slfmm_ctx* ctx, void* d_packed, uint64_t n, uint64_t n_snps, uint64_t ld,
                           slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && out && d_packed, SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(ld % 16 == 0 && ld >= cdiv(n, 4), SLFMM_E_ARG, "row stride must be a multiple of 16 >= ceil(n/4)");
    auto* g = new slfmm_geno();
    g->ctx = ctx; g->d = (uint8_t*)d_packed; g->n = n; g->p = n_snps; g->ld = ld; g->owned = false;
    ctx->live_genos += 1;
    *out = g;
  });
}

int slfmm_geno_open_bed_host(slfmm_ctx* ctx, const char* bed_path, uint64_t chunk_snps, slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && bed_path && out, SLFMM_E_ARG, "NULL argument");
    Mapped m;
    uint64_t n = 0, p = 0;
    bed_open(bed_path, n, p, &m);
    SLFMM_CHECK(n >= 1, SLFMM_E_ARG, "a host-resident panel needs at least one sample");
    auto* g = new slfmm_geno();
    g->ctx = ctx; g->n = n; g->p = p; g->ld = geno_ld(n); g->bps = cdiv(n, 4);
    g->host = m.p + 3; g->chunk_snps = chunk_snps;
    g->map_base = (void*)m.p; g->map_len = m.len;
    m.p = nullptr; m.len = 0;   // the panel owns the mapping now (the descriptor may close)
    ctx->live_genos += 1;
    *out = g;
  });
}

int slfmm_geno_wrap_host(slfmm_ctx* ctx, const uint8_t* packed, uint64_t n, uint64_t n_snps, uint64_t chunk_snps,
                         slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && out && (packed || n_snps == 0) && n >= 1, SLFMM_E_ARG, "NULL argument");
    auto* g = new slfmm_geno();
    g->ctx = ctx; g->n = n; g->p = n_snps; g->ld = geno_ld(n); g->bps = cdiv(n, 4);
    g->host = packed; g->chunk_snps = chunk_snps;
    g->host_pinned = n_snps > 0 && is_pinned(packed) && is_pinned(packed + n_snps * g->bps - 1);
    ctx->live_genos += 1;
    *out = g;
  });
}

int slfmm_geno_read_packed(slfmm_ctx* ctx, const slfmm_geno* g, uint64_t j0, uint64_t rows, uint8_t* out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && (out || rows == 0), SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(j0 <= g->p && rows <= g->p - j0, SLFMM_E_ARG, "SNP row range out of range");
    const uint64_t bps = cdiv(g->n, 4);
    if (rows == 0) return;
    if (g->host) {
      parallel_memcpy(out, g->host + j0 * bps, rows * bps);
      if (g->n & 3) {   // padding codes of the last byte read as zero, as on the device
        const uint8_t keep = (uint8_t)((1u << (2 * (g->n & 3))) - 1u);
        for (uint64_t r = 0; r < rows; ++r) out[r * bps + bps - 1] &= keep;
      }
      return;
    }
    SLFMM_HIP(hipMemcpy2DAsync(out, bps, g->d + j0 * g->ld, g->ld, bps, rows, hipMemcpyDeviceToHost, ctx->stream));
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

void slfmm_geno_free(slfmm_geno* g) {
  if (!g) return;
  slfmm_ctx* c = g->ctx;
  if (g->host && c) {   // queued copies may still read the rows
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  }
  if (g->map_base) munmap(g->map_base, g->map_len);
  if (g->owned && g->d) {
    if (c) {
      (void)hipSetDevice(c->device);
      (void)hipStreamSynchronize(c->stream);
    }
    (void)hipFree(g->d);
  }
  delete g;
  if (c && --c->live_genos == 0 && c->closed) delete c;   // context destroyed before its last panel
}

int slfmm_geno_info(const slfmm_geno* g, uint64_t* n, uint64_t* p, uint64_t* ld) {
  if (!g) return SLFMM_E_ARG;
  if (n) *n = g->n;
  if (p) *p = g->p;
  if (ld) *ld = g->ld;
  return SLFMM_OK;
}

int slfmm_decode_stats(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t n_idx,
                       int32_t* n_miss, int32_t* sum_g, int32_t* n_hom) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && snp_idx, SLFMM_E_ARG, "NULL argument");
    for (uint64_t i = 0; i < n_idx; ++i)
      SLFMM_CHECK(snp_idx[i] >= 0 && (uint64_t)snp_idx[i] < g->p, SLFMM_E_ARG, "SNP index out of range");
    hipStream_t s = ctx->stream;
    uint8_t* Y = ctx->buf<uint8_t>("api_Y", n_idx * g->ld + kSlack);
    if (g->host) {   // rows straight from host memory
      SLFMM_HIP(hipMemsetAsync(Y, 0, n_idx * g->ld + kSlack, s));
      host_rows_to_device(ctx, g, snp_idx, n_idx, Y);
    } else {
      int64_t* didx = to_dev(ctx, "api_idx", snp_idx, n_idx);
      SLFMM_HIP(hipMemsetAsync(Y + n_idx * g->ld, 0, kSlack, s));
      geno_gather(ctx, g, didx, n_idx, Y);
    }
    int32_t* a = ctx->buf<int32_t>("api_s0", n_idx);
    int32_t* b = ctx->buf<int32_t>("api_s1", n_idx);
    int32_t* c = ctx->buf<int32_t>("api_s2", n_idx);
    launch_stats(s, Y, g->ld, g->n, n_idx, a, b, c);
    to_host(ctx, n_miss, a, n_idx);
    to_host(ctx, sum_g, b, n_idx);
    to_host(ctx, n_hom, c, n_idx);
    SLFMM_HIP(hipStreamSynchronize(s));
  });
}

static int decode_common(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t c, int norm,
                         const double* fill, double* out, int raw) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && snp_idx && out, SLFMM_E_ARG, "NULL argument");
    for (uint64_t i = 0; i < c; ++i)
      SLFMM_CHECK(snp_idx[i] >= 0 && (uint64_t)snp_idx[i] < g->p, SLFMM_E_ARG, "SNP index out of range");
    hipStream_t s = ctx->stream;
    const uint8_t* src = g->d;
    std::vector<int64_t> iota;
    const int64_t* hidx = snp_idx;
    if (g->host) {   // the requested rows from host memory, then decode them in order
      uint8_t* Y = ctx->buf<uint8_t>("api_hostrows", c * g->ld + kSlack);
      SLFMM_HIP(hipMemsetAsync(Y, 0, c * g->ld + kSlack, s));
      host_rows_to_device(ctx, g, snp_idx, c, Y);
      iota.resize(c);
      for (uint64_t i = 0; i < c; ++i) iota[i] = (int64_t)i;
      hidx = iota.data();
      src = Y;
    }
    int64_t* didx = to_dev(ctx, "api_idx", hidx, c);
    const double* dfill = fill ? to_dev(ctx, "api_fill", fill, g->n * c) : nullptr;
    double* dout = ctx->buf<double>("api_out", g->n * c);
    launch_decode_f64(s, src, g->ld, g->n, didx, c, norm, dfill, dout, raw);
    to_host(ctx, out, dout, g->n * c);
    SLFMM_HIP(hipStreamSynchronize(s));
  });
}

int slfmm_decode(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t c, int32_t norm,
                 const double* fill, double* out) {
  return decode_common(ctx, g, snp_idx, c, norm, fill, out, 0);
}

int slfmm_decode_raw(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* snp_idx, uint64_t c, double* out) {
  return decode_common(ctx, g, snp_idx, c, 0, nullptr, out, 1);
}

int slfmm_center_covariates(const double* x, uint64_t n, uint64_t d, int32_t scale, double* out) {
  if (!x || !out || n == 0) return SLFMM_E_ARG;
  center_host(x, n, d, scale, out);
  return SLFMM_OK;
}

int slfmm_safe_inv(slfmm_ctx* ctx, const double* a, uint64_t m, double* out, int32_t* regularized) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && a && out, SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(m >= 1 && m <= 64, SLFMM_E_ARG, "safe_inv: order must be 1..64");
    double* da = to_dev(ctx, "api_inv_a", a, m * m);
    double* dout = ctx->buf<double>("api_inv_o", m * m);
    safe_inv_dev(ctx, da, m, dout,
                 ctx->check_slot("Warning: matrix is singular, adding epsilon diagonal regularization",
                                 SLFMM_E_SINGULAR, "inversion failed even with epsilon regularization"));
    to_host(ctx, out, dout, m * m);
    const std::vector<int> st = ctx->finish_checks();
    if (regularized) *regularized = st[0] == 1;
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int slfmm_precompute(slfmm_ctx* ctx, const double* xs, uint64_t n, uint64_t d, double lambda, double* sigma,
                     double* d_lambda, double* q_thin, double* ridge_inv) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && xs, SLFMM_E_ARG, "NULL argument");
    double* dx = to_dev(ctx, "api_xs", xs, n * d);
    double* ds = ctx->buf<double>("api_sig", d);
    double* dl = ctx->buf<double>("api_dl", d);
    double* dq = ctx->buf<double>("api_qd", n * d);
    double* dr = ctx->buf<double>("api_rg", d * d);
    precompute_dev(ctx, dx, n, d, lambda, ds, dl, dq, dr);
    to_host(ctx, sigma, ds, d);
    to_host(ctx, d_lambda, dl, d);
    to_host(ctx, q_thin, dq, n * d);
    to_host(ctx, ridge_inv, dr, d * d);
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int slfmm_estimate_factors(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* subset_idx, uint64_t n_idx,
                           const double* xs, uint64_t d, const slfmm_config* cfg, const double* b0,
                           double* u_hat, uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && xs && u_hat, SLFMM_E_ARG, "NULL argument");
    check_cfg(cfg);
    const uint64_t n = g->n, l = cfg->k + cfg->oversampling;
    double* dx = to_dev(ctx, "api_xs", xs, n * d);
    const double* db0 = b0 ? to_dev(ctx, "api_b0", b0, n * l) : nullptr;
    double* du = (flags & SLFMM_OUT_DEVICE) ? u_hat : ctx->buf<double>("api_uhat", n * cfg->k);
    rsvd_run(ctx, g, subset_idx, subset_idx ? n_idx : g->p, dx, d, *cfg, db0, du);
    if (!(flags & SLFMM_OUT_DEVICE)) to_host(ctx, u_hat, du, n * cfg->k);
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

static void assoc_host(slfmm_ctx* ctx, const slfmm_geno* g, const double* dx, uint64_t d, const double* du,
                       const slfmm_config* cfg, slfmm_assoc_out* out, uint32_t flags) {
  const uint64_t p = g->p;
  const bool dev = flags & SLFMM_OUT_DEVICE;
  double* beta = dev ? out->beta : ctx->buf<double>("api_beta", p * d);
  double* t = dev ? out->tstat : ctx->buf<double>("api_t", p * d);
  double* r2 = dev ? out->r2 : ctx->buf<double>("api_r2", p * 3);
  double* pv = dev ? out->pvalue : ctx->buf<double>("api_p", p * d);
  uint64_t* hist = (dev && out->hist) ? out->hist : (out->hist ? ctx->buf<uint64_t>("api_hist", d * kBins) : nullptr);
  double* gpt = dev ? out->gif_per_trait : ctx->buf<double>("api_gpt", d);
  double* gif = dev ? out->gif : ctx->buf<double>("api_gif", 1);
  SLFMM_CHECK(beta && t && r2 && pv && gpt && gif, SLFMM_E_ARG, "NULL output pointer");
  assoc_run(ctx, g, dx, d, du, cfg->k, cfg->lambda, *cfg, beta, t, r2, pv, hist, gpt, gif);
  if (!dev) {
    to_host(ctx, out->beta, beta, p * d);
    to_host(ctx, out->tstat, t, p * d);
    to_host(ctx, out->r2, r2, p * 3);
    to_host(ctx, out->pvalue, pv, p * d);
    if (out->hist) to_host(ctx, out->hist, hist, d * kBins);
    to_host(ctx, out->gif_per_trait, gpt, d);
    to_host(ctx, out->gif, gif, 1);
  }
}

// Queues the device work of one fit (randomized SVD, association, GIF) on
// ctx's stream.  With device-resident panels and device outputs the work is
// captured into a HIP graph once per (shapes, configuration, pointers, pool
// generation, first status slot) and replayed by later calls: one graph
// launch instead of ~200 kernel launches.  The host inputs are uploaded
// before (outside) the graph, so a replay reads this call's values.  A
// capture that would need a new device buffer is dropped and the work is
// queued directly; the next call captures again.
extern "C++" {
namespace {
constexpr size_t kMaxGraphs = 96;   // cached fit graphs per context (per stream of a batch)

template <class T>
void key_add(std::string& k, const T& v) {
  k.append(reinterpret_cast<const char*>(&v), sizeof(T));
}
}  // namespace
}

static void queue_fit(slfmm_ctx* ctx, const slfmm_geno* est, const int64_t* idx, uint64_t n_idx,
                      const slfmm_geno* g, const double* dx, uint64_t d, const slfmm_config& cfg,
                      const double* db0, double* du, slfmm_assoc_out* out, uint32_t flags) {
  auto direct = [&] {
    rsvd_run(ctx, est, idx, n_idx, dx, d, cfg, db0, du);   // src/lib.rs:231
    assoc_host(ctx, g, dx, d, du, &cfg, out, flags);       // src/lib.rs:254
  };
  const bool graphable = (flags & SLFMM_OUT_DEVICE) && !(flags & SLFMM_EAGER) && !est->host && !g->host &&
                         !ctx->profiling && !ctx->arena;
  if (!graphable) {
    direct();
    return;
  }
  // the estimation indices go to their device buffer now; an identity subset needs none
  bool identity = idx == nullptr || n_idx == est->p;
  for (uint64_t i = 0; identity && idx && i < n_idx; ++i) identity = idx[i] == (int64_t)i;
  if (!identity) {
    int64_t* didx = ctx->buf<int64_t>("rs_Y_idx", n_idx);
    ctx->upload(didx, idx, n_idx * 8);
  }
  std::string key;
  key_add(key, est); key_add(key, est->d); key_add(key, est->n); key_add(key, est->p); key_add(key, est->ld);
  key_add(key, g); key_add(key, g->d); key_add(key, g->n); key_add(key, g->p); key_add(key, g->ld);
  key_add(key, n_idx); key_add(key, identity); key_add(key, d); key_add(key, cfg);
  key_add(key, dx); key_add(key, db0); key_add(key, du); key_add(key, *out); key_add(key, flags);
  const slfmm_tuning& tu = ctx->tuning;   // launch shapes are captured with the graph
  key_add(key, tu.snp_two_tiles_min_blocks); key_add(key, tu.fwd_target_wg); key_add(key, tu.gram_max_blocks);
  const size_t slot0 = ctx->checks.size();
  key_add(key, slot0);
  auto it = ctx->graphs.find(key);
  if (it != ctx->graphs.end() && it->second.pool_gen == ctx->pool_gen) {
    SLFMM_HIP(hipGraphLaunch(it->second.exec, ctx->stream));
    ctx->checks.insert(ctx->checks.end(), it->second.checks.begin(), it->second.checks.end());
    ++ctx->graphs_replayed;
    return;
  }
  if (it != ctx->graphs.end() || ctx->graphs.size() >= kMaxGraphs) {
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));   // an earlier replay may still run
    if (it != ctx->graphs.end()) {
      (void)hipGraphExecDestroy(it->second.exec);
      ctx->graphs.erase(it);
    } else {
      for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second.exec);
      ctx->graphs.clear();
    }
  }
  const uint64_t gen0 = ctx->pool_gen;
  SLFMM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  ctx->staged_idx = identity ? nullptr : idx;
  bool aborted = false;
  try {
    direct();
  } catch (const std::exception&) {   // CaptureAbort, or an error that the direct run reports again
    aborted = true;
  }
  ctx->capturing = false;
  ctx->staged_idx = nullptr;
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &gr);
  if (aborted || ec != hipSuccess || !gr) {
    if (gr) (void)hipGraphDestroy(gr);
    (void)hipGetLastError();
    ctx->checks.resize(slot0);
    ctx->staged_idx = identity ? nullptr : idx;   // the indices are already in their buffer
    try {
      direct();
    } catch (...) {
      ctx->staged_idx = nullptr;
      throw;
    }
    ctx->staged_idx = nullptr;
    return;
  }
  slfmm_ctx::FitGraph fg;
  fg.pool_gen = gen0;
  fg.checks.assign(ctx->checks.begin() + slot0, ctx->checks.end());
  const hipError_t ei = hipGraphInstantiate(&fg.exec, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  SLFMM_HIP(ei);
  SLFMM_HIP(hipGraphLaunch(fg.exec, ctx->stream));
  ctx->graphs.emplace(key, std::move(fg));
  ++ctx->graphs_captured;
}

int slfmm_test_associations(slfmm_ctx* ctx, const slfmm_geno* g, const double* xs, uint64_t d,
                            const double* u_hat, const slfmm_config* cfg, slfmm_assoc_out* out, uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && xs && u_hat && out, SLFMM_E_ARG, "NULL argument");
    check_cfg(cfg);
    double* dx = to_dev(ctx, "api_xs", xs, g->n * d);
    double* du = to_dev(ctx, "api_u", u_hat, g->n * cfg->k);
    assoc_host(ctx, g, dx, d, du, cfg, out, flags);
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int slfmm_fit_lfmm2(slfmm_ctx* ctx, const slfmm_geno* est, const int64_t* subset_idx, uint64_t n_idx,
                    const slfmm_geno* g, const double* x, uint64_t d, const slfmm_config* cfg, const double* b0,
                    double* u_hat, slfmm_assoc_out* out, uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && est && g && x && out, SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(est->n == g->n, SLFMM_E_ARG, "estimation and test panels must have the same samples");
    check_cfg(cfg);
    const uint64_t n = g->n, K = cfg->k, l = K + cfg->oversampling;
    std::vector<double> xs(n * d);
    center_host(x, n, d, cfg->scale_cov, xs.data());                     // src/lib.rs:170
    double* dx = to_dev(ctx, "api_xs", xs.data(), n * d);
    const double* db0 = b0 ? to_dev(ctx, "api_b0", b0, n * l) : nullptr;
    const bool dev = flags & SLFMM_OUT_DEVICE;
    double* du = (dev && u_hat) ? u_hat : ctx->buf<double>("api_uhat", n * K);
    queue_fit(ctx, est, subset_idx, subset_idx ? n_idx : est->p, g, dx, d, *cfg, db0, du, out, flags);
    if (!dev && u_hat) to_host(ctx, u_hat, du, n * K);
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int slfmm_fit_lfmm2_batch(slfmm_ctx* ctx, slfmm_problem* probs, uint64_t count, int n_streams, uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && (probs || count == 0), SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(n_streams >= 1 && n_streams <= 16, SLFMM_E_ARG, "n_streams must be 1..16");
    for (uint64_t i = 0; i < count; ++i) {
      const slfmm_problem& P = probs[i];
      const std::string tag = "problem " + std::to_string(i) + ": ";
      if (!(P.est && P.g && P.x && P.d >= 1)) throw Error(SLFMM_E_ARG, tag + "NULL argument");
      if (P.est->n != P.g->n) throw Error(SLFMM_E_ARG, tag + "estimation and test panels must have the same samples");
      try {
        check_cfg(&P.cfg);
      } catch (const Error& e) {
        throw Error(e.code, tag + e.what());
      }
    }
    while ((int)ctx->workers.size() < n_streams) {
      slfmm_ctx* w = new slfmm_ctx();
      w->device = ctx->device;
      w->tuning = ctx->tuning;
      w->batch_worker = true;
      if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
        delete w;
        throw Error(SLFMM_E_HIP, "stream creation failed");
      }
      ctx->workers.push_back(w);
    }
    // panels and inputs queued on the context's own stream come first
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<slfmm_ctx*> ws(ctx->workers.begin(), ctx->workers.begin() + n_streams);
    for (slfmm_ctx* w : ws) {
      w->tuning = ctx->tuning;
      w->checks.clear();
      w->upload_reset();
    }
    std::vector<std::vector<double>> xs(count);   // host sources of the queued copies
    const bool dev = flags & SLFMM_OUT_DEVICE;
    auto drain = [&] {
      for (slfmm_ctx* w : ws) (void)hipStreamSynchronize(w->stream);
    };
    // Longest-processing-time-first schedule: problems in decreasing
    // estimated device time, each to the stream with the least queued time.
    // Estimate: a fixed cost per fit (its chain of small dense solves,
    // tuning.batch_fit_ms) plus the genotype bytes its passes stream at
    // tuning.batch_gb_per_ms (the association pass once, the estimation rows
    // 2 q + 2 times).
    std::vector<double> cost(count);
    for (uint64_t i = 0; i < count; ++i) {
      const slfmm_problem& P = probs[i];
      const double bps = (double)((P.g->n + 3) / 4);
      const double est_rows = P.subset_idx ? (double)P.n_idx : (double)P.est->p;
      cost[i] = ctx->tuning.batch_fit_ms + ((double)P.g->p * bps + (2.0 * P.cfg.n_power_iter + 2.0) * est_rows * bps) /
                                               (ctx->tuning.batch_gb_per_ms * 1e9);
    }
    std::vector<uint64_t> order(count);
    for (uint64_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return cost[a] > cost[b]; });
    std::vector<double> load(n_streams, 0.0);
    std::vector<int> stream_of(count, 0);
    for (uint64_t i : order) {
      int best = 0;
      for (int k = 1; k < n_streams; ++k)
        if (load[k] < load[best]) best = k;
      stream_of[i] = best;
      load[best] += cost[i];
    }
    for (uint64_t i : order) {
      slfmm_problem& P = probs[i];
      slfmm_ctx* w = ws[stream_of[i]];
      w->check_prefix = "problem " + std::to_string(i) + ": ";
      try {
        const uint64_t n = P.g->n, K = P.cfg.k, l = K + P.cfg.oversampling, d = P.d;
        xs[i].resize(n * d);
        center_host(P.x, n, d, P.cfg.scale_cov, xs[i].data());
        // the next problem on this stream reuses these buffers only after this
        // problem's work (stream order)
        double* dx = to_dev(w, "api_xs", xs[i].data(), n * d);
        const double* db0 = P.b0 ? to_dev(w, "api_b0", P.b0, n * l) : nullptr;
        double* du = (dev && P.u_hat) ? P.u_hat : w->buf<double>("api_uhat", n * K);
        queue_fit(w, P.est, P.subset_idx, P.subset_idx ? P.n_idx : P.est->p, P.g, dx, d, P.cfg, db0, du, &P.out,
                  flags);
        if (!dev && P.u_hat) to_host(w, P.u_hat, du, n * K);
      } catch (const Error& e) {
        const std::string tag = w->check_prefix;
        drain();
        for (slfmm_ctx* v : ws) { v->checks.clear(); v->check_prefix.clear(); }
        throw Error(e.code, tag + e.what());
      } catch (...) {
        drain();
        for (slfmm_ctx* v : ws) { v->checks.clear(); v->check_prefix.clear(); }
        throw;
      }
    }
    std::string err;
    int code = SLFMM_OK;
    for (slfmm_ctx* w : ws) {
      w->check_prefix.clear();
      try {
        w->finish_checks();   // synchronises the stream
      } catch (const Error& e) {
        if (code == SLFMM_OK) { code = e.code; err = e.what(); }
      }
    }
    drain();
    if (code != SLFMM_OK) throw Error(code, err);
  });
}

int slfmm_nmf_estimate(slfmm_ctx* ctx, const slfmm_geno* g, const int64_t* subset_idx, uint64_t n_idx,
                       const slfmm_nmf_config* cfg, const double* omega, double* w, double* h,
                       double* cv_mae, double* cv_mae_mean, uint64_t* mask_count, uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && cfg && w && h, SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(cfg->k >= 1 && cfg->k <= 16, SLFMM_E_ARG, "NMF rank k must be 1..16");
    const uint64_t n = g->n, K = cfg->k;
    uint64_t pe = 0;
    const uint8_t* Y = gather_subset(ctx, g, subset_idx, subset_idx ? n_idx : g->p, "nmf_Y", &pe);
    const double* dom = omega ? to_dev(ctx, "api_nmf_omega", omega, n * (K + 10)) : nullptr;
    const bool dev = flags & SLFMM_OUT_DEVICE;
    double* dW = dev ? w : ctx->buf<double>("api_nmf_w", n * K);
    double* dH = dev ? h : ctx->buf<double>("api_nmf_h", pe * K);
    double* dHt = ctx->buf<double>("api_nmf_ht", pe * K);
    const uint64_t it = cfg->n_iter;
    double* dcv = ctx->buf<double>("api_nmf_cv", 3 * std::max<uint64_t>(it, 1));
    nmf_run(ctx, Y, g->ld, n, pe, *cfg, dom, dW, dHt, dcv, (flags & SLFMM_NMF_INT8) != 0);
    nmf_transpose_h(ctx->stream, dHt, pe, (int)K, dH);
    if (!dev) {
      to_host(ctx, w, dW, n * K);
      to_host(ctx, h, dH, pe * K);
    }
    std::vector<double> hcv(3 * std::max<uint64_t>(it, 1), 0.0);
    if (it) to_host(ctx, hcv.data(), dcv, 3 * it);
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t total = 0;
    for (uint64_t i = 0; i < it; ++i) {
      if (cv_mae) cv_mae[i] = hcv[3 * i];
      total += (uint64_t)hcv[3 * i + 2];
    }
    if (cv_mae_mean) *cv_mae_mean = it ? hcv[3 * (it - 1) + 1] : 0.0;
    if (mask_count) *mask_count = total;
  });
}

int slfmm_fit_lfmm2_nmf(slfmm_ctx* ctx, const slfmm_geno* est, const int64_t* subset_idx, uint64_t n_idx,
                        const slfmm_geno* g, const double* x, uint64_t d, const slfmm_config* cfg,
                        const slfmm_nmf_config* nmf, const double* b0, const double* nmf_omega, double* u_hat,
                        slfmm_assoc_out* out, double* nmf_cv_mae, double* gwas_cv, uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && est && g && x && out && cfg && nmf, SLFMM_E_ARG, "NULL argument");
    require_device(g);
    SLFMM_CHECK(est->n == g->n, SLFMM_E_ARG, "estimation and test panels must have the same samples");
    slfmm_config c = *cfg;
    c.impute = SLFMM_IMPUTE_MEAN;   // the fills are passed explicitly below
    check_cfg(&c);
    SLFMM_CHECK(nmf->k >= 1 && nmf->k <= 16, SLFMM_E_ARG, "NMF rank k must be 1..16");
    hipStream_t s = ctx->stream;
    const uint64_t n = g->n, K = c.k, l = K + c.oversampling, kn = nmf->k;
    std::vector<double> xs(n * d);
    center_host(x, n, d, c.scale_cov, xs.data());                       // src/lib.rs:170
    double* dx = to_dev(ctx, "api_xs", xs.data(), n * d);
    const double* db0 = b0 ? to_dev(ctx, "api_b0", b0, n * l) : nullptr;
    const double* dom = nmf_omega ? to_dev(ctx, "api_nmf_omega", nmf_omega, n * (kn + 10)) : nullptr;
    // NMF estimation on the estimation rows (src/lib.rs:186-198)
    uint64_t pe = 0;
    const int64_t* idx = subset_idx;
    const uint64_t nidx = subset_idx ? n_idx : est->p;
    const uint8_t* Y = gather_subset(ctx, est, idx, nidx, "nmf_Y", &pe);
    double* W = ctx->buf<double>("fit_nmf_w", n * kn);
    double* Ht = ctx->buf<double>("fit_nmf_ht", pe * kn);
    double* dcv = ctx->buf<double>("fit_nmf_cv", 3 * std::max<uint64_t>(nmf->n_iter, 1));
    nmf_run(ctx, Y, est->ld, n, pe, *nmf, dom, W, Ht, dcv, (flags & SLFMM_NMF_INT8) != 0);
    // W^+ = safe_inv(W^T W) W^T, kept as its transpose Wp = W (W^T W)^-1 (src/lib.rs:199-206)
    double* WtW = ctx->buf<double>("fit_nmf_wtw", kn * kn);
    double* Winv = ctx->buf<double>("fit_nmf_winv", kn * kn);
    double* Wp = ctx->buf<double>("fit_nmf_wp", n * kn);
    ts_gram(ctx, W, kn, kn, W, kn, kn, n, WtW);
    safe_inv_dev(ctx, WtW, kn, Winv,
                 ctx->check_slot("Warning: W^T W for NMF pinv is singular, adding epsilon diagonal regularization",
                                 SLFMM_E_SINGULAR, "W^T W for NMF pinv inversion failed even with epsilon regularization"));
    ts_gemm(s, W, kn, kn, Winv, kn, kn, n, 1.0, 0.0, Wp, kn);
    // RSVD on the NMF-filled estimation rows (NmfInRam, src/lib.rs:223-231)
    const bool dev = flags & SLFMM_OUT_DEVICE;
    double* du = (dev && u_hat) ? u_hat : ctx->buf<double>("api_uhat", n * K);
    FillSpec fill;
    fill.W = W; fill.Ht = Ht; fill.K = (int)kn; fill.n = n;
    rsvd_run(ctx, est, idx, nidx, dx, d, c, db0, du, &fill);
    // association with NMF on-the-fly fill and the GWAS-phase CV (src/lib.rs:233-254)
    NmfAssoc na;
    na.W = W; na.Wp = Wp; na.K = (int)kn;
    double* dgcv = ctx->buf<double>("fit_gwas_cv", 3);
    if (gwas_cv) {
      na.cv_rate = nmf->cv_rate; na.cv_seed = c.seed; na.chunk = c.chunk_size; na.cv_out = dgcv;
    }
    const uint64_t p = g->p;
    double* beta = dev ? out->beta : ctx->buf<double>("api_beta", p * d);
    double* t = dev ? out->tstat : ctx->buf<double>("api_t", p * d);
    double* r2 = dev ? out->r2 : ctx->buf<double>("api_r2", p * 3);
    double* pv = dev ? out->pvalue : ctx->buf<double>("api_p", p * d);
    uint64_t* hist = (dev && out->hist) ? out->hist : (out->hist ? ctx->buf<uint64_t>("api_hist", d * kBins) : nullptr);
    double* gpt = dev ? out->gif_per_trait : ctx->buf<double>("api_gpt", d);
    double* gif = dev ? out->gif : ctx->buf<double>("api_gif", 1);
    SLFMM_CHECK(beta && t && r2 && pv && gpt && gif, SLFMM_E_ARG, "NULL output pointer");
    assoc_run(ctx, g, dx, d, du, K, c.lambda, c, beta, t, r2, pv, hist, gpt, gif, &na);
    if (!dev) {
      to_host(ctx, out->beta, beta, p * d);
      to_host(ctx, out->tstat, t, p * d);
      to_host(ctx, out->r2, r2, p * 3);
      to_host(ctx, out->pvalue, pv, p * d);
      if (out->hist) to_host(ctx, out->hist, hist, d * kBins);
      to_host(ctx, out->gif_per_trait, gpt, d);
      to_host(ctx, out->gif, gif, 1);
      if (u_hat) to_host(ctx, u_hat, du, n * K);
    }
    std::vector<double> hcv(3 * std::max<uint64_t>(nmf->n_iter, 1), 0.0);
    if (nmf->n_iter) to_host(ctx, hcv.data(), dcv, 3 * nmf->n_iter);
    if (gwas_cv) to_host(ctx, gwas_cv, dgcv, 3);
    SLFMM_HIP(hipStreamSynchronize(s));
    if (nmf_cv_mae) {
      for (uint64_t i = 0; i < nmf->n_iter; ++i) nmf_cv_mae[i] = hcv[3 * i];
      // mean-imputation error on the final iteration's mask (results.nmf_cv_mean, src/lib.rs:259)
      nmf_cv_mae[nmf->n_iter] = nmf->n_iter ? hcv[3 * (nmf->n_iter - 1) + 1] : 0.0;
    }
  });
}

int slfmm_debug_mfma_i8(slfmm_ctx* ctx, const int8_t* a, const int8_t* b, int32_t* d) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && a && b && d, SLFMM_E_ARG, "NULL argument");
    probe_mfma_i8(ctx, a, b, d);
  });
}

int slfmm_debug_mfma_rate(slfmm_ctx* ctx, int32_t kind, uint64_t iters, double* tops) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && tops, SLFMM_E_ARG, "NULL argument");
    mfma_rate(ctx, kind, iters, tops);
  });
}

int slfmm_debug_i8_layout(int32_t m, int32_t k3, int32_t km, int32_t* dims, int32_t* rows, uint64_t cap) {
  return guarded(nullptr, [&] {
    SLFMM_CHECK(dims, SLFMM_E_ARG, "NULL argument");
    i8_layout(m, k3, km, dims, rows, cap);
  });
}

int slfmm_set_tuning(slfmm_ctx* ctx, const slfmm_tuning* t) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && t, SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(t->fwd_target_wg >= 1 && t->fwd_target_wg <= 65536, SLFMM_E_ARG, "fwd_target_wg must be 1..65536");
    SLFMM_CHECK(t->gram_max_blocks >= 1 && t->gram_max_blocks <= 65536, SLFMM_E_ARG,
                "gram_max_blocks must be 1..65536");
    SLFMM_CHECK(t->batch_fit_ms >= 0.0 && t->batch_gb_per_ms > 0.0, SLFMM_E_ARG,
                "batch_fit_ms must be >= 0 and batch_gb_per_ms > 0");
    ctx->tuning = *t;
    for (slfmm_ctx* w : ctx->workers) w->tuning = *t;
  });
}

int slfmm_get_tuning(slfmm_ctx* ctx, slfmm_tuning* t) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && t, SLFMM_E_ARG, "NULL argument");
    *t = ctx->tuning;
  });
}

int slfmm_profile_enable(slfmm_ctx* ctx, int32_t on) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx, SLFMM_E_ARG, "NULL context");
    ctx->profiling = on != 0;
  });
}

int slfmm_profile_read(slfmm_ctx* ctx, const char* name, double* total_ms, uint64_t* launches, int32_t reset) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && name, SLFMM_E_ARG, "NULL argument");
    ctx->collect();
    auto& t = ctx->timers[name];
    if (total_ms) *total_ms = t.ms;
    if (launches) *launches = t.launches;
    if (reset) { t.ms = 0.0; t.launches = 0; }
  });
}

int slfmm_graph_stats(slfmm_ctx* ctx, uint64_t* captured, uint64_t* replayed) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx, SLFMM_E_ARG, "NULL context");
    uint64_t c = ctx->graphs_captured, r = ctx->graphs_replayed;
    for (const slfmm_ctx* w : ctx->workers) {
      c += w->graphs_captured;
      r += w->graphs_replayed;
    }
    if (captured) *captured = c;
    if (replayed) *replayed = r;
  });
}

int slfmm_synth_bn(slfmm_ctx* ctx, uint64_t n, uint64_t n_snps, uint32_t n_pops, double fst_lo, double fst_hi,
                   double maf_lo, double maf_hi, double miss_rate, uint64_t seed, slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && out && n > 0 && n_pops > 0, SLFMM_E_ARG, "bad argument");
    SLFMM_CHECK(fst_lo > 0.0 && fst_hi < 1.0 && fst_lo <= fst_hi, SLFMM_E_ARG, "Fst range must lie in (0,1)");
    auto* g = new slfmm_geno();
    g->n = n; g->p = n_snps;
    try {
      geno_alloc(ctx, g);
      launch_synth_bn(ctx->stream, g->d, g->ld, n, n_snps, n_pops, fst_lo, fst_hi, maf_lo, maf_hi, miss_rate, seed);
      SLFMM_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      slfmm_geno_free(g);
      throw;
    }
    *out = g;
  });
}

int slfmm_synth_bn_ex(slfmm_ctx* ctx, uint64_t n, uint64_t n_snps, const slfmm_synth_spec* spec, slfmm_geno** out) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && out && spec && n > 0 && spec->n_pops > 0 && spec->n_pops <= 64, SLFMM_E_ARG, "bad argument");
    SLFMM_CHECK(spec->fst_lo > 0.0 && spec->fst_hi < 1.0 && spec->fst_lo <= spec->fst_hi, SLFMM_E_ARG,
                "Fst range must lie in (0,1)");
    SLFMM_CHECK(!spec->maf_beta || (spec->maf_a > 0.0 && spec->maf_b > 0.0 && spec->maf_min >= 0.0 &&
                                    spec->maf_min < 0.5),
                SLFMM_E_ARG, "Beta MAF parameters must be positive");
    SLFMM_CHECK(spec->snp_miss_a <= 0.0 || spec->snp_miss_b > 0.0, SLFMM_E_ARG, "per-SNP missingness Beta(a, b) needs b > 0");
    auto* g = new slfmm_geno();
    g->n = n; g->p = n_snps;
    try {
      geno_alloc(ctx, g);
      launch_synth_ex(ctx->stream, g->d, g->ld, n, n_snps, *spec);
    } catch (...) {
      slfmm_geno_free(g);
      throw;
    }
    *out = g;
  });
}

}  // extern "C"
