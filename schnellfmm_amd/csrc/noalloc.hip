// Allocation-free association pass: slfmm_test_associations_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the association
// pass with every scratch buffer carved from one file-scope device array.
// Each GPU the library runs on holds its own copy of the array (kNoAllocBytes)
// until the process ends; its device address is resolved once per device
// when a context is created.  Calls that use it share it: the last call's
// contents stay, every buffer a kernel expects zeroed is cleared by the pass
// itself, and calls that overlap on one GPU (two contexts, two threads) race.
#include <mutex>

#include "internal.h"

namespace slfmm {

constexpr size_t kNoAllocBytes = 96ull << 20;   // 96 MiB per device
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kNoAllocDevices] = {};

void noalloc_register(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  SLFMM_CHECK(device >= 0 && device < kNoAllocDevices, SLFMM_E_ARG, "device index out of range");
  if (g_noalloc_ptr[device]) return;
  void* p = nullptr;
  SLFMM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));   // the device's own copy
  g_noalloc_ptr[device] = p;
}

static void* noalloc_base(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  SLFMM_CHECK(device >= 0 && device < kNoAllocDevices && g_noalloc_ptr[device], SLFMM_E_ARG,
              "allocation-free scratch not registered for this device");
  return g_noalloc_ptr[device];
}

// Routes the context's scratch requests to the static array for one call and
// restores the context afterwards (also when the call throws).  The solver
// status words and the timers of the pooled path are set aside: the status
// words come from the array, and no timing events are created.
struct ArenaScope {
  slfmm_ctx* ctx;
  int* status;
  bool profiling;
  ArenaScope(slfmm_ctx* c, void* base, size_t cap) : ctx(c), status(c->status), profiling(c->profiling) {
    ctx->arena = static_cast<unsigned char*>(base);
    ctx->arena_cap = cap;
    ctx->arena_top = 0;
    ctx->arena_map.clear();
    ctx->status = nullptr;
    ctx->profiling = false;
  }
  ~ArenaScope() {
    ctx->arena = nullptr;
    ctx->arena_cap = ctx->arena_top = 0;
    ctx->arena_map.clear();
    ctx->status = status;
    ctx->profiling = profiling;
  }
};

// Upper bound of the scratch bytes of the int8 association pass (checked
// before anything is queued): the basis, design and CholeskyQR copies
// (n x m doubles each), the packed operand, the t^2 histogram with its
// partial counts, and small matrices.
static size_t noalloc_need(uint64_t n, uint64_t d, uint64_t k) {
  const uint64_t m = d + k, mc = m + 1;
  const uint64_t dense = 8 * n * (2 * m + mc + m + d + k);    // cc, cf, Q, QR work, staged xs and u_hat
  const uint64_t pack = rup(n, 256) * 4 * 16 * 16;             // <= 16 row tiles of int8 fragments
  const uint64_t hist = 8 * d * (uint64_t)kBins + 4ull * 256 * 32768 + 8 * d * 1024;
  return (size_t)(dense + pack + hist + (8ull << 20));        // small matrices, partials, status words
}

}  // namespace slfmm

using namespace slfmm;

extern "C" int slfmm_test_associations_noalloc(slfmm_ctx* ctx, const slfmm_geno* g, const double* xs, uint64_t d,
                                               const double* u_hat, const slfmm_config* cfg, slfmm_assoc_out* out,
                                               uint32_t flags) {
  return guarded(ctx, [&] {
    SLFMM_CHECK(ctx && g && xs && u_hat && out && cfg, SLFMM_E_ARG, "NULL argument");
    SLFMM_CHECK(flags & SLFMM_OUT_DEVICE, SLFMM_E_ARG,
                "the allocation-free association pass writes device outputs (SLFMM_OUT_DEVICE)");
    SLFMM_CHECK(g->d && !g->host, SLFMM_E_ARG, "the allocation-free association pass needs a device-resident panel");
    SLFMM_CHECK(cfg->precision == SLFMM_PREC_I8 && cfg->k + d <= 32, SLFMM_E_ARG,
                "the allocation-free association pass runs the int8 kernels (SLFMM_PREC_I8, K + d <= 32)");
    SLFMM_CHECK(out->beta && out->tstat && out->r2 && out->pvalue && out->gif_per_trait && out->gif, SLFMM_E_ARG,
                "NULL output pointer");
    const size_t need = noalloc_need(g->n, d, cfg->k);
    SLFMM_CHECK(need <= kNoAllocBytes, SLFMM_E_ARG,
                "the allocation-free association pass needs about " + std::to_string(need >> 20) +
                    " MiB of scratch for n = " + std::to_string(g->n) + ", K + d = " + std::to_string(cfg->k + d) +
                    "; its static buffer holds " + std::to_string(kNoAllocBytes >> 20) + " MiB");
    ArenaScope scope(ctx, noalloc_base(ctx->device), kNoAllocBytes);
    double* dx = ctx->buf<double>("api_xs", g->n * d);
    double* du = ctx->buf<double>("api_u", g->n * cfg->k);
    SLFMM_HIP(hipMemcpyAsync(dx, xs, g->n * d * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    SLFMM_HIP(hipMemcpyAsync(du, u_hat, g->n * cfg->k * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    assoc_run(ctx, g, dx, d, du, cfg->k, cfg->lambda, *cfg, out->beta, out->tstat, out->r2, out->pvalue, out->hist,
              out->gif_per_trait, out->gif);
    ctx->finish_checks();   // reads the status words while they are still in the array
    SLFMM_HIP(hipStreamSynchronize(ctx->stream));
  });
}
