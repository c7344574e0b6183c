// noalloc.hip -- the annealing entry point that allocates no device memory.
//
// Callers that cannot call hipMalloc (code under stream capture, tight
// serving loops, harnesses that forbid allocation) anneal through
// pmmc_anneal_noalloc: the same kernels, problem preparation and results as
// pmmc_heatmap_anneal / pmmc_arcs_anneal / pmmc_smooth_anneal, with every
// device buffer carved from one file-scope device array of kScratchBytes
// instead of the context's arena.  Each GPU holds its own copy of the array
// for the life of the process; its device address is looked up once per
// device, when the first context on that device is created.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "engine_host.hpp"

// 64 MiB per device: the arc headline problem (400 beads) fits with 13,000
// chains, the 19,951-bead smooth segment with 256
static constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace pmmc {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char *g_ptr[kMaxDevices] = {};  // scratch's device address, per device

}  // namespace

// Called by pmmc_ctx_create with `device` current.  The host-side name of the
// array is only a stand-in; its device address comes from the runtime.
pmmc_status scratch_init(pmmc_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(ctx, PMMC_ERR_INVALID, "scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[ctx->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(ctx, PMMC_ERR_HIP, std::string("scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    }
    g_ptr[ctx->device] = static_cast<char *>(p);
  }
  ctx->scratch = g_ptr[ctx->device];
  ctx->scratch_cap = kScratchBytes;
  return PMMC_OK;
}

}  // namespace pmmc

using namespace pmmc;

extern "C" pmmc_status pmmc_anneal_noalloc(pmmc_ctx *ctx, int32_t level, const void *problem,
                                           const pmmc_schedule *s, const pmmc_run *run,
                                           float *pos_out, pmmc_anneal_stats *stats) {
  if (!ctx || !problem || !s || !run || !pos_out) return fail(ctx, PMMC_ERR_INVALID, "noalloc: null argument");
  if (!ctx->scratch) return fail(ctx, PMMC_ERR_HIP, "noalloc: device scratch not set up: " + ctx->scratch_error);
  std::vector<HostProblem> hp(1);
  pmmc_status ps = prepare_problem(ctx, level, problem, *s, *run, hp[0]);
  if (ps != PMMC_OK) return ps;
  const size_t cap = ctx->scratch_cap;
  if (run->n_chains <= 0 && batch_bytes(hp) > cap) {
    // the default chain count, reduced to the most whole wavefronts that fit
    int lo = 1, hi = (hp[0].n_chains + 63) / 64;
    while (lo < hi) {
      const int mid = lo + (hi - lo + 1) / 2;
      hp[0].n_chains = mid * 64;
      if (batch_bytes(hp) <= cap) lo = mid;
      else hi = mid - 1;
    }
    hp[0].n_chains = lo * 64;
  }
  const size_t need = batch_bytes(hp);
  if (need > cap)
    return fail(ctx, PMMC_ERR_NOMEM,
                "noalloc: the problem and its " + std::to_string(hp[0].n_chains) + " chains need " +
                    std::to_string(need) + " bytes of device scratch; at most " +
                    std::to_string(cap) + " are available (fewer chains or beads)");
  std::vector<float *> outs{pos_out};
  std::vector<pmmc_anneal_stats *> st{stats};
  BatchShape sh;
  return run_problems(ctx, hp, outs, st, sh, DevSpan{ctx->scratch, cap});
}
