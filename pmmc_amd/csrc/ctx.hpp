// ctx.hpp -- the opaque pmmc_ctx (device, stream, grow-only arena) and the
// error plumbing of the C ABI.  Replaces the reference's GpuResourceCache
// (LooperSolver.hpp:54-106, ParallelMonteCarloArcs.cu:84-213): one context
// owns every device buffer, buffers only grow, and nothing is freed between
// calls.  Errors become pmmc_status codes (the reference exit()s,
// ParallelMonteCarloArcs.cu:60-70).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pmmc_hip.h"

// side streams a context creates (the server's default of four streams per
// step: the context's own and these three)
constexpr int kCtxSideStreams = 3;

struct pmmc_ctx {
  int device = 0;
  int cu_count = 0;
  size_t hbm_bytes = 0;  // the device's global memory
  hipStream_t stream = nullptr;
  // streams the server spreads a step's group launches over, created with
  // the context: HIP gives each new stream one of the process's hardware
  // queues in turn, and a hipGraph launch keeps one for its own stream from
  // then on, so streams created later can land two on one queue and run
  // their launches one after another
  std::vector<hipStream_t> side;
  uint64_t default_seed = 0;
  std::string last_error;
  // grow-only device arena, carved per call
  char *arena = nullptr;
  size_t arena_cap = 0;
  size_t arena_top = 0;
  // grow-only pinned host buffer the batch set-up stages the problems'
  // static tables in (one DMA upload; reused by every call, whose copy has
  // completed before the call returns)
  char *host_stage = nullptr;
  size_t host_stage_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  // the device's static scratch array (pmmc_anneal_noalloc): its device
  // address, looked up when the context is created, or null with the reason
  // in scratch_error
  char *scratch = nullptr;
  size_t scratch_cap = 0;
  std::string scratch_error;
};

namespace pmmc {

inline pmmc_status fail(pmmc_ctx *c, pmmc_status st, const std::string &msg) {
  if (c) c->last_error = msg;
  return st;
}

// f() with a failed host allocation turned into PMMC_ERR_NOMEM: no C++
// exception leaves an entry point of the C ABI.
template <class F>
pmmc_status no_throw(pmmc_ctx *c, F &&f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    return fail(c, PMMC_ERR_NOMEM, "host memory exhausted");
  } catch (const std::length_error &) {
    return fail(c, PMMC_ERR_INVALID, "size too large for a host array");
  }
}

#define PMMC_HIP_TRY(ctx, call)                                                          \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::pmmc::fail((ctx), PMMC_ERR_HIP,                                           \
                          std::string(#call) + ": " + hipGetErrorString(e_));           \
  } while (0)

// Ensure the arena holds `bytes`; contents are discarded when it grows.
inline pmmc_status arena_reserve(pmmc_ctx *c, size_t bytes) {
  if (bytes <= c->arena_cap) return PMMC_OK;
  if (c->arena) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, PMMC_ERR_HIP, hipGetErrorString(e));
    (void)hipFree(c->arena);
    c->arena = nullptr;
    c->arena_cap = 0;
  }
  size_t cap = bytes + (bytes >> 3) + (1u << 20);
  if (hipMalloc(&c->arena, cap) != hipSuccess) {
    c->arena = nullptr;
    (void)hipGetLastError();  // a later launch check must not see this failure
    return fail(c, PMMC_ERR_NOMEM, "device arena allocation failed");
  }
  c->arena_cap = cap;
  return PMMC_OK;
}

// Look up the device address of the static scratch array of ctx->device
// (the device must be current); sets ctx->scratch / scratch_cap.
pmmc_status scratch_init(pmmc_ctx *ctx);

// Bump allocator over a reserved span (256-byte aligned).
struct Carver {
  char *base;
  size_t top = 0;
  size_t cap;
  template <class T>
  T *take(size_t count) {
    top = (top + 255) & ~size_t(255);
    T *p = reinterpret_cast<T *>(base + top);
    top += sizeof(T) * count;
    return p;
  }
};

// Bytes a Carver needs for a list of (sizeof, count) requests.
struct Sizer {
  size_t top = 0;
  template <class T>
  void add(size_t count) {
    top = (top + 255) & ~size_t(255);
    top += sizeof(T) * count;
  }
};

}  // namespace pmmc
