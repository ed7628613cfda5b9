// LBVH rebuild that allocates nothing: its scratch is a file-scope device
// array instead of the context's grow-only arena, for callers that cannot
// call hipMalloc (code under HIP graph capture, serving loops, harnesses that
// forbid allocation).
#include <mutex>

#include "common.hpp"
#include "lbvh.hpp"
#include "wide.hpp"

// 64 MiB: a single-mesh LBVH rebuild of up to 1.97M triangles (63-bit keys,
// 8-bit digits in 4096-key tiles: 34 B per triangle), held by every device
// the library runs on until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;

// In host code `scratch` names a host-side stand-in, not device memory: only
// the address hipGetSymbolAddress returns (g_ptr) is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` on each device, taken once per device with that
// device current (each device holds its own copy of the array).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

// Called by trace_ctx_create with `device` current.
hipError_t noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    void* p = nullptr;
    g_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[device] = g_status[device] == hipSuccess ? p : nullptr;
    g_done[device] = true;
  }
  return g_status[device];
}

extern "C" {

int trace_lbvh_rebuild_noalloc(trace_ctx* ctx, const trace_mesh* mesh, const double* h_enclosing, int key_mode,
                               trace_bvh* bvh) {
  if (!ctx) return TRACE_E_INVALID;
  if (!mesh || !bvh) TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild_noalloc: null argument");
  if (key_mode != TRACE_KEY_REF63 && key_mode != TRACE_KEY_M30)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild_noalloc: unknown key mode %d", key_mode);
  if (bvh->key_mode < 0 || bvh->base)
    TRACE_FAIL(ctx, TRACE_E_INVALID, "trace_lbvh_rebuild_noalloc: not an LBVH handle");
  if (mesh->ntri != bvh->n_prims || mesh->nseg != bvh->nseg)
    TRACE_FAIL(ctx, TRACE_E_INVALID,
               "trace_lbvh_rebuild_noalloc: mesh shape differs from the one the tree was built for");
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_done[ctx->device])
      TRACE_FAIL(ctx, TRACE_E_HIP, "trace_lbvh_rebuild_noalloc: scratch address of device %d not taken",
                 ctx->device);
    if (g_status[ctx->device] != hipSuccess)
      TRACE_FAIL(ctx, TRACE_E_HIP, "trace_lbvh_rebuild_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                 hipGetErrorString(g_status[ctx->device]));
    ptr = g_ptr[ctx->device];
  }
  // both passes (the LBVH's and the widening's) take their scratch from the
  // start of the array; refuse sizes beyond it before anything is enqueued
  const int32_t n_roots = bvh->nseg > 1 ? bvh->nseg : 1;
  const size_t need = std::max(lbvh::scratch_bytes(ctx, mesh, key_mode),
                               wide::scratch_bytes(int64_t(bvh->nodes.n), n_roots));
  if (need > kScratchBytes)
    TRACE_FAIL(ctx, TRACE_E_UNSUPPORTED,
               "trace_lbvh_rebuild_noalloc: %lld triangles need %zu bytes of scratch, more than the %zu the "
               "library holds; use trace_lbvh_rebuild",
               static_cast<long long>(mesh->ntri), need, kScratchBytes);
  Workspace ws;
  ws.base = static_cast<char*>(ptr);
  ws.bytes = kScratchBytes;
  return lbvh::build(ctx, mesh, h_enclosing, key_mode, bvh, &ws);
}

}  // extern "C"
