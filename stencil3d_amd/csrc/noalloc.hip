// noalloc.hip -- sth_leapfrog_noalloc: the LEAPFROG sweep on fields the
// caller owns, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// The only device memory the sweep needs beyond its fields is the work
// queue's ticket counters of the persistent launch.  They live in `scratch`,
// a static device array (one per GPU the library runs on, held until the
// process ends); host code addresses it through the pointer
// hipGetSymbolAddress returns for the current device, never through the
// array's name.  Every call clears the counters with a kernel (no memset),
// so calls can be captured; calls that overlap on one GPU race on them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "stencil_hip.h"
#include "sth_internal.h"

namespace {

// Eight work-queue counters, 64 bytes apart (one per XCD band; the FP32 and
// FP64 sweeps use the first).
constexpr size_t kScratchWords = 16 * sth::kXcds;
static __device__ __align__(256) uint32_t scratch[kScratchWords];
static_assert(sizeof(scratch) == STH_NOALLOC_SCRATCH_BYTES, "scratch size as documented");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t *g_scratch[kMaxDevices] = {};  // per device: the address of `scratch`
int g_cus[kMaxDevices] = {};

// The current device's copy of `scratch` and its CU count, looked up once
// per device with that device current.
int scratch_of(int dev, uint32_t **ptr, int *cus) {
  if (dev < 0 || dev >= kMaxDevices) return STH_ERR_NODEV;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return sth::detail::hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol);
    }
    int n = 0;
    e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return sth::detail::hip_status(e);
    g_scratch[dev] = (uint32_t *)p;
    g_cus[dev] = n;
  }
  *ptr = g_scratch[dev];
  *cus = g_cus[dev];
  return STH_OK;
}

}  // namespace

extern "C" int sth_leapfrog_noalloc(const sth_geom *g, void *u, void *v, const void *s,
                                    uint32_t steps, float c0, float c1, const float *c2,
                                    void *stream) {
  using namespace sth;
  using namespace sth::detail;
  if (!g || !u || !v || !s || u == v || u == s || v == s) return STH_ERR_ARG;
  if (g->dtype != STH_FP32 && g->dtype != STH_FP64) return STH_ERR_UNSUPPORTED;
  sth_layout_info L;
  int st = compute_layout(*g, L);
  if (st) return st;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return STH_ERR_NODEV;
  uint32_t *counters = nullptr;
  int cus = 0;
  if ((st = scratch_of(dev, &counters, &cus))) return st;
  if (steps == 0) return STH_OK;

  // A context on the stack over the caller's buffers: no allocation, no
  // fused passes (they need the twins of u and v), no timing.
  sth_ctx c;
  c.g = *g;
  c.L = L;
  c.R = (int)L.radius;
  c.eb = (int)L.elem_bytes;
  c.fmt = g->dtype == STH_FP64 ? FMT_F64 : FMT_F32;
  c.device = dev;
  c.stream = (hipStream_t)stream;
  c.base[0] = u;
  c.base[1] = v;
  c.base[2] = const_cast<void *>(s);
  c.ubuf = 0;
  c.no_fuse = true;
  c.tickets_dev = counters;
  c.cus = cus;
  c.tune.xcd_remap = 1;
  // the counters start at zero for this call's ticket bases
  hipError_t e = launch_zero_u32(counters, (int)kScratchWords, c.stream);
  if (e != hipSuccess) return hip_status(e);
  const Plan p = make_plan(&c, c.tune);
  if (!p.persist) return STH_ERR_STATE;
  return run_leap(&c, p, steps, c0, c1, c2);
}
