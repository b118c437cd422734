// Allocation-free overlap, for callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops, harnesses that forbid it): the
// split-K partial sums of S = X^H Y go to a file-scope device array instead
// of a buffer the library allocates.  Its address is resolved once per
// device, with that device current, when a context is created
// (derek_hip_init -> noalloc_setup); kernels only ever receive that
// address, never the array's host-side name.
//
// The array holds kNoAllocScratchElems double2 (32 MiB) on every GPU the
// library runs on, until the process ends.  All calls on one GPU share it:
// calls that overlap in time on one GPU (different streams or threads) race.
// Every partial the reduction reads is written by the product kernel of the
// same call, so nothing has to be cleared between calls.
#include <mutex>

#include "context.hpp"

static __device__ __align__(256) double2 scratch[derek::kNoAllocScratchElems];

namespace derek {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch[kMaxDevices] = {};
}  // namespace

void noalloc_setup(int dev) {
  DEREK_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return;
  void* p = nullptr;
  DEREK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw StateError("hipGetSymbolAddress returned a null address for the scratch array");
  // the overlap kernels' resident-slot queries run here once, not inside
  // the allocation-free call
  (void)overlap_work_elems(1, 1, 1, 1);
  g_scratch[dev] = reinterpret_cast<double2*>(p);
}

void overlap_noalloc(derek_ctx* c, int64_t npw, int m, int n, const double2* X, const double2* Y, double2* S) {
  double2* work = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) work = g_scratch[c->device];
  }
  if (!work) throw StateError("scratch array of the context's device not resolved");
  const int64_t mn = (int64_t)m * n;
  DEREK_REQUIRE(mn <= kNoAllocScratchElems, "overlap_noalloc: m * n = " + std::to_string(mn) +
                                                " exceeds the scratch capacity of " +
                                                std::to_string(kNoAllocScratchElems) + " complex values");
  const int64_t max_chunks = kNoAllocScratchElems / mn;
  const bool herm = (X == Y) && (m == n);
  // launched directly: the profiler's event pool may create events
  overlap<double>(X, npw, 0, Y, npw, 0, m, n, npw, nullptr, S, m, 0, 1, work, kNoAllocScratchElems, herm, c->s, 0,
                  max_chunks);
}

}  // namespace derek
