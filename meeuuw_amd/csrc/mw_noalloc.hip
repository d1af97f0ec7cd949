// Allocation-free variant of the particle sort, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight stepping loops).
//
// The sort's scratch (the per-element counters and the unstable scatter
// buffer) lives in one file-scope device array per GPU, of a fixed size, held
// until the process ends.  Host code never uses the array's name as a device
// address: its address on each device is looked up once, with that device
// current, and kept per device.  Calls on one GPU share that array, so calls
// that overlap on one GPU (different streams or threads) race.

#include <mutex>

#include "mw_common.h"
#include "mw_capi_impl.h"
#include "meeuuw_hip.h"

namespace {

constexpr long kMaxParticles = 1L << 22;  // 4 194 304
constexpr long kMaxElements = 1L << 20;   // 1 048 576
// counters: kMaxElements + 1 ints; prefix-sum tile sums: one per 4096
// elements, plus one; scatter buffer: kMaxParticles ints (each rounded up to 64)
constexpr long kCntInts = ((kMaxElements + 1 + 63) / 64) * 64;
constexpr long kTsumInts = ((kMaxElements / 4096 + 2 + 63) / 64) * 64;
constexpr long kScratchInts = kCntInts + kTsumInts + kMaxParticles;

__device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int* g_ptr[kMaxDevices] = {nullptr};

// The array's device address on the current device (looked up on the first
// call on that device).
int scratch_on_current_device(int** out) {
  int dev = -1;
  MW_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return mw::fail(MW_EINVAL, "sort (no-alloc): device index");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    MW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return mw::fail(MW_EHIP, "sort (no-alloc): no device address for the scratch array");
    g_ptr[dev] = static_cast<int*>(p);
  }
  *out = g_ptr[dev];
  return MW_OK;
}

}  // namespace

extern "C" {

int mw_sort_particles_by_element_noalloc(int np, int nel, const uint8_t* active,
                                         const int32_t* iel, int32_t* order, int32_t* seg,
                                         void* stream) {
  using namespace mw;
  if (np <= 0 || nel <= 0 || !active || !iel || !order || !seg)
    return fail(MW_EINVAL, "sort_particles_by_element_noalloc: bad arguments");
  if ((long)np > kMaxParticles || (long)nel > kMaxElements)
    return fail(MW_ERANGE,
                "sort_particles_by_element_noalloc: at most 4194304 particles and 1048576 "
                "elements (the size of its fixed scratch array)");
  int* base = nullptr;
  RC(scratch_on_current_device(&base));
  // counters are cleared by the sort itself (cnt_zero = false): nothing is
  // assumed about what the previous call left in the array
  return sort_by_element_scratch(np, nel, active, iel, order, seg, base, base + kCntInts,
                                 base + kCntInts + kTsumInts, false, (hipStream_t)stream);
}

}  // extern "C"
