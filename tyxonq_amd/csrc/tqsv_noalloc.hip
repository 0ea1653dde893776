// tqsv_noalloc.hip - tqsv_dot_noalloc: the tqsv_dot reduction for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
//
// tqsv_dot keeps its block partial sums in a per-thread buffer that it
// allocates (hipMalloc) on first use and copies the result to the host with a
// synchronisation.  This variant uses the same two kernels (pauli_dot_kernel,
// final_sum_kernel) with the partials in `scratch`, a fixed array of this
// library's device code (one copy per GPU, held until the process ends),
// writes the result to a device buffer of the caller and does not
// synchronise.  It calls no allocating routine.
//
// Host code cannot use the array's name as a device address: the device
// address of each GPU's copy is looked up once per device, with that device
// current, on the first call there (hipGetSymbolAddress), and only that
// address is passed to kernels.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (different streams, or threads, without ordering between them) race.  Calls
// on one stream are ordered and safe.  Every partial that the final sum reads
// is written by the first kernel of the same call, so nothing has to be
// cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/tqsv.h"
#include "tqsv_internal.h"

namespace tqsv {

// two doubles (re, im) per block of the partial-sum kernel; reduce_grid(n)
// is at most 2048 blocks for every n (32 KiB per GPU)
constexpr size_t kScratchDoubles = 2 * 2048;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch_dev[kMaxDevices] = {};

// device address of the current device's copy of `scratch`
int scratch_address(double** out) {
  int dev = -1;
  TQSV_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(TQSV_EDEVICE, "tqsv_dot_noalloc: device index " + std::to_string(dev) +
                                  " beyond the supported " + std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "tqsv_dot_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) return fail(TQSV_EDEVICE, "tqsv_dot_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace
}  // namespace tqsv

extern "C" int tqsv_dot_noalloc(const void* u, const void* v, int n, uint64_t xmask,
                                uint64_t zmask, double* out_dev, void* stream) {
  using namespace tqsv;
  if (!u || !v || !out_dev) return fail(TQSV_EINVAL, "tqsv_dot_noalloc: NULL argument");
  if (n < 1 || n > 34) return fail(TQSV_EINVAL, "tqsv_dot_noalloc: n must be in [1, 34]");
  if ((xmask | zmask) >> n) return fail(TQSV_EINVAL, "tqsv_dot_noalloc: mask out of range");
  const uint32_t g = reduce_grid(n);
  if (2 * (size_t)g > kScratchDoubles)
    return fail(TQSV_EINVAL, "tqsv_dot_noalloc: " + std::to_string(2 * (size_t)g) +
                                 " partial sums exceed the fixed scratch of " +
                                 std::to_string(kScratchDoubles) + " doubles");
  double* d = nullptr;
  const int rc = scratch_address(&d);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  TQSV_HIP(k_pauli_dot(reinterpret_cast<const float2*>(u), reinterpret_cast<const float2*>(v), n,
                       xmask, zmask, d, st));
  TQSV_HIP(k_final_sum(d, g, 2, out_dev, st));
  return 0;
}
