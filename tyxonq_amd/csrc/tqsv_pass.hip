// tqsv_pass.hip - host dispatch of the tile-pass kernel variants.  The kernel
// templates are in tqsv_pass_impl.h; each tqsv_pass_v*.hip translation unit
// instantiates a few variants so they compile in parallel.
#include <hip/hip_runtime.h>

#include "tqsv_internal.h"
#include "tqsv_prog.h"

namespace tqsv {

template <int TB, bool DUAL, typename IDX>
hipError_t launch_one(const PassLaunch& pl, hipStream_t st);

size_t pass_lds_bytes(int tb, bool dual, uint32_t nred) {
  const size_t tile = ((size_t)1 << (tb + RB)) * sizeof(float2);
  const size_t nwave = ((size_t)1 << tb) >= 64 ? ((size_t)1 << tb) / 64 : 1;
  return (dual ? 2 : 1) * tile + nwave * nred * sizeof(double);
}

hipError_t launch_pass(int tb, bool dual, bool idx64, const PassLaunch& pl, hipStream_t st) {
  if (idx64) {
    if (dual) return hipErrorInvalidValue;
    switch (tb) {
      case 8: return launch_one<8, false, uint64_t>(pl, st);
      case 9: return launch_one<9, false, uint64_t>(pl, st);
      default: return hipErrorInvalidValue;
    }
  }
  if (dual) {
    switch (tb) {
      case 6: return launch_one<6, true, uint32_t>(pl, st);
      case 7: return launch_one<7, true, uint32_t>(pl, st);
      case 8: return launch_one<8, true, uint32_t>(pl, st);
      default: return hipErrorInvalidValue;
    }
  }
  switch (tb) {
    case 6: return launch_one<6, false, uint32_t>(pl, st);
    case 7: return launch_one<7, false, uint32_t>(pl, st);
    case 8: return launch_one<8, false, uint32_t>(pl, st);
    case 9: return launch_one<9, false, uint32_t>(pl, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tqsv
