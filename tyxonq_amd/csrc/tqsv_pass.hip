// tqsv_pass.hip - host dispatch of the tile-pass kernel variants.  The kernel
// templates are in tqsv_pass_impl.h; each tqsv_pass_*.hip translation unit
// instantiates one variant so they compile in parallel.
#include <hip/hip_runtime.h>

#include "tqsv_internal.h"
#include "tqsv_prog.h"

namespace tqsv {

template <int TB, int MODE, typename IDX, bool LEAN>
hipError_t launch_one(const PassLaunch& pl, hipStream_t st);

size_t pass_lds_bytes(int tb, int mode, uint32_t nred) {
  const size_t tile = ((size_t)1 << (tb + RB)) * sizeof(float2);
  const size_t nwave = ((size_t)1 << tb) >= 64 ? ((size_t)1 << tb) / 64 : 1;
  return (mode == MODE_DUAL ? 2 : 1) * tile + (mode == MODE_GATE ? 0 : nwave * nred * sizeof(double));
}

bool pass_variant_exists(int tb, int mode, bool idx64) {
  if (idx64) return (mode == MODE_GATE || mode == MODE_RED) && tb == 8;
  switch (mode) {
    case MODE_GATE: return tb >= 6 && tb <= 9;
    case MODE_RED: return tb >= 6 && tb <= 8;
    case MODE_DUAL: return tb >= 6 && tb <= 8;
    default: return false;
  }
}

hipError_t launch_pass(int tb, int mode, bool idx64, bool lean, const PassLaunch& pl,
                       hipStream_t st) {
  if (lean && !idx64) {
    if (mode == MODE_GATE) switch (tb) {
        case 6: return launch_one<6, MODE_GATE, uint32_t, true>(pl, st);
        case 7: return launch_one<7, MODE_GATE, uint32_t, true>(pl, st);
        case 8: return launch_one<8, MODE_GATE, uint32_t, true>(pl, st);
        case 9: return launch_one<9, MODE_GATE, uint32_t, true>(pl, st);
        default: break;
      }
    if (mode == MODE_DUAL) switch (tb) {
        case 6: return launch_one<6, MODE_DUAL, uint32_t, true>(pl, st);
        case 7: return launch_one<7, MODE_DUAL, uint32_t, true>(pl, st);
        case 8: return launch_one<8, MODE_DUAL, uint32_t, true>(pl, st);
        default: break;
      }
  }
  if (idx64) {
    if (mode == MODE_GATE && tb == 8) return launch_one<8, MODE_GATE, uint64_t, false>(pl, st);
    if (mode == MODE_RED && tb == 8) return launch_one<8, MODE_RED, uint64_t, false>(pl, st);
    return hipErrorInvalidValue;
  }
  switch (mode) {
    case MODE_GATE:
      switch (tb) {
        case 6: return launch_one<6, MODE_GATE, uint32_t, false>(pl, st);
        case 7: return launch_one<7, MODE_GATE, uint32_t, false>(pl, st);
        case 8: return launch_one<8, MODE_GATE, uint32_t, false>(pl, st);
        case 9: return launch_one<9, MODE_GATE, uint32_t, false>(pl, st);
        default: return hipErrorInvalidValue;
      }
    case MODE_RED:
      switch (tb) {
        case 6: return launch_one<6, MODE_RED, uint32_t, false>(pl, st);
        case 7: return launch_one<7, MODE_RED, uint32_t, false>(pl, st);
        case 8: return launch_one<8, MODE_RED, uint32_t, false>(pl, st);
        default: return hipErrorInvalidValue;
      }
    case MODE_DUAL:
      switch (tb) {
        case 6: return launch_one<6, MODE_DUAL, uint32_t, false>(pl, st);
        case 7: return launch_one<7, MODE_DUAL, uint32_t, false>(pl, st);
        case 8: return launch_one<8, MODE_DUAL, uint32_t, false>(pl, st);
        default: return hipErrorInvalidValue;
      }
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tqsv
