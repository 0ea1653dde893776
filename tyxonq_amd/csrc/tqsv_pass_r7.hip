// Instantiates the tile-pass kernel for TB=7, MODE_RED, IDX=uint32_t.
#include "tqsv_pass_impl.h"

namespace tqsv {
template hipError_t launch_one<7, MODE_RED, uint32_t, false>(const PassLaunch& pl, hipStream_t st);
}
