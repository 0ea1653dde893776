// Instantiates the tile-pass kernel for TB=8, MODE_DUAL, IDX=uint32_t.
#include "tqsv_pass_impl.h"

namespace tqsv {
template hipError_t launch_one<8, MODE_DUAL, uint32_t, false>(const PassLaunch& pl, hipStream_t st);
}
