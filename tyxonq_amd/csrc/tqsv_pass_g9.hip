// Instantiates the tile-pass kernel for TB=9, MODE_GATE, IDX=uint32_t.
#include "tqsv_pass_impl.h"

namespace tqsv {
template hipError_t launch_one<9, MODE_GATE, uint32_t, false>(const PassLaunch& pl, hipStream_t st);
}
