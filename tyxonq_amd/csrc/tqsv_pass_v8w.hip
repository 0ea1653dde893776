// Instantiates the tile-pass kernel for TB=8, DUAL=false, IDX=uint64_t.
#include "tqsv_pass_impl.h"

namespace tqsv {
template hipError_t launch_one<8, false, uint64_t>(const PassLaunch& pl, hipStream_t st);
}
