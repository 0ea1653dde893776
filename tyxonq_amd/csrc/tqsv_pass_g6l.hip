// Instantiates the lean tile-pass kernel for TB=6, MODE_GATE, IDX=uint32_t.
#include "tqsv_pass_impl.h"

namespace tqsv {
template hipError_t launch_one<6, MODE_GATE, uint32_t, true>(const PassLaunch& pl, hipStream_t st);
}
