// Instantiates every pass kernel for one axis length (WDM_INST_N = A*B) in
// FP32 and FP64; compiled once per entry of WDM_FOR_EACH_SIZE by build.py.
#include "wdm_ops.hpp"

#define WDM_DEFINE_SIZE_EXPAND(N, A, B) WDM_DEFINE_SIZE(N, A, B)

namespace wdm {
WDM_DEFINE_SIZE_EXPAND(WDM_INST_N, WDM_INST_A, WDM_INST_B)
}
