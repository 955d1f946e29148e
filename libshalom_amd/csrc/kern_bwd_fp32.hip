// kern_bwd_fp32.hip -- fused backward launches for the fp32 mode (all four
// transpose combinations, 64 x 64 and 128 x 128 tiles).
#include "backward.hpp"

namespace ls {

void register_backward_fp32(BwdTable &t) { register_backward<P_FP32>(t); }

}  // namespace ls
