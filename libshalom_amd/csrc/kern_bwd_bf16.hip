// kern_bwd_bf16.hip -- fused backward launches for the bf16 mode (all four
// transpose combinations, 64 x 64 and 128 x 128 tiles).
#include "backward.hpp"

namespace ls {

void register_backward_bf16(BwdTable &t) { register_backward<P_BF16>(t); }

}  // namespace ls
