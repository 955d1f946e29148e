// kern_bf16x3.hip -- tile-GEMM instantiations for the bf16x3 mode (one translation
// unit per precision so the kernels compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_bf16x3(LaunchTable &t)
{
    register_trans<P_BF16X3, 0, 0>(t);
    register_trans<P_BF16X3, 0, 1>(t);
    register_trans<P_BF16X3, 1, 0>(t);
    register_trans<P_BF16X3, 1, 1>(t);
}

}  // namespace ls
