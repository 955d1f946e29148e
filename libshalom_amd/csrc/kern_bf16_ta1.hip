// kern_bf16_ta1.hip -- tile-GEMM instantiations for the bf16 mode with
// transa = 1 (one translation unit per precision and transa so the kernels
// compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_bf16_ta1(LaunchTable &t)
{
    register_trans<P_BF16, 1, 0>(t);
    register_trans<P_BF16, 1, 1>(t);
}

}  // namespace ls
