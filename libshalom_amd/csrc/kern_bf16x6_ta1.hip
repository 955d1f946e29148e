// kern_bf16x6_ta1.hip -- tile-GEMM instantiations for the bf16x6 mode with
// transa = 1 (one translation unit per precision and transa so the kernels
// compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_bf16x6_ta1(LaunchTable &t)
{
    register_trans<P_BF16X6, 1, 0>(t);
    register_trans<P_BF16X6, 1, 1>(t);
}

}  // namespace ls
