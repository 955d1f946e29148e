// kern_fp32_ta1.hip -- tile-GEMM instantiations for the fp32 mode with
// transa = 1 (one translation unit per precision and transa so the kernels
// compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_fp32_ta1(LaunchTable &t)
{
    register_trans<P_FP32, 1, 0>(t);
    register_trans<P_FP32, 1, 1>(t);
}

}  // namespace ls
