// kern_fp32_ta0.hip -- tile-GEMM instantiations for the fp32 mode with
// transa = 0 (one translation unit per precision and transa so the kernels
// compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_fp32_ta0(LaunchTable &t)
{
    register_trans<P_FP32, 0, 0>(t);
    register_trans<P_FP32, 0, 1>(t);
}

}  // namespace ls
