// kern_fp32.hip -- tile-GEMM instantiations for the fp32 mode (one translation
// unit per precision so the kernels compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_fp32(LaunchTable &t)
{
    register_trans<P_FP32, 0, 0>(t);
    register_trans<P_FP32, 0, 1>(t);
    register_trans<P_FP32, 1, 0>(t);
    register_trans<P_FP32, 1, 1>(t);
}

}  // namespace ls
