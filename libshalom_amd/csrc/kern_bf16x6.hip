// kern_bf16x6.hip -- tile-GEMM instantiations for the bf16x6 mode (one translation
// unit per precision so the kernels compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_bf16x6(LaunchTable &t)
{
    register_trans<P_BF16X6, 0, 0>(t);
    register_trans<P_BF16X6, 0, 1>(t);
    register_trans<P_BF16X6, 1, 0>(t);
    register_trans<P_BF16X6, 1, 1>(t);
}

}  // namespace ls
