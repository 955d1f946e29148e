// kern_bf16.hip -- tile-GEMM instantiations for the bf16 mode (one translation
// unit per precision so the kernels compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_bf16(LaunchTable &t)
{
    register_trans<P_BF16, 0, 0>(t);
    register_trans<P_BF16, 0, 1>(t);
    register_trans<P_BF16, 1, 0>(t);
    register_trans<P_BF16, 1, 1>(t);
}

}  // namespace ls
