// kern_fp64.hip -- tile-GEMM instantiations for the fp64 mode (one translation
// unit per precision so the kernels compile in parallel).
#include "kernels.hpp"

namespace ls {

void register_fp64(LaunchTable &t)
{
    register_trans<P_FP64, 0, 0>(t);
    register_trans<P_FP64, 0, 1>(t);
    register_trans<P_FP64, 1, 0>(t);
    register_trans<P_FP64, 1, 1>(t);
}

}  // namespace ls
