// kern_skinny.hip -- the short-side kernel's launcher table (skinny.hpp) and
// its FP32 instantiations: both transposes of A and B, tall and wide C, 1, 2
// or 4 column blocks of 16 for the short side, four or eight waves.  The
// split-BF16 modes are in kern_skinny_split.hip.
#include "skinny.hpp"

namespace ls {

template <int TA, int TB, int WIDE>
static void register_fp32(SkinnyTable &t)
{
    t.fn[P_FP32][TA][TB][WIDE][0][0] = &launch_skinny<P_FP32, TA, TB, WIDE, 1, 4>;
    t.fn[P_FP32][TA][TB][WIDE][1][0] = &launch_skinny<P_FP32, TA, TB, WIDE, 2, 4>;
    t.fn[P_FP32][TA][TB][WIDE][2][0] = &launch_skinny<P_FP32, TA, TB, WIDE, 4, 4>;
    t.fn[P_FP32][TA][TB][WIDE][0][1] = &launch_skinny<P_FP32, TA, TB, WIDE, 1, 8>;
    t.fn[P_FP32][TA][TB][WIDE][1][1] = &launch_skinny<P_FP32, TA, TB, WIDE, 2, 8>;
    t.fn[P_FP32][TA][TB][WIDE][2][1] = &launch_skinny<P_FP32, TA, TB, WIDE, 4, 8>;
}

void register_skinny_fp32(SkinnyTable &t)
{
    register_fp32<0, 0, 0>(t);
    register_fp32<0, 1, 0>(t);
    register_fp32<1, 0, 0>(t);
    register_fp32<1, 1, 0>(t);
    register_fp32<0, 0, 1>(t);
    register_fp32<0, 1, 1>(t);
    register_fp32<1, 0, 1>(t);
    register_fp32<1, 1, 1>(t);
}

SkinnyTable &skinny_table()
{
    static SkinnyTable t = [] {
        SkinnyTable s{};
        register_skinny_fp32(s);
        register_skinny_split(s);
        return s;
    }();
    return t;
}

}  // namespace ls
