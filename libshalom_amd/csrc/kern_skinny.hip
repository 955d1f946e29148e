// kern_skinny.hip -- instantiations of the short-side FP32 GEMM kernel
// (skinny.hpp): both transposes of A and B, tall and wide C, 1, 2 or 4
// column blocks of 16 for the short side, four or eight waves.
#include "skinny.hpp"

namespace ls {

template <int TA, int TB, int WIDE>
static void register_skinny(SkinnyTable &t)
{
    t.fn[TA][TB][WIDE][0][0] = &launch_skinny<TA, TB, WIDE, 1, 4>;
    t.fn[TA][TB][WIDE][1][0] = &launch_skinny<TA, TB, WIDE, 2, 4>;
    t.fn[TA][TB][WIDE][2][0] = &launch_skinny<TA, TB, WIDE, 4, 4>;
    t.fn[TA][TB][WIDE][0][1] = &launch_skinny<TA, TB, WIDE, 1, 8>;
    t.fn[TA][TB][WIDE][1][1] = &launch_skinny<TA, TB, WIDE, 2, 8>;
    t.fn[TA][TB][WIDE][2][1] = &launch_skinny<TA, TB, WIDE, 4, 8>;
}

SkinnyTable &skinny_table()
{
    static SkinnyTable t = [] {
        SkinnyTable s{};
        register_skinny<0, 0, 0>(s);
        register_skinny<0, 1, 0>(s);
        register_skinny<1, 0, 0>(s);
        register_skinny<1, 1, 0>(s);
        register_skinny<0, 0, 1>(s);
        register_skinny<0, 1, 1>(s);
        register_skinny<1, 0, 1>(s);
        register_skinny<1, 1, 1>(s);
        return s;
    }();
    return t;
}

}  // namespace ls
