// kern_skinny_split.hip -- the short-side kernel (skinny.hpp) in the split
// BF16 modes (BF16X3, BF16, BF16X6): both transposes, tall and wide C, 1 or
// 2 column blocks of 16 (the planner keeps wider short sides of these modes
// on the tile kernels), four waves.
#include "skinny.hpp"

namespace ls {

template <int PREC, int TA, int TB, int WIDE>
static void register_split(SkinnyTable &t)
{
    t.fn[PREC][TA][TB][WIDE][0][0] = &launch_skinny<PREC, TA, TB, WIDE, 1, 4>;
    t.fn[PREC][TA][TB][WIDE][1][0] = &launch_skinny<PREC, TA, TB, WIDE, 2, 4>;
}

template <int PREC>
static void register_prec(SkinnyTable &t)
{
    register_split<PREC, 0, 0, 0>(t);
    register_split<PREC, 0, 1, 0>(t);
    register_split<PREC, 1, 0, 0>(t);
    register_split<PREC, 1, 1, 0>(t);
    register_split<PREC, 0, 0, 1>(t);
    register_split<PREC, 0, 1, 1>(t);
    register_split<PREC, 1, 0, 1>(t);
    register_split<PREC, 1, 1, 1>(t);
}

void register_skinny_split(SkinnyTable &t)
{
    register_prec<P_BF16X3>(t);
    register_prec<P_BF16>(t);
    register_prec<P_BF16X6>(t);
}

}  // namespace ls
