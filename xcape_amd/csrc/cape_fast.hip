// MATH_FAST grid kernels: hardware exp2/log2/reciprocal, FMA contraction.
#define XC_MATH MATH_FAST
#define XC_NS fast
#include "cape_grid_impl.h"

namespace xcape {
hipError_t launch_cape_grid_fast(const CapeGridArgs &a, int kind, int block, hipStream_t s)
{
    return fast::launch_t<float>(a, kind, block, s);
}
}  // namespace xcape
