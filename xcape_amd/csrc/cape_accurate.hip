// MATH_ACCURATE grid kernels (FP32 and the FP64 yardstick).  Built with
// -ffp-contract=off: every +,-,*,/ rounds as the reference's does.
#define XC_MATH MATH_ACCURATE
#define XC_NS acc
#include "cape_grid_impl.h"

namespace xcape {
hipError_t launch_cape_grid_accurate(const CapeGridArgs &a, bool fp64, int kind, int block,
                                     hipStream_t s)
{
    if (fp64) {
        if (kind != GRID_CAPE && kind != GRID_SRH) return hipErrorInvalidValue;
        return acc::launch_t<double>(a, kind, block, s);
    }
    return acc::launch_t<float>(a, kind, block, s);
}
}  // namespace xcape
