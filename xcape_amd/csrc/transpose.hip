// Column-contiguous (ncol, nk) -> level-major (nk, ncol) copies of up to five
// profile arrays, staged through LDS in 64 x 64 tiles: a wave reads 64
// consecutive levels of one column (256 contiguous bytes) and writes 64
// consecutive columns of one level, so both directions are coalesced.  The
// tile rows are padded to 65 words so that the transposed LDS reads hit 64
// distinct banks.
//
// HBM traffic: 8 bytes per element (one read, one write).
#include <hip/hip_runtime.h>

#include "cape_launch.h"

namespace xcape {
namespace {

constexpr int TILE = 64;
constexpr int MAX_ARR = 5;

struct TransposeArgs {
    const float *src[MAX_ARR];
    float *dst[MAX_ARR];
    int64_t ncol, nk;
};

// blockIdx.x: column tile, blockIdx.y: level tile, blockIdx.z: array
__global__ void __launch_bounds__(256) transpose_levels_kernel(TransposeArgs a)
{
    __shared__ float tile[TILE][TILE + 1];
    const float *src = a.src[blockIdx.z];
    float *dst = a.dst[blockIdx.z];
    const int64_t c0 = (int64_t)blockIdx.x * TILE;
    const int64_t k0 = (int64_t)blockIdx.y * TILE;
    const int tx = threadIdx.x & (TILE - 1), ty = threadIdx.x >> 6;
    const int64_t k = k0 + tx;
#pragma unroll 4
    for (int r = ty; r < TILE; r += 4) {
        const int64_t c = c0 + r;
        if (c < a.ncol && k < a.nk) tile[r][tx] = src[c * a.nk + k];
    }
    __syncthreads();
    const int64_t c = c0 + tx;
#pragma unroll 4
    for (int r = ty; r < TILE; r += 4) {
        const int64_t kk = k0 + r;
        if (c < a.ncol && kk < a.nk) dst[kk * a.ncol + c] = tile[tx][r];
    }
}

}  // namespace

hipError_t launch_transpose_levels(const float *const *src, float *const *dst, int narr,
                                   int64_t ncol, int64_t nk, hipStream_t s)
{
    if (narr < 1 || narr > MAX_ARR || ncol < 0 || nk < 0) return hipErrorInvalidValue;
    if (ncol == 0 || nk == 0) return hipSuccess;
    const int64_t gx = (ncol + TILE - 1) / TILE, gy = (nk + TILE - 1) / TILE;
    if (gx > 0x7fffffff || gy > 65535) return hipErrorInvalidValue;
    TransposeArgs a{};
    for (int i = 0; i < narr; ++i) {
        a.src[i] = src[i];
        a.dst[i] = dst[i];
    }
    a.ncol = ncol;
    a.nk = nk;
    transpose_levels_kernel<<<dim3((unsigned)gx, (unsigned)gy, (unsigned)narr), 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace xcape
