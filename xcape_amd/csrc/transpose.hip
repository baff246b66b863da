// Column-contiguous (ncol, nk) -> level-major (nk, ncol) copies of up to five
// profile arrays, staged through LDS.
//
// Slab kernel (nk <= 255): a workgroup copies 64 whole columns of one array.
// Their nk * 64 values are one contiguous run of the source, read with
// consecutive lanes on consecutive words, into LDS as 64 rows of an odd
// pitch; each level is then written as 64 consecutive columns (256 bytes).
// Every load and store is coalesced and no thread idles on a partial tile
// (137 levels in 64-level tiles leave 55 of every 192 tile rows empty).
//
// Tile kernel (any nk): 64 x 64 tiles, a wave reading 64 consecutive levels
// of one column and writing 64 consecutive columns of one level; the tile
// rows are padded to 65 words so that the transposed LDS reads hit distinct
// banks.
//
// HBM traffic: 8 bytes per element (one read, one write).  The slab kernel's
// loads and stores are non-temporal: each element passes once, and the
// level-major copy, larger than the caches, is read back from HBM by the next
// kernels either way (the helicity-only call ran 0.1 ms faster so).
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
    int64_t ld;   // the destination's row length (>= ncol)
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
        if (c < a.ncol && kk < a.nk) dst[kk * a.ld + c] = tile[tx][r];
    }
}

constexpr int SLAB = 64;          // columns per workgroup of the slab kernel
constexpr int SLAB_MAX_PITCH = 255;

// blockIdx.x: 64-column slab, blockIdx.y: array; LDS: SLAB * pitch floats.
// VEC4 (16-byte aligned sources, nk odd, pitch == nk): the run goes into LDS
// as it lies, 16 bytes per lane on both sides (column c's level k at
// c * nk + k; the run of slab j starts 256 nk bytes into the array), and the
// odd nk keeps the column-wise reads below on distinct banks.
template <bool VEC4>
__global__ void __launch_bounds__(256) transpose_slab_kernel(TransposeArgs a, int pitch)
{
    extern __shared__ float slab[];
    const float *src = a.src[blockIdx.y];
    float *dst = a.dst[blockIdx.y];
    const int64_t c0 = (int64_t)blockIdx.x * SLAB;
    const int nk = (int)a.nk;
    const int nc = (int)(a.ncol - c0 < SLAB ? a.ncol - c0 : SLAB);
    const int total = nc * nk;
    const float *s = src + c0 * a.nk;
    if constexpr (VEC4) {
        const int total4 = total >> 2;
        typedef float v4f __attribute__((ext_vector_type(4)));
        const v4f *s4 = reinterpret_cast<const v4f *>(s);
        v4f *l4 = reinterpret_cast<v4f *>(slab);
#pragma unroll 4
        for (int q = threadIdx.x; q < total4; q += 256) l4[q] = __builtin_nontemporal_load(&s4[q]);
        for (int e = (total4 << 2) + (int)threadIdx.x; e < total; e += 256) slab[e] = s[e];
    } else {
        // element e of the run is level k of column c: walk (c, k) along with e
        int c = threadIdx.x / nk, k = threadIdx.x - c * nk;
        const int dc = 256 / nk, dk = 256 - dc * nk;
#pragma unroll 4
        for (int e = threadIdx.x; e < total; e += 256) {
            slab[c * pitch + k] = s[e];
            k += dk;
            c += dc;
            if (k >= nk) {
                k -= nk;
                ++c;
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < nc) {
        float *d = dst + c0 + lane;
#pragma unroll 4
        for (int kk = w; kk < nk; kk += 4)
            __builtin_nontemporal_store(slab[lane * pitch + kk], &d[(int64_t)kk * a.ld]);
    }
}

}  // namespace

hipError_t launch_transpose_levels(const float *const *src, float *const *dst, int narr,
                                   int64_t ncol, int64_t nk, hipStream_t s, int64_t ld)
{
    if (ld < 0) ld = ncol;
    if (narr < 1 || narr > MAX_ARR || ncol < 0 || nk < 0 || ld < ncol) return hipErrorInvalidValue;
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
    a.ld = ld;
    const int64_t pitch = nk | 1;   // odd: the column-wise LDS reads hit distinct banks
    if (pitch <= SLAB_MAX_PITCH) {
        const int64_t sx = (ncol + SLAB - 1) / SLAB;
        if (sx > 0x7fffffff) return hipErrorInvalidValue;
        bool vec4 = (nk & 1) != 0;
        for (int i = 0; i < narr; ++i) vec4 = vec4 && (reinterpret_cast<uintptr_t>(src[i]) % 16 == 0);
        const dim3 grid((unsigned)sx, (unsigned)narr);
        const size_t lds = sizeof(float) * SLAB * (size_t)pitch;
        if (vec4)
            transpose_slab_kernel<true><<<grid, 256, lds, s>>>(a, (int)pitch);
        else
            transpose_slab_kernel<false><<<grid, 256, lds, s>>>(a, (int)pitch);
    } else {
        transpose_levels_kernel<<<dim3((unsigned)gx, (unsigned)gy, (unsigned)narr), 256, 0, s>>>(
            a);
    }
    return hipGetLastError();
}

}  // namespace xcape
