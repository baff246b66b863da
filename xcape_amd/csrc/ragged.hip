// Ragged batches: columns with different level counts, parcel sources and
// adiabats in one graph-replayed sequence of persistent launches.
//
// Caller layout (CSR): column i owns levels offsets[i] .. offsets[i+1]-1 of
// the concatenated p, t, td arrays (hPa, degC); surface values, source and
// adiabat are per column.
//
// Plan (built once on the host): the columns are sorted, stably, by adiabat,
// then source, then descending level count, so that the 64 lanes of a
// wavefront run the same code path on columns of similar length and long
// columns are dealt first.  Each adiabat is a segment of that order with its
// own kernel instantiation (one ICE/PSEUDO code path: a kernel that switched
// over all four per column would carry all four in its registers).
//
// Run (one captured hipGraph, replayed):
//  1. reset the segments' work counters (a one-thread kernel: no memset node);
//  2. stage the batch level-major in sorted order: level k of sorted column j
//     at [k * ncol + j] of a padded (max level count, ncol) copy, through
//     64 x 64 LDS tiles so that the CSR reads (consecutive levels of one
//     column) and the level-major writes (consecutive columns of one level)
//     are both coalesced; the surface values are gathered in the same order.
//     A wavefront's 64 lanes then read 64 consecutive words per level (the
//     CSR layout read one column per lane makes every lane fetch its own
//     cache lines, which the L2 cannot keep for large batches);
//  3. per segment, the ascent, the segments' launches as parallel branches
//     of the graph (forked after the staging, joined before the graph ends):
//     one segment alone is often too small to fill the device, and a
//     segment's tail then overlaps the others' work.
//     * The reference's method: a persistent ascent kernel, a grid sized to
//       fill the device once, every wavefront taking the next `chunk`
//       columns of its segment from the counter until the segment is drained
//       (every wave reaches the exit test, so the grid always drains).
//       Optionally the first `tile` levels of a wave's columns are copied to
//       LDS first (the autotuner decides whether that pays).
//     * The warm start: the segment is a grid of its own to the grid path's
//       warm kernels (setup pass, surrogate, persistent per-lane ascent),
//       with each column's level count, source and output slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <type_traits>
#include <vector>

#include "../../include/xcape_hip.h"
#include "capi_common.h"
#include "cape_column.h"
#include "cape_launch.h"

namespace xcape {
namespace {

constexpr int MAX_TILE = 64;   // upper bound of the LDS level-tile depth
constexpr int NSEG = 4;        // one segment per adiabat
constexpr int STAGE_TILE = 64;

struct StageArgs {
    const int64_t *offsets;
    const int32_t *order, *nk_sorted;
    const float *p, *t, *td, *ps, *ts, *tds;
    float *lp, *lt, *ltd;          // (maxnk, ncol) level-major, sorted order
    float *sps, *sts, *stds;       // (ncol,) sorted order
    int64_t ncol, maxnk;
};

// blockIdx.x: 64 sorted columns, blockIdx.y: 64 levels
__global__ void __launch_bounds__(256) stage_kernel(StageArgs a)
{
    __shared__ float tile[STAGE_TILE][STAGE_TILE + 1];
    const int64_t j0 = (int64_t)blockIdx.x * STAGE_TILE;
    const int64_t k0 = (int64_t)blockIdx.y * STAGE_TILE;
    const int tx = threadIdx.x & (STAGE_TILE - 1), ty = threadIdx.x >> 6;
    if (blockIdx.y == 0 && threadIdx.x < STAGE_TILE) {
        const int64_t j = j0 + tx;
        if (j < a.ncol) {
            const int32_t c = a.order[j];
            a.sps[j] = a.ps[c];
            a.sts[j] = a.ts[c];
            a.stds[j] = a.tds[c];
        }
    }
    for (int arr = 0; arr < 3; ++arr) {
        const float *src = arr == 0 ? a.p : (arr == 1 ? a.t : a.td);
        float *dst = arr == 0 ? a.lp : (arr == 1 ? a.lt : a.ltd);
        for (int r = ty; r < STAGE_TILE; r += 4) {
            const int64_t j = j0 + r, k = k0 + tx;
            if (j < a.ncol && k < a.nk_sorted[j]) tile[r][tx] = src[a.offsets[a.order[j]] + k];
        }
        __syncthreads();
        for (int r = ty; r < STAGE_TILE; r += 4) {
            const int64_t k = k0 + r, j = j0 + tx;
            // entries past a column's level count are never read
            if (j < a.ncol && k < a.maxnk) dst[k * a.ncol + j] = tile[tx][r];
        }
        __syncthreads();
    }
}

struct RaggedArgs {
    const int32_t *order, *nk_sorted, *src_sorted;
    const float *lp, *lt, *ltd, *sps, *sts, *stds;
    int64_t ncol;                  // the level stride of the staged copies
    int64_t seg_begin, seg_end;    // this launch's segment of the sorted order
    float pinc, ml_depth;
    float *cape, *cin, *zout;
    int32_t *mulvl;
    unsigned int *counter;         // this segment's work counter
    unsigned long long *stats;
    int chunk;                     // columns a wavefront takes per grab
    int tile;                      // levels staged in LDS (0 = none)
    int cpw;                       // columns per wave pass: 64, or 16 / 32 (other lanes idle)
};

// Sorted column j of the staged batch: level k >= 1 at [(k-1) * ls + j] of
// the level-major copies.
struct StagedCol {
    const float *p, *t, *td;
    uint32_t j, ls;
    float ps, ts, tds;
    __device__ __forceinline__ uint32_t at(int k) const { return (uint32_t)(k - 1) * ls + j; }
    __device__ __forceinline__ float P(int k) const { return k == 0 ? ps : p[at(k)]; }
    __device__ __forceinline__ float Tc(int k) const { return k == 0 ? ts : t[at(k)]; }
    __device__ __forceinline__ float Tdc(int k) const { return k == 0 ? tds : td[at(k)]; }
    __device__ __forceinline__ Raw raw_above(int k) const
    {
        return Raw{p[at(k)], t[at(k)], td[at(k)]};
    }
};

// The same with the levels below `tile` read from this lane's LDS slice.
struct TiledCol {
    const float *p, *t, *td;
    const float *lp, *lt, *ltd;
    uint32_t j, ls;
    int tile;
    float ps, ts, tds;
    __device__ __forceinline__ uint32_t at(int k) const { return (uint32_t)(k - 1) * ls + j; }
    __device__ __forceinline__ float P(int k) const
    {
        return k == 0 ? ps : (k - 1 < tile ? lp[(k - 1) * 64] : p[at(k)]);
    }
    __device__ __forceinline__ float Tc(int k) const
    {
        return k == 0 ? ts : (k - 1 < tile ? lt[(k - 1) * 64] : t[at(k)]);
    }
    __device__ __forceinline__ float Tdc(int k) const
    {
        return k == 0 ? tds : (k - 1 < tile ? ltd[(k - 1) * 64] : td[at(k)]);
    }
    __device__ __forceinline__ Raw raw_above(int k) const
    {
        return k - 1 < tile ? Raw{lp[(k - 1) * 64], lt[(k - 1) * 64], ltd[(k - 1) * 64]}
                            : Raw{p[at(k)], t[at(k)], td[at(k)]};
    }
};

__global__ void reset_counters(unsigned int *c)
{
    if (threadIdx.x < NSEG) c[threadIdx.x] = 0u;
}

// The reference's loops (cape_column).  TILED: stage the first a.tile levels
// in LDS.  STATS: accumulate the work counters.
template <int MATH, bool ICE, bool PSEUDO, bool TILED, bool STATS>
__global__ void __launch_bounds__(256) ragged_kernel(RaggedArgs a)
{
    // LDS: per wave, 3 arrays x tile levels x 64 lanes, level-major so the
    // wave's 64 lanes read 64 consecutive words (distinct banks)
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int wtile = a.tile > 0 ? a.tile : 1;
    float *lp = lds + (size_t)wid * 3 * wtile * 64 + lane;
    float *lt = lp + wtile * 64, *ltd = lt + wtile * 64;
    LaneStats st;
    unsigned long long ran = 0;
    for (;;) {
        unsigned int base = 0;
        if (lane == 0) base = atomicAdd(a.counter, (unsigned int)a.chunk);
        base = __shfl(base, 0, 64);
        const int64_t b0 = a.seg_begin + (int64_t)base;
        if (b0 >= a.seg_end) break;
        for (int sub = 0; sub < a.chunk; sub += a.cpw) {
            const int64_t j = b0 + sub + lane;
            const bool active = lane < a.cpw && j < a.seg_end;
            const int nk_in = active ? a.nk_sorted[j] : 0;
            const int tile = TILED ? min(a.tile, nk_in) : 0;
            // levels 0..tile-1: the wave's lanes read consecutive columns
            if constexpr (TILED) {
                for (int k = 0; k < a.tile; ++k) {
                    if (k < tile) {
                        const int64_t g = (int64_t)k * a.ncol + j;
                        lp[k * 64] = a.lp[g];
                        lt[k * 64] = a.lt[g];
                        ltd[k * 64] = a.ltd[g];
                    }
                }
            }
            // each lane reads back only its own LDS slots: no barrier needed
            if (!active) continue;
            const int32_t col = a.order[j];
            const float ts = a.sts[j];
            if (!(ts > 0.0f)) {
                a.cape[col] = 0.0f;
                a.cin[col] = 0.0f;
                a.zout[col] = 0.0f;
                a.mulvl[col] = 0;
                continue;
            }
            ++ran;
            using Col = std::conditional_t<TILED, TiledCol, StagedCol>;
            Col c;
            if constexpr (TILED)
                c = TiledCol{a.lp, a.lt, a.ltd, lp, lt, ltd, (uint32_t)j, (uint32_t)a.ncol,
                             tile, a.sps[j], ts, a.stds[j]};
            else
                c = StagedCol{a.lp, a.lt, a.ltd, (uint32_t)j, (uint32_t)a.ncol, a.sps[j], ts,
                              a.stds[j]};
            ColumnOut<float> o{};
            const int nk = nk_in + 1;
            const int src = a.src_sorted[j];
            cape_column<float, MATH, ICE, PSEUDO>(c, nk, a.pinc, src, a.ml_depth, o, st);
            a.mulvl[col] = o.mulvl;
            a.cape[col] = o.cape;
            a.cin[col] = o.cin;
            a.zout[col] = o.zout;
        }
    }
    if constexpr (STATS) {
        unsigned long long v[7] = {ran, st.levels, st.substeps, st.iterations, st.nonconverged,
                                   st.fallbacks, st.first};
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            unsigned long long x = v[c];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
            if (lane == 0 && x) atomicAdd(&a.stats[c], x);
        }
    }
}

using Kernel = void (*)(RaggedArgs);

template <int MATH, bool TILED, bool STATS>
Kernel pick_adiabat(int adiabat)
{
    switch (adiabat) {
    case 1: return ragged_kernel<MATH, false, true, TILED, STATS>;
    case 2: return ragged_kernel<MATH, false, false, TILED, STATS>;
    case 3: return ragged_kernel<MATH, true, true, TILED, STATS>;
    default: return ragged_kernel<MATH, true, false, TILED, STATS>;
    }
}

template <int MATH>
Kernel pick_math(int adiabat, bool tiled, bool stats)
{
    if (tiled) return stats ? pick_adiabat<MATH, true, true>(adiabat)
                            : pick_adiabat<MATH, true, false>(adiabat);
    return stats ? pick_adiabat<MATH, false, true>(adiabat)
                 : pick_adiabat<MATH, false, false>(adiabat);
}

// the plain method's kernel for one segment
Kernel pick(bool fast, int adiabat, bool tiled, bool stats)
{
    return fast ? pick_math<MATH_FAST>(adiabat, tiled, stats)
                : pick_math<MATH_ACCURATE>(adiabat, tiled, stats);
}

}  // namespace
}  // namespace xcape

using namespace xcape;

struct xcape_ragged_plan {
    StageArgs stage{};
    RaggedArgs args{};                 // common part of every segment's launch
    CapeGridArgs warm_args[NSEG] = {}; // warm plans: each segment as a grid
    float *d_warm = nullptr;           // warm plans: the segments' scratch
    unsigned int *d_work = nullptr;    // warm plans: 8 counters per segment
    int64_t seg_begin[NSEG] = {}, seg_end[NSEG] = {};
    int seg_grid[NSEG] = {};
    int32_t *d_order = nullptr, *d_nk = nullptr, *d_src = nullptr;
    float *d_stage = nullptr;          // lp, lt, ltd, sps, sts, stds
    unsigned int *d_counter = nullptr;
    unsigned long long *d_stats = nullptr;
    int block = 256;
    size_t lds_bytes = 0;
    bool fast = false, warm = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// seg_streams: one stream per segment to fork the segments' launches onto
// (after the staging on s, joined back to s), or nullptr to queue them all
// on s.
hipError_t enqueue(const xcape_ragged_plan *pl, bool stats, hipStream_t s,
                   const hipStream_t *seg_streams = nullptr, hipEvent_t fork = nullptr,
                   const hipEvent_t *join = nullptr)
{
    reset_counters<<<1, 64, 0, s>>>(pl->d_counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const StageArgs &sa = pl->stage;
    const dim3 sgrid((unsigned)((sa.ncol + STAGE_TILE - 1) / STAGE_TILE),
                     (unsigned)((sa.maxnk + STAGE_TILE - 1) / STAGE_TILE));
    stage_kernel<<<sgrid, 256, 0, s>>>(sa);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (seg_streams && (e = hipEventRecord(fork, s)) != hipSuccess) return e;
    for (int g = 0; g < NSEG; ++g) {
        if (pl->seg_end[g] <= pl->seg_begin[g]) continue;
        hipStream_t sg = s;
        if (seg_streams) {
            sg = seg_streams[g];
            if ((e = hipStreamWaitEvent(sg, fork, 0)) != hipSuccess) return e;
        }
        if (pl->warm) {
            CapeGridArgs w = pl->warm_args[g];
            w.stats = stats ? pl->d_stats : nullptr;
            e = pl->fast ? launch_cape_grid_fast(w, GRID_CAPE, pl->block, sg)
                         : launch_cape_grid_accurate(w, false, GRID_CAPE, pl->block, sg);
        } else {
            RaggedArgs a = pl->args;
            a.seg_begin = pl->seg_begin[g];
            a.seg_end = pl->seg_end[g];
            a.counter = pl->d_counter + g;
            a.stats = stats ? pl->d_stats : nullptr;
            const Kernel k = pick(pl->fast, g + 1, pl->args.tile > 0, stats);
            void *kargs[] = {&a};
            e = hipLaunchKernel(reinterpret_cast<const void *>(k), dim3(pl->seg_grid[g]),
                                dim3(pl->block), kargs, pl->lds_bytes, sg);
        }
        if (e != hipSuccess) return e;
        if (seg_streams && ((e = hipEventRecord(join[g], sg)) != hipSuccess ||
                            (e = hipStreamWaitEvent(s, join[g], 0)) != hipSuccess))
            return e;
    }
    return hipSuccess;
}

void destroy(xcape_ragged_plan *pl)
{
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->d_order) (void)hipFree(pl->d_order);
    if (pl->d_nk) (void)hipFree(pl->d_nk);
    if (pl->d_src) (void)hipFree(pl->d_src);
    if (pl->d_stage) (void)hipFree(pl->d_stage);
    if (pl->d_counter) (void)hipFree(pl->d_counter);
    if (pl->d_stats) (void)hipFree(pl->d_stats);
    if (pl->d_warm) (void)hipFree(pl->d_warm);
    if (pl->d_work) (void)hipFree(pl->d_work);
    delete pl;
}

}  // namespace

extern "C" {

int xcape_ragged_plan_create(const int64_t *offsets_host, const int32_t *source_host,
                             const int32_t *adiabat_host, int64_t ncol,
                             const int64_t *offsets, const int32_t *source,
                             const int32_t *adiabat, const float *p, const float *t,
                             const float *td, const float *ps, const float *ts,
                             const float *tds, float pinc, float ml_depth, float *cape,
                             float *cin, int32_t *mulev, float *zmulev, int chunk, int tile,
                             int cols_per_block, uint32_t flags, xcape_ragged_plan **out)
{
    if (!out) return fail(XCAPE_EINVAL, "out is NULL");
    *out = nullptr;
    if (ncol < 1 || ncol > INT32_MAX - 256)
        return fail(XCAPE_EINVAL, "ncol=%lld out of range", (long long)ncol);
    if (!offsets_host || !source_host || !adiabat_host || !offsets || !source || !adiabat ||
        !p || !t || !td || !ps || !ts || !tds || !cape || !cin || !mulev || !zmulev)
        return fail(XCAPE_EINVAL, "NULL argument");
    if (!(pinc >= 1.0f) || !std::isfinite(pinc))
        return fail(XCAPE_EINVAL, "pinc=%g Pa: must be finite and >= 1 Pa", (double)pinc);
    if (chunk != 64 && chunk != 128 && chunk != 256)
        return fail(XCAPE_EINVAL, "chunk=%d not in {64, 128, 256}", chunk);
    if (tile < 0 || tile > MAX_TILE)
        return fail(XCAPE_EINVAL, "tile=%d outside 0..%d", tile, MAX_TILE);
    if (cols_per_block != 16 && cols_per_block != 32 && cols_per_block != 64 &&
        cols_per_block != 128 && cols_per_block != 256)
        return fail(XCAPE_EINVAL, "cols_per_block=%d not in {16, 32, 64, 128, 256}",
                    cols_per_block);
    if ((flags & XCAPE_WARM_START) && cols_per_block < 64)
        return fail(XCAPE_EINVAL, "warm-started plans take 64, 128 or 256 columns per block "
                                  "(cols_per_block=%d)", cols_per_block);
    // a block of fewer than 64 columns is one wave with the other lanes idle
    const int waves_per_block = cols_per_block < 64 ? 1 : cols_per_block / 64;
    if (flags & ~(uint32_t)(XCAPE_FAST_MATH | XCAPE_DEVICE_PTRS | XCAPE_WARM_START |
                            XCAPE_NO_SURROGATE))
        return fail(XCAPE_EINVAL, "unsupported flags 0x%x", flags);
    // validate the layout on the host copies and build the work order
    if (offsets_host[0] != 0) return fail(XCAPE_EINVAL, "offsets[0] must be 0");
    std::vector<int32_t> order((size_t)ncol);
    int64_t maxnk = 0;
    for (int64_t i = 0; i < ncol; ++i) {
        const int64_t n = offsets_host[i + 1] - offsets_host[i];
        if (n < 1 || n > (1 << 20))
            return fail(XCAPE_EINVAL, "column %lld has %lld levels", (long long)i, (long long)n);
        if (source_host[i] < 1 || source_host[i] > 3)
            return fail(XCAPE_EINVAL, "source[%lld]=%d not in 1..3", (long long)i, source_host[i]);
        if (adiabat_host[i] < 1 || adiabat_host[i] > 4)
            return fail(XCAPE_EINVAL, "adiabat[%lld]=%d not in 1..4", (long long)i,
                        adiabat_host[i]);
        maxnk = std::max(maxnk, n);
        order[(size_t)i] = (int32_t)i;
    }
    // the staged copies are indexed with 32-bit offsets
    if ((uint64_t)maxnk * (uint64_t)ncol >= ((uint64_t)1 << 32))
        return fail(XCAPE_EINVAL, "batch too large: %lld columns x %lld levels (split it)",
                    (long long)ncol, (long long)maxnk);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (adiabat_host[x] != adiabat_host[y]) return adiabat_host[x] < adiabat_host[y];
        if (source_host[x] != source_host[y]) return source_host[x] < source_host[y];
        return offsets_host[x + 1] - offsets_host[x] > offsets_host[y + 1] - offsets_host[y];
    });
    std::vector<int32_t> nk_sorted((size_t)ncol), src_sorted((size_t)ncol);
    for (int64_t j = 0; j < ncol; ++j) {
        const int32_t c = order[(size_t)j];
        nk_sorted[(size_t)j] = (int32_t)(offsets_host[c + 1] - offsets_host[c]);
        src_sorted[(size_t)j] = source_host[c];
    }
    int rc = ensure_device();
    if (rc) return rc;
    auto *pl = new xcape_ragged_plan;
    for (int g = 0; g < NSEG; ++g) {
        const auto lower = [&](int a) {
            return std::lower_bound(order.begin(), order.end(), a,
                                    [&](int32_t c, int v) { return adiabat_host[c] < v; }) -
                   order.begin();
        };
        pl->seg_begin[g] = lower(g + 1);
        pl->seg_end[g] = lower(g + 2);
    }
    const size_t lvl = (size_t)maxnk * (size_t)ncol;
    const size_t stage_floats = 3 * lvl + 3 * (size_t)ncol;
    const size_t bi = sizeof(int32_t) * (size_t)ncol;
    hipError_t e;
    if ((e = hipMalloc((void **)&pl->d_order, bi)) || (e = hipMalloc((void **)&pl->d_nk, bi)) ||
        (e = hipMalloc((void **)&pl->d_src, bi)) ||
        (e = hipMalloc((void **)&pl->d_stage, sizeof(float) * stage_floats)) ||
        (e = hipMalloc((void **)&pl->d_counter, sizeof(unsigned int) * NSEG)) ||
        (e = hipMalloc((void **)&pl->d_stats, sizeof(unsigned long long) * ST_COUNT)) ||
        (e = hipMemcpy(pl->d_order, order.data(), bi, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(pl->d_nk, nk_sorted.data(), bi, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(pl->d_src, src_sorted.data(), bi, hipMemcpyHostToDevice))) {
        destroy(pl);
        return hip_fail(e, "plan allocation");
    }
    pl->fast = flags & XCAPE_FAST_MATH;
    pl->warm = flags & XCAPE_WARM_START;
    pl->block = 64 * waves_per_block;
    pl->lds_bytes = tile > 0 ? sizeof(float) * 3 * (size_t)tile * 64 * waves_per_block : 0;
    StageArgs &sa = pl->stage;
    sa.offsets = offsets; sa.order = pl->d_order; sa.nk_sorted = pl->d_nk;
    sa.p = p; sa.t = t; sa.td = td; sa.ps = ps; sa.ts = ts; sa.tds = tds;
    sa.lp = pl->d_stage; sa.lt = sa.lp + lvl; sa.ltd = sa.lt + lvl;
    sa.sps = sa.ltd + lvl; sa.sts = sa.sps + ncol; sa.stds = sa.sts + ncol;
    sa.ncol = ncol; sa.maxnk = maxnk;
    RaggedArgs &a = pl->args;
    a.order = pl->d_order; a.nk_sorted = pl->d_nk; a.src_sorted = pl->d_src;
    a.lp = sa.lp; a.lt = sa.lt; a.ltd = sa.ltd; a.sps = sa.sps; a.sts = sa.sts; a.stds = sa.stds;
    a.ncol = ncol; a.pinc = pinc; a.ml_depth = ml_depth;
    a.cape = cape; a.cin = cin; a.zout = zmulev; a.mulvl = mulev;
    a.chunk = chunk; a.tile = tile;
    a.cpw = cols_per_block < 64 ? cols_per_block : 64;
    if (pl->warm) {
        // each segment as a grid of its own for the warm kernels: its columns
        // start seg_begin into each level row of the staged copies
        size_t per_seg_floats[NSEG] = {}, total = 0;
        for (int g = 0; g < NSEG; ++g) {
            const int64_t n = pl->seg_end[g] - pl->seg_begin[g];
            per_seg_floats[g] = n > 0 ? (size_t)warm_scratch_floats(maxnk, false) * (size_t)n : 0;
            total += per_seg_floats[g];
        }
        if ((e = hipMalloc((void **)&pl->d_warm, sizeof(float) * (total ? total : 1))) ||
            (e = hipMalloc((void **)&pl->d_work, sizeof(unsigned int) * 8 * NSEG))) {
            destroy(pl);
            return hip_fail(e, "plan allocation");
        }
        size_t off = 0;
        for (int g = 0; g < NSEG; ++g) {
            const int64_t b = pl->seg_begin[g], n = pl->seg_end[g] - b;
            if (n <= 0) continue;
            // the warm kernels index level environments with 32-bit offsets
            if ((uint64_t)(maxnk + 1) * (uint64_t)n >= ((uint64_t)1 << 32)) {
                destroy(pl);
                return fail(XCAPE_EINVAL, "batch too large for the warm start (split it)");
            }
            CapeGridArgs &w = pl->warm_args[g];
            w.p = sa.lp + b; w.t = sa.lt + b; w.td = sa.ltd + b;
            w.ps = sa.sps + b; w.ts = sa.sts + b; w.tds = sa.stds + b;
            w.nk_in = maxnk; w.ncol = n; w.ls = ncol;
            w.nk_col = pl->d_nk + b; w.src_col = pl->d_src + b; w.out_index = pl->d_order + b;
            w.pinc = pinc; w.ml_depth = ml_depth; w.source = 1; w.adiabat = g + 1;
            w.cape = cape; w.cin = cin; w.zout = zmulev; w.mulvl = mulev;
            w.warm = 1;
            w.fallback_tol = 1.0e-3f;
            w.surrogate = (flags & XCAPE_NO_SURROGATE) ? nullptr : surrogate_device();
            float *ws = pl->d_warm + off;
            const size_t lv = (size_t)(maxnk + 1) * (size_t)n;
            w.env = ws;
            w.par = ws + 4 * lv;
            w.env_p = nullptr;
            w.work = pl->d_work + 8 * g;
            off += per_seg_floats[g];
        }
    }
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    for (int g = 0; g < NSEG; ++g) {
        const int64_t n = pl->seg_end[g] - pl->seg_begin[g];
        if (n <= 0) continue;
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu,
            reinterpret_cast<const void *>(pick(pl->fast, g + 1, tile > 0, false)),
            pl->block, pl->lds_bytes);
        if (e != hipSuccess || per_cu < 1) per_cu = 1;
        // enough blocks to fill every CU once, never more than the work needs
        const int64_t waves_needed = (n + chunk - 1) / chunk;   // (chunk: columns per grab)
        const int64_t blocks_needed = (waves_needed + waves_per_block - 1) / waves_per_block;
        pl->seg_grid[g] =
            (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ncu * per_cu, blocks_needed));
    }
    *out = pl;
    return XCAPE_OK;
}

int xcape_ragged_plan_run(xcape_ragged_plan *pl, xcape_stats *stats, void *stream)
{
    if (!pl) return fail(XCAPE_EINVAL, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream
    hipError_t e;
    if (stats) {
        // counted run: direct launches with the counters attached (not the graph)
        if ((e = zero_counters(pl->d_stats, ST_COUNT, s)) || (e = enqueue(pl, true, s)))
            return hip_fail(e, "ragged launch");
        unsigned long long h[ST_COUNT];
        if ((e = hipMemcpyAsync(h, pl->d_stats, sizeof h, hipMemcpyDeviceToHost, s)) ||
            (e = hipStreamSynchronize(s)))
            return hip_fail(e, "ragged stats");
        copy_stats(h, stats);
        return XCAPE_OK;
    }
    if (!pl->exec) {
        // capture counter reset + staging + the segments' kernels once, the
        // segments forked onto streams of their own (parallel graph branches)
        hipStream_t cs = nullptr, ss[NSEG] = {};
        hipEvent_t fork = nullptr, join[NSEG] = {};
        e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
        for (int g = 0; g < NSEG && e == hipSuccess; ++g)
            if ((e = hipStreamCreateWithFlags(&ss[g], hipStreamNonBlocking)) == hipSuccess)
                e = hipEventCreateWithFlags(&join[g], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
        if (e == hipSuccess) {
            e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t le = enqueue(pl, false, cs, ss, fork, join);
                e = hipStreamEndCapture(cs, &pl->graph);
                if (le != hipSuccess) e = le;
            }
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
        for (int g = 0; g < NSEG; ++g) {
            if (join[g]) (void)hipEventDestroy(join[g]);
            if (ss[g]) (void)hipStreamDestroy(ss[g]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (cs) (void)hipStreamDestroy(cs);
        if (e != hipSuccess) return hip_fail(e, "graph capture");
    }
    if ((e = hipGraphLaunch(pl->exec, s))) return hip_fail(e, "graph launch");
    return XCAPE_OK;
}

int xcape_ragged_plan_info(const xcape_ragged_plan *pl, int *grid, int *block)
{
    if (!pl) return fail(XCAPE_EINVAL, "plan is NULL");
    int gmax = 0;
    for (int g = 0; g < NSEG; ++g) gmax = std::max(gmax, pl->seg_grid[g]);
    if (grid) *grid = gmax;
    if (block) *block = pl->block;
    return XCAPE_OK;
}

int xcape_ragged_plan_destroy(xcape_ragged_plan *pl)
{
    destroy(pl);
    return XCAPE_OK;
}

}  // extern "C"
