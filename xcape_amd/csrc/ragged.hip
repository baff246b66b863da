// Ragged batches: columns with different level counts, parcel sources and
// adiabats in one persistent launch.
//
// Layout (CSR): column i owns levels offsets[i] .. offsets[i+1]-1 of the
// concatenated p, t, td arrays (hPa, degC); surface values, source and
// adiabat are per column.
//
// Scheduling: a plan sorts the columns once (host, stable) by adiabat, then
// source, then descending level count, so that the 64 lanes of a wavefront
// run the same code path on columns of similar length, and long columns are
// dealt first.  The kernel is persistent: a grid sized to fill the device
// once, every wavefront taking the next `chunk` sorted columns from a device
// counter until the batch is drained (every wave reaches the exit test, so
// the grid always drains).  The counter is reset by a one-thread kernel in
// the same stream (graph-safe: no memset node).
//
// Profiles are staged through LDS: when a wavefront takes a chunk it loads
// the first `tile` levels of its columns into LDS with each column's levels
// read as contiguous 16-byte vectors, so the source selection and the lowest
// ascent levels (where most columns start) read LDS; levels above the tile
// are read from global memory (L2-resident after the first touch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/xcape_hip.h"
#include "capi_common.h"
#include "cape_column.h"
#include "cape_launch.h"

namespace xcape {
namespace {

constexpr int MAX_TILE = 64;   // upper bound of the LDS level-tile depth

struct RaggedArgs {
    const int64_t *offsets;
    const int32_t *order;      // sorted column order
    const int32_t *source, *adiabat;
    const float *p, *t, *td, *ps, *ts, *tds;
    int64_t ncol;
    float pinc, ml_depth;
    float *cape, *cin, *zout;
    int32_t *mulvl;
    unsigned int *counter;
    unsigned long long *stats;
    int chunk;                 // columns a wavefront takes per grab (64 or 128)
    int tile;                  // levels staged in LDS (0 = none)
    int warm;                  // warm-started iteration (see cape_column_warm)
    float fallback_tol;
};

template <int MATH, bool ICE, bool PSEUDO, class C>
__device__ __forceinline__ void run_column(const RaggedArgs &a, const C &c, int nk, int src,
                                           ColumnOut<float> &o, LaneStats &st)
{
    if (a.warm) {
        Parcel<float> pc;
        if (parcel_source<float, MATH>(c, nk, src, a.ml_depth, pc))
            cape_column_warm<float, MATH, ICE, PSEUDO>(c, nk, a.pinc, pc, NAN, a.fallback_tol,
                                                       o, st);
        else {
            o.cape = 0.0f; o.cin = 0.0f; o.mulvl = pc.mulvl; o.zout = -999999.0f;
        }
    } else {
        cape_column<float, MATH, ICE, PSEUDO>(c, nk, a.pinc, src, a.ml_depth, o, st);
    }
}

// Column accessor: levels below `tile` from this lane's LDS slice, the rest
// from global memory.
struct TiledCol {
    const float *p, *t, *td;       // global, level k-1 at [k-1]
    const float *lp, *lt, *ltd;    // LDS slice, level k-1 at [(k-1) * 64]
    float ps, ts, tds;
    int tile;
    __device__ __forceinline__ float P(int k) const
    {
        return k == 0 ? ps : (k - 1 < tile ? lp[(k - 1) * 64] : p[k - 1]);
    }
    __device__ __forceinline__ float Tc(int k) const
    {
        return k == 0 ? ts : (k - 1 < tile ? lt[(k - 1) * 64] : t[k - 1]);
    }
    __device__ __forceinline__ float Tdc(int k) const
    {
        return k == 0 ? tds : (k - 1 < tile ? ltd[(k - 1) * 64] : td[k - 1]);
    }
    __device__ __forceinline__ Raw raw_above(int k) const
    {
        return k - 1 < tile ? Raw{lp[(k - 1) * 64], lt[(k - 1) * 64], ltd[(k - 1) * 64]}
                            : Raw{p[k - 1], t[k - 1], td[k - 1]};
    }
};

__global__ void reset_counter(unsigned int *c) { *c = 0u; }

template <int MATH>
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
        if ((int64_t)base >= a.ncol) break;
        for (int sub = 0; sub < a.chunk; sub += 64) {
            const int64_t i = (int64_t)base + sub + lane;
            const bool active = i < a.ncol;
            int col = 0, nk_in = 0;
            int64_t off = 0;
            if (active) {
                col = a.order[i];
                off = a.offsets[col];
                nk_in = (int)(a.offsets[col + 1] - off);
            }
            const int tile = min(a.tile, nk_in);
            // stage levels 0..tile-1 of this lane's column: 16-byte reads where
            // the column start is aligned, scalar reads otherwise
            if (active) {
                const float *gp = a.p + off, *gt = a.t + off, *gtd = a.td + off;
                int k = 0;
                if ((off & 3) == 0) {
                    for (; k + 4 <= tile; k += 4) {
                        const float4 vp = *reinterpret_cast<const float4 *>(gp + k);
                        const float4 vt = *reinterpret_cast<const float4 *>(gt + k);
                        const float4 vd = *reinterpret_cast<const float4 *>(gtd + k);
                        lp[(k + 0) * 64] = vp.x; lp[(k + 1) * 64] = vp.y;
                        lp[(k + 2) * 64] = vp.z; lp[(k + 3) * 64] = vp.w;
                        lt[(k + 0) * 64] = vt.x; lt[(k + 1) * 64] = vt.y;
                        lt[(k + 2) * 64] = vt.z; lt[(k + 3) * 64] = vt.w;
                        ltd[(k + 0) * 64] = vd.x; ltd[(k + 1) * 64] = vd.y;
                        ltd[(k + 2) * 64] = vd.z; ltd[(k + 3) * 64] = vd.w;
                    }
                }
                for (; k < tile; ++k) {
                    lp[k * 64] = gp[k];
                    lt[k * 64] = gt[k];
                    ltd[k * 64] = gtd[k];
                }
            }
            // each lane reads back only its own LDS slots: no barrier needed
            if (active) {
                const float ts = a.ts[col];
                if (ts > 0.0f) {
                    ++ran;
                    TiledCol c{a.p + off, a.t + off, a.td + off, lp, lt, ltd,
                               a.ps[col], ts, a.tds[col], tile};
                    ColumnOut<float> o;
                    const int nk = nk_in + 1;
                    const int src = a.source[col];
                    switch (a.adiabat[col]) {
                    case 1: run_column<MATH, false, true>(a, c, nk, src, o, st); break;
                    case 2: run_column<MATH, false, false>(a, c, nk, src, o, st); break;
                    case 3: run_column<MATH, true, true>(a, c, nk, src, o, st); break;
                    default: run_column<MATH, true, false>(a, c, nk, src, o, st); break;
                    }
                    a.cape[col] = o.cape;
                    a.cin[col] = o.cin;
                    a.mulvl[col] = o.mulvl;
                    a.zout[col] = o.zout;
                } else {
                    a.cape[col] = 0.0f;
                    a.cin[col] = 0.0f;
                    a.zout[col] = 0.0f;
                    a.mulvl[col] = 0;
                }
            }
        }
    }
    if (a.stats) {
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

}  // namespace
}  // namespace xcape

using namespace xcape;

struct xcape_ragged_plan {
    RaggedArgs args{};
    int32_t *d_order = nullptr;
    unsigned int *d_counter = nullptr;
    unsigned long long *d_stats = nullptr;
    int grid = 0, block = 256;
    size_t lds_bytes = 0;
    bool fast = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

hipError_t enqueue(const xcape_ragged_plan *pl, hipStream_t s)
{
    reset_counter<<<1, 1, 0, s>>>(pl->d_counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (pl->fast)
        ragged_kernel<MATH_FAST><<<pl->grid, pl->block, pl->lds_bytes, s>>>(pl->args);
    else
        ragged_kernel<MATH_ACCURATE><<<pl->grid, pl->block, pl->lds_bytes, s>>>(pl->args);
    return hipGetLastError();
}

void destroy(xcape_ragged_plan *pl)
{
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->d_order) (void)hipFree(pl->d_order);
    if (pl->d_counter) (void)hipFree(pl->d_counter);
    if (pl->d_stats) (void)hipFree(pl->d_stats);
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
                             int waves_per_block, uint32_t flags, xcape_ragged_plan **out)
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
    if (waves_per_block < 1 || waves_per_block > 4)
        return fail(XCAPE_EINVAL, "waves_per_block=%d outside 1..4", waves_per_block);
    if (flags & ~(uint32_t)(XCAPE_FAST_MATH | XCAPE_DEVICE_PTRS | XCAPE_WARM_START))
        return fail(XCAPE_EINVAL, "unsupported flags 0x%x", flags);
    // validate the layout on the host copies and build the work order
    if (offsets_host[0] != 0) return fail(XCAPE_EINVAL, "offsets[0] must be 0");
    std::vector<int32_t> order((size_t)ncol);
    for (int64_t i = 0; i < ncol; ++i) {
        const int64_t n = offsets_host[i + 1] - offsets_host[i];
        if (n < 1 || n > (1 << 20))
            return fail(XCAPE_EINVAL, "column %lld has %lld levels", (long long)i, (long long)n);
        if (source_host[i] < 1 || source_host[i] > 3)
            return fail(XCAPE_EINVAL, "source[%lld]=%d not in 1..3", (long long)i, source_host[i]);
        if (adiabat_host[i] < 1 || adiabat_host[i] > 4)
            return fail(XCAPE_EINVAL, "adiabat[%lld]=%d not in 1..4", (long long)i,
                        adiabat_host[i]);
        order[(size_t)i] = (int32_t)i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (adiabat_host[x] != adiabat_host[y]) return adiabat_host[x] < adiabat_host[y];
        if (source_host[x] != source_host[y]) return source_host[x] < source_host[y];
        return offsets_host[x + 1] - offsets_host[x] > offsets_host[y + 1] - offsets_host[y];
    });
    int rc = ensure_device();
    if (rc) return rc;
    auto *pl = new xcape_ragged_plan;
    hipError_t e;
    if ((e = hipMalloc((void **)&pl->d_order, sizeof(int32_t) * (size_t)ncol)) ||
        (e = hipMalloc((void **)&pl->d_counter, sizeof(unsigned int))) ||
        (e = hipMalloc((void **)&pl->d_stats, sizeof(unsigned long long) * ST_COUNT)) ||
        (e = hipMemcpy(pl->d_order, order.data(), sizeof(int32_t) * (size_t)ncol,
                       hipMemcpyHostToDevice))) {
        destroy(pl);
        return hip_fail(e, "plan allocation");
    }
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    pl->block = 64 * waves_per_block;
    pl->lds_bytes = sizeof(float) * 3 * (size_t)(tile > 0 ? tile : 1) * 64 * waves_per_block;
    int per_cu = 0;
    if (flags & XCAPE_FAST_MATH)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ragged_kernel<MATH_FAST>,
                                                         pl->block, pl->lds_bytes);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ragged_kernel<MATH_ACCURATE>,
                                                         pl->block, pl->lds_bytes);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    // enough blocks to fill every CU once, never more than the work needs
    const int64_t waves_needed = (ncol + chunk - 1) / chunk;
    const int64_t blocks_needed = (waves_needed + waves_per_block - 1) / waves_per_block;
    pl->grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ncu * per_cu, blocks_needed));
    pl->fast = flags & XCAPE_FAST_MATH;
    RaggedArgs &a = pl->args;
    a.offsets = offsets; a.order = pl->d_order; a.source = source; a.adiabat = adiabat;
    a.p = p; a.t = t; a.td = td; a.ps = ps; a.ts = ts; a.tds = tds;
    a.ncol = ncol; a.pinc = pinc; a.ml_depth = ml_depth;
    a.cape = cape; a.cin = cin; a.zout = zmulev; a.mulvl = mulev;
    a.counter = pl->d_counter; a.stats = nullptr; a.chunk = chunk; a.tile = tile;
    a.warm = (flags & XCAPE_WARM_START) ? 1 : 0;
    a.fallback_tol = 1.0e-3f;
    *out = pl;
    return XCAPE_OK;
}

int xcape_ragged_plan_run(xcape_ragged_plan *pl, xcape_stats *stats, void *stream)
{
    if (!pl) return fail(XCAPE_EINVAL, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;   // NULL: the default stream
    hipError_t e;
    if (stats) {
        // counted run: plain launch with the counters attached (not the graph)
        xcape_ragged_plan tmp = *pl;
        tmp.args.stats = pl->d_stats;
        if ((e = zero_counters(pl->d_stats, ST_COUNT, s)) || (e = enqueue(&tmp, s)))
            return hip_fail(e, "ragged launch");
        unsigned long long h[ST_COUNT];
        if ((e = hipMemcpyAsync(h, pl->d_stats, sizeof h, hipMemcpyDeviceToHost, s)) ||
            (e = hipStreamSynchronize(s)))
            return hip_fail(e, "ragged stats");
        copy_stats(h, stats);
        return XCAPE_OK;
    }
    if (!pl->exec) {
        // capture counter reset + persistent kernel once, replay afterwards
        hipStream_t cs;
        if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)))
            return hip_fail(e, "capture stream");
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t le = enqueue(pl, cs);
            e = hipStreamEndCapture(cs, &pl->graph);
            if (le != hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) return hip_fail(e, "graph capture");
    }
    if ((e = hipGraphLaunch(pl->exec, s))) return hip_fail(e, "graph launch");
    return XCAPE_OK;
}

int xcape_ragged_plan_info(const xcape_ragged_plan *pl, int *grid, int *block)
{
    if (!pl) return fail(XCAPE_EINVAL, "plan is NULL");
    if (grid) *grid = pl->grid;
    if (block) *block = pl->block;
    return XCAPE_OK;
}

int xcape_ragged_plan_destroy(xcape_ragged_plan *pl)
{
    destroy(pl);
    return XCAPE_OK;
}

}  // extern "C"
