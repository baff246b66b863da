// Rectangular-grid kernels (sigma or shared pressure levels), included by
// one translation unit per math mode with XC_MATH defined.
//
// Mapping: one column per lane, 64 columns per wavefront, blockDim columns
// per workgroup.  The ascent is compute-bound (one column's ~2000
// fixed-point iterations against ~400 bytes of profile), so it reads the
// profiles in whichever layout the caller holds them.  The helicity chain is
// light arithmetic over many levels: it is given level-major profiles (the
// C-ABI stages column-contiguous input through transpose_levels first), so
// that the 64 lanes of a wave read 64 consecutive words.
//
// The CAPE, helicity and surrogate kernels are separate launches: each keeps
// its own register budget (the FP64 helicity chain and the MFMA evaluation
// would otherwise cap the ascent's occupancy).
#pragma once
#include "cape_column.h"
#include "cape_launch.h"
#include "mlp_device.h"
#include "srh_column.h"

namespace xcape {
namespace XC_NS {

// Level k of a column (k = 0: the surface values).  Level k >= 1 is element
// off + (k - 1) * ls of the (uniform) array bases t, td, u, v, and element
// poff + (k - 1) * pls of p (the shared level vector on a pressure grid).
// IDX: uint32_t when every element index fits, so that a lane carries one
// 32-bit offset per array family instead of a 64-bit pointer per array.
template <typename IDX>
struct SigmaCol {
    const float *p, *t, *td, *u, *v;
    IDX off, poff, ls, pls;
    float ps, ts, tds, us, vs;
    __device__ __forceinline__ IDX at(int k) const { return off + (IDX)(k - 1) * ls; }
    __device__ __forceinline__ float P(int k) const
    {
        return k == 0 ? ps : p[poff + (IDX)(k - 1) * pls];
    }
    __device__ __forceinline__ float Tc(int k) const { return k == 0 ? ts : t[at(k)]; }
    __device__ __forceinline__ float Tdc(int k) const { return k == 0 ? tds : td[at(k)]; }
    __device__ __forceinline__ float U(int k) const { return k == 0 ? us : u[at(k)]; }
    __device__ __forceinline__ float V(int k) const { return k == 0 ? vs : v[at(k)]; }
    // a level above the surface (k >= 1): no surface select
    __device__ __forceinline__ Raw raw_above(int k) const
    {
        return Raw{p[poff + (IDX)(k - 1) * pls], t[at(k)], td[at(k)]};
    }
};

// First level at or above the ground, 1-based: the reference's
// argmin(masked_less(ps - p, 0)) + 1 (core.py:286-289), with numpy's rules:
// a NaN difference is unmasked and wins the argmin; all-masked gives 0.
__device__ __forceinline__ int start_level(const float *p1d, int nk_in, float ps)
{
    int best = -1;
    float bv = 0.0f;
    for (int j = 0; j < nk_in; ++j) {
        float d = ps - p1d[j];
        if (d != d) return j + 1;
        if (d >= 0.0f && (best < 0 || d < bv)) {
            best = j;
            bv = d;
        }
    }
    return (best < 0 ? 0 : best) + 1;
}

__device__ __forceinline__ void block_stats_add(const LaneStats &st, bool ran,
                                                unsigned long long *stats)
{
    constexpr int NC = 7;
    __shared__ unsigned long long red[NC][16];
    unsigned long long v[NC] = {(unsigned long long)ran, st.levels, st.substeps, st.iterations,
                                st.nonconverged, st.fallbacks, st.first};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        unsigned long long x = v[c];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[c][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        unsigned long long s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
        if (s) atomicAdd(&stats[threadIdx.x], s);
    }
}

// Column set-up shared by the kernels: the accessor of column `col` and its
// level count (surface included).  Returns false for lanes past the grid.
template <typename IDX>
__device__ __forceinline__ bool column_at(const CapeGridArgs &a, int64_t col, SigmaCol<IDX> &c,
                                          int &nk)
{
    if (col >= a.ncol) return false;
    const float ps = a.ps[col];
    c.ls = (IDX)a.ls;
    IDX off = a.ls == 1 ? (IDX)col * (IDX)a.nk_in : (IDX)col;
    if (a.p1d) {
        int s1 = a.start ? a.start[col] : start_level(a.p1d, (int)a.nk_in, ps);
        // a start level outside 1..nk_in cannot come from the rule above;
        // clamp caller-supplied ones so the column is never read outside
        s1 = s1 < 1 ? 1 : (s1 > (int)a.nk_in ? (int)a.nk_in : s1);
        const int s0 = s1 - 1;
        c.p = a.p1d;
        c.poff = (IDX)s0;
        c.pls = 1;
        off += (IDX)s0 * c.ls;
        nk = (int)a.nk_in - s0 + 1;
    } else {
        c.p = a.p;
        c.poff = off;
        c.pls = c.ls;
        nk = (int)a.nk_in + 1;
    }
    c.off = off;
    c.t = a.t;
    c.td = a.td;
    c.u = a.u;
    c.v = a.v;
    c.ps = ps;
    c.ts = a.ts[col];
    c.tds = a.tds[col];
    return true;
}

// First-sub-step inputs of the surrogate for a parcel: ln p2, the parcel's
// theta-e at its start (the reference's parcel formula, getthe(p, t, t, qv),
// :505) and its total water, for the sub-step p -> p - dp (p = pc.e.p) of the
// first ascent level (:408-429); lp1 = ln(p2/p1) of that sub-step (0: none).
template <typename T, class C>
__device__ __forceinline__ void surrogate_inputs(const C &c, int nk, const Parcel<T> &pc,
                                                 float pinc, float &f0, float &f1, float &f2,
                                                 T &lp1)
{
    using F = Fn<T, XC_MATH>;
    const Env<T> e1 = env_from<T, XC_MATH>(c.raw_above(pc.kmax + 1));
    T dp = pc.e.p - e1.p;
    const int nl = substep_count<T, XC_MATH>(dp, T(pinc));
    const T p2 = pc.e.p - dp;
    if (nl >= 1 && p2 > T(0.0)) {
        lp1 = F::log(F::div(p2, pc.e.p));
        f0 = (float)F::log(p2);
        f1 = (float)thetae<T, XC_MATH>(pc.e.p, pc.t2, pc.t2, pc.qv2);
        f2 = (float)pc.qv2;
    }
}

// CAPE/CIN, one column per lane.  MODE: 0 the reference's loops with the
// cold start (cape_column); 1 the warm-started ascent (cape_column_warm);
// 2 the same loop with every sub-step cold (to separate the loop's share of
// the speed-up from the prediction's).  SUR (with MODE 1): a surrogate is
// loaded; every wave first evaluates it for its 64 columns' first sub-steps
// (one wave-wide BF16 MFMA evaluation, a point every lane of the block
// reaches), and each lane starts its ascent from that iterate.  STATS:
// accumulate the work counters into a.stats (a separate instantiation, so
// that the counters do not occupy registers in the timed kernels).
template <typename T, typename IDX, bool ICE, bool PSEUDO, int MODE, bool SUR, bool STATS>
__global__ void __launch_bounds__(256) cape_grid_kernel(CapeGridArgs a)
{
    extern __shared__ uint32_t dyn_lds[];
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    SigmaCol<IDX> c{};
    int nk = 0;
    const bool in_grid = column_at(a, col, c, nk);
    T *cape = static_cast<T *>(a.cape), *cin = static_cast<T *>(a.cin),
      *zout = static_cast<T *>(a.zout);
    if constexpr (MODE == 0) {
        LaneStats st;
        bool ran = false;
        if (in_grid) {
            ColumnOut<T> o;
            if (c.ts > 0.0f) {
                ran = true;
                cape_column<T, XC_MATH, ICE, PSEUDO>(c, nk, T(a.pinc), a.source, T(a.ml_depth),
                                                     o, st);
            } else {
                o.cape = T(0); o.cin = T(0); o.zout = T(0); o.mulvl = 0;
            }
            cape[col] = o.cape;
            cin[col] = o.cin;
            a.mulvl[col] = o.mulvl;
            zout[col] = o.zout;
        }
        if constexpr (STATS) block_stats_add(st, ran, a.stats);
    } else {
        const bool ran = in_grid && c.ts > 0.0f;
        Parcel<T> pc;
        bool ok = false;
        T lp1 = T(0.0);
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
        if (ran) {
            ok = parcel_source<T, XC_MATH>(c, nk, a.source, T(a.ml_depth), pc);
            // stored before the ascent: one register less through it
            a.mulvl[col] = pc.mulvl;
            if (SUR && ok && pc.kmax < nk - 1) surrogate_inputs(c, nk, pc, a.pinc, f0, f1, f2, lp1);
        }
        float r = 0.0f;
        if constexpr (SUR)
            r = surrogate_eval_wave(static_cast<const SurrogateDev *>(a.surrogate), f0, f1, f2,
                                    dyn_lds + (threadIdx.x >> 6) * MLP_LDS_WORDS_PER_WAVE);
        LaneStats st;
        if (in_grid) {
            ColumnOut<T> o{};
            if (ok) {
                const T g = lp1 != T(0.0) ? pc.th2 + T(r) * lp1 : T(NAN);
                cape_column_warm<T, XC_MATH, ICE, PSEUDO>(c, nk, T(a.pinc), pc, g,
                                                          T(a.fallback_tol), o, st, MODE == 1);
            } else if (ran) {
                // unknown source (unreachable through the C-ABI's checks)
                o.zout = T(-999999.0);
            } else {
                a.mulvl[col] = 0;
            }
            cape[col] = o.cape;
            cin[col] = o.cin;
            zout[col] = o.zout;
        }
        if constexpr (STATS) block_stats_add(st, ran, a.stats);
    }
}

// Storm-relative helicity chain, one column per lane.  TB: the storm-motion
// precision (float: the reference's; double: the FP64 re-run).
template <typename TB>
__global__ void __launch_bounds__(256) srh_grid_kernel(CapeGridArgs a)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    SigmaCol<uint64_t> c{};
    int nk = 0;
    if (!column_at(a, col, c, nk)) return;
    c.us = a.us[col];
    c.vs = a.vs[col];
    SrhOutT<TB> so;
    srh_column<TB>(c, nk, a.depth, 2.0, so);
    a.srh_rm[col] = so.srh_rm;
    a.srh_lm[col] = so.srh_lm;
    if (a.rm) {
        TB *rm = static_cast<TB *>(a.rm), *lm = static_cast<TB *>(a.lm),
           *mw = static_cast<TB *>(a.mean6);
        rm[2 * col] = so.rm_u;
        rm[2 * col + 1] = so.rm_v;
        lm[2 * col] = so.lm_u;
        lm[2 * col + 1] = so.lm_v;
        mw[2 * col] = so.mean_u;
        mw[2 * col + 1] = so.mean_v;
    }
}

template <typename T, typename IDX, int MODE, bool SUR, bool STATS>
hipError_t launch_cape_s(const CapeGridArgs &a, int block, hipStream_t s)
{
    dim3 grid((unsigned)((a.ncol + block - 1) / block));
    const size_t lds = SUR ? sizeof(uint32_t) * MLP_LDS_WORDS_PER_WAVE * (size_t)(block / 64) : 0;
    switch (a.adiabat) {
    case 1: cape_grid_kernel<T, IDX, false, true, MODE, SUR, STATS><<<grid, block, lds, s>>>(a); break;
    case 2: cape_grid_kernel<T, IDX, false, false, MODE, SUR, STATS><<<grid, block, lds, s>>>(a); break;
    case 3: cape_grid_kernel<T, IDX, true, true, MODE, SUR, STATS><<<grid, block, lds, s>>>(a); break;
    case 4: cape_grid_kernel<T, IDX, true, false, MODE, SUR, STATS><<<grid, block, lds, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// 32-bit element offsets when every index of the profile arrays fits (the
// stats launches, which only count work, always use 64-bit ones)
template <typename T, int MODE, bool SUR>
hipError_t launch_cape(const CapeGridArgs &a, int block, hipStream_t s)
{
    if (a.stats) return launch_cape_s<T, uint64_t, MODE, SUR, true>(a, block, s);
    const bool fits = (uint64_t)a.ncol * (uint64_t)a.nk_in < ((uint64_t)1 << 32);
    return fits ? launch_cape_s<T, uint32_t, MODE, SUR, false>(a, block, s)
                : launch_cape_s<T, uint64_t, MODE, SUR, false>(a, block, s);
}

template <typename T>
hipError_t launch_t(const CapeGridArgs &a, int kind, int block, hipStream_t s)
{
    if (a.ncol <= 0) return hipSuccess;
    if constexpr (sizeof(T) == 8) {
        // the FP64 yardstick: CAPE with the reference's cold start, and the
        // helicity chain with FP64 storm motion
        if (kind == GRID_SRH) {
            dim3 grid((unsigned)((a.ncol + block - 1) / block));
            srh_grid_kernel<double><<<grid, block, 0, s>>>(a);
            return hipGetLastError();
        }
        if (kind != GRID_CAPE || a.warm) return hipErrorInvalidValue;
        return launch_cape<T, 0, false>(a, block, s);
    } else {
        dim3 grid((unsigned)((a.ncol + block - 1) / block));
        switch (kind) {
        case GRID_CAPE:
            if (a.warm == 0) return launch_cape<T, 0, false>(a, block, s);
            if (a.warm == 2) return launch_cape<T, 2, false>(a, block, s);
            return a.surrogate ? launch_cape<T, 1, true>(a, block, s)
                               : launch_cape<T, 1, false>(a, block, s);
        case GRID_SRH:
            srh_grid_kernel<float><<<grid, block, 0, s>>>(a);
            return hipGetLastError();
        default: return hipErrorInvalidValue;
        }
    }
}

}  // namespace XC_NS
}  // namespace xcape
