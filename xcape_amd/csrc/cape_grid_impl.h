// Rectangular-grid kernels (sigma or shared pressure levels), included by
// one translation unit per math mode with XC_MATH defined.
//
// Every kernel here reads level-major profiles, level k of column j at
// k * ls + j (the C-ABI stages column-contiguous input through the LDS
// transpose first), so that the 64 lanes of a wave read 64 consecutive words
// per level; one column per lane.  The plain method is one launch
// (cape_grid_kernel); the warm start is a setup pass, the surrogate's
// evaluation and a persistent ascent (see the warm section below).  The
// CAPE, helicity and surrogate work are separate launches: each keeps its own
// register budget (the FP64 helicity chain and the MFMA evaluation would
// otherwise cap the ascent's occupancy).
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
// TI: the element type of the profile and surface arrays (float; double for
// the helicity chain's float64 inputs, CapeGridArgs::in_f64).
template <typename IDX, typename TI = float>
struct SigmaCol {
    const TI *p, *t, *td, *u, *v;
    IDX off, poff, ls, pls;
    TI ps, ts, tds, us, vs;
    __device__ __forceinline__ IDX at(int k) const { return off + (IDX)(k - 1) * ls; }
    __device__ __forceinline__ TI P(int k) const
    {
        return k == 0 ? ps : p[poff + (IDX)(k - 1) * pls];
    }
    __device__ __forceinline__ TI Tc(int k) const { return k == 0 ? ts : t[at(k)]; }
    __device__ __forceinline__ TI Tdc(int k) const { return k == 0 ? tds : td[at(k)]; }
    __device__ __forceinline__ TI U(int k) const { return k == 0 ? us : u[at(k)]; }
    __device__ __forceinline__ TI V(int k) const { return k == 0 ? vs : v[at(k)]; }
    // a level above the surface (k >= 1): no surface select
    __device__ __forceinline__ Raw raw_above(int k) const
    {
        return Raw{(float)p[poff + (IDX)(k - 1) * pls], (float)t[at(k)], (float)td[at(k)]};
    }
};

// An input array of CapeGridArgs as its element type TI (the arrays are
// declared float; with in_f64 they hold doubles).
template <typename TI>
__device__ __forceinline__ const TI *in_arr(const float *a)
{
    return reinterpret_cast<const TI *>(a);
}

// First level at or above the ground, 1-based: the reference's
// argmin(masked_less(ps - p, 0)) + 1 (core.py:286-289), with numpy's rules:
// a NaN difference is unmasked and wins the argmin; all-masked gives 0.
template <typename TI>
__device__ __forceinline__ int start_level(const TI *p1d, int nk_in, TI ps)
{
    int best = -1;
    TI bv = TI(0);
    for (int j = 0; j < nk_in; ++j) {
        TI d = ps - p1d[j];
        if (d != d) return j + 1;
        if (d >= TI(0) && (best < 0 || d < bv)) {
            best = j;
            bv = d;
        }
    }
    return (best < 0 ? 0 : best) + 1;
}

__device__ __forceinline__ void block_stats_add(const LaneStats &st, bool ran,
                                                unsigned long long *stats)
{
    constexpr int NC = ST_USED;
    __shared__ unsigned long long red[NC][16];
    unsigned long long v[NC] = {(unsigned long long)ran, st.levels, st.substeps, st.iterations,
                                st.nonconverged, st.fallbacks, st.first, st.wave_trips};
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        v[ST_BLK_LANES + b] = st.blk_lanes[b];
        v[ST_BLK_WAVES + b] = st.blk_waves[b];
    }
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
template <typename IDX, typename TI = float>
__device__ __forceinline__ bool column_at(const CapeGridArgs &a, int64_t col,
                                          SigmaCol<IDX, TI> &c, int &nk)
{
    if (col >= a.ncol) return false;
    const TI ps = in_arr<TI>(a.ps)[col];
    c.ls = (IDX)a.ls;
    IDX off = a.ls == 1 ? (IDX)col * (IDX)a.nk_in : (IDX)col;
    if (a.p1d) {
        int s1 = a.start ? a.start[col] : start_level(in_arr<TI>(a.p1d), (int)a.nk_in, ps);
        // a start level outside 1..nk_in cannot come from the rule above;
        // clamp caller-supplied ones so the column is never read outside
        s1 = s1 < 1 ? 1 : (s1 > (int)a.nk_in ? (int)a.nk_in : s1);
        const int s0 = s1 - 1;
        c.p = in_arr<TI>(a.p1d);
        c.poff = (IDX)s0;
        c.pls = 1;
        off += (IDX)s0 * c.ls;
        nk = (int)a.nk_in - s0 + 1;
    } else {
        if (a.start) {
            // p per column from its own start level (loopcape_pl,
            // CAPE_CODE_pressure_lev.f90:69-75), clamped as above
            int s1 = a.start[col];
            s1 = s1 < 1 ? 1 : (s1 > (int)a.nk_in ? (int)a.nk_in : s1);
            off += (IDX)(s1 - 1) * c.ls;
            nk = (int)a.nk_in - (s1 - 1) + 1;
        } else {
            nk = (a.nk_col ? (int)a.nk_col[col] : (int)a.nk_in) + 1;
        }
        c.p = in_arr<TI>(a.p);
        c.poff = off;
        c.pls = c.ls;
    }
    c.off = off;
    c.t = in_arr<TI>(a.t);
    c.td = in_arr<TI>(a.td);
    c.u = in_arr<TI>(a.u);
    c.v = in_arr<TI>(a.v);
    c.ps = ps;
    c.ts = in_arr<TI>(a.ts)[col];
    c.tds = in_arr<TI>(a.tds)[col];
    return true;
}

// The output slot of column `col` (a.out_index maps a batch's work order back
// to its callers' order).
__device__ __forceinline__ int64_t out_col(const CapeGridArgs &a, int64_t col)
{
    return a.out_index ? (int64_t)a.out_index[col] : col;
}

// CAPE/CIN with the reference's loops and cold start (cape_column), one
// column per lane.  STATS: accumulate the work counters into a.stats (a
// separate instantiation, so that the counters do not occupy registers in
// the timed kernels).  LEVELS: also the LFC and EL (a.p_lfc, a.p_el).
template <typename T, typename IDX, bool ICE, bool PSEUDO, bool STATS, bool LEVELS = false>
__global__ void __launch_bounds__(256) cape_grid_kernel(CapeGridArgs a)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    SigmaCol<IDX> c{};
    int nk = 0;
    const bool in_grid = column_at(a, col, c, nk);
    T *cape = static_cast<T *>(a.cape), *cin = static_cast<T *>(a.cin),
      *zout = static_cast<T *>(a.zout);
    LaneStats st;
    bool ran = false;
    if (in_grid) {
        ColumnOut<T> o;
        if (c.ts > 0.0f) {
            ran = true;
            cape_column<T, XC_MATH, ICE, PSEUDO, SigmaCol<IDX>, NoLog, LEVELS>(
                c, nk, T(a.pinc), a.source, T(a.ml_depth), o, st);
        } else {
            o.cape = T(0); o.cin = T(0); o.zout = T(0); o.mulvl = 0;
            o.p_lfc = T(NAN); o.p_el = T(NAN);
        }
        cape[col] = o.cape;
        cin[col] = o.cin;
        a.mulvl[col] = o.mulvl;
        zout[col] = o.zout;
        if constexpr (LEVELS) {
            static_cast<T *>(a.p_lfc)[col] = o.p_lfc;
            static_cast<T *>(a.p_el)[col] = o.p_el;
        }
    }
    if constexpr (STATS) block_stats_add(st, ran, a.stats);
}

// ---------------------------------------------------------------------------
// Warm-started CAPE/CIN: two launches.
//
// warm_setup_kernel, one column per lane, every lane of a wave on the same
// level: the parcel source (parcel_source), every level from the parcel's up
// as the ascent enters it (Level: pi, thv, the sub-step size and count of
// the layer below it, the 100 hPa stop test), written level-major to a.env
// as 16-byte records (level k of column j at record k * ncol + j, k = 0 the
// surface), so that a lane's level is one load of one cache line; and, with
// a loaded surrogate (SUR), its first iterate for the ascent's first
// sub-step, one wave-wide BF16 MFMA evaluation per 64 columns at the end of
// the pass.  Columns with nothing to ascend get their final outputs here.
//
// warm_ascent_kernel, persistent: each lane takes a column from the a.work
// counters, runs the WarmAscent state machine on it trip by trip, writes its
// outputs when it ends and takes the next column, until the counter passes
// ncol.  A level's set-up is then one load, issued a level ahead: the
// environment's pow, exp and divisions and the sub-step count's divisions
// ran in the setup pass with every lane on the same level, instead of under
// the mask of the few lanes that reach a new level in a given trip.
// ---------------------------------------------------------------------------

// Level records as float4 loads and stores (the layout of LevelRec).
__device__ __forceinline__ float4 level_rec(const Level<float> &lv)
{
    const LevelRec r = level_pack(lv);
    return make_float4(r.dp, r.pi, r.thv, __uint_as_float(r.meta));
}

__device__ __forceinline__ LevelRec level_rec_of(const float4 r)
{
    return LevelRec{r.x, r.y, r.z, __float_as_uint(r.w)};
}

// A column's parcel as the setup pass leaves it for the ascent: 64 bytes, four
// 16-byte loads with no dependence between them.  kmax = -1: nothing to
// ascend.  guess: the surrogate's first iterate (NaN: none).
struct ParcelRec {
    int32_t kmax, nk;
    float th2, t2;
    float qv2, b2, zk, guess;
    float e_p, e_pi, e_thv, pad0;     // the environment at kmax
    float4 next;                      // the level record of kmax + 1
};
static_assert(sizeof(ParcelRec) == 64, "parcel record layout");

struct EnvCol {
    static constexpr bool has_env = true;
    const float *rec, *p;
    uint32_t col, ls;
    __device__ __forceinline__ LevelRec rec_at(int k) const
    {
        return level_rec_of(reinterpret_cast<const float4 *>(rec)[(uint32_t)k * ls + col]);
    }
    __device__ __forceinline__ float p_at(int k) const { return p[(uint32_t)k * ls + col]; }
};

// The setup pass's sink of the source scan's environments: each level's
// record, stored as the scan evaluates the levels 0, 1, ... in order (*cover:
// the levels stored so far, *p_cov the pressure of the last).  It asks the
// most-unstable scan for every level, so that the pass reads each level of
// the profiles once.
struct RecordSink {
    static constexpr bool every_level = true;
    float4 *env;
    float *envp;
    uint32_t ls, j;
    float pinc;
    int *cover;
    float *p_cov;
    __device__ __forceinline__ void operator()(int k, const Env<float> &e) const
    {
        if (k == *cover) {
            env[(uint32_t)k * ls + j] = level_rec(level_from<float, XC_MATH>(e, *p_cov, pinc));
            if (envp) envp[(uint32_t)k * ls + j] = e.p;
            *cover = k + 1;
            *p_cov = e.p;
        }
    }
};

template <bool ICE, bool PSEUDO, bool SUR, bool STATS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
warm_setup_kernel(CapeGridArgs a)
{
    using T = float;
    using F = Fn<T, XC_MATH>;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    SigmaCol<uint64_t> c{};
    int nk = 0;
    const bool in_grid = column_at(a, col, c, nk);
    const bool ran = in_grid && c.ts > 0.0f;
    Parcel<T> pc;
    bool ok = false;
    // records: row length ls, this column's place j in it (a setup pass over
    // part of the records' columns: rec_ls, rec_col0)
    const uint32_t ls = a.rec_ls ? (uint32_t)a.rec_ls : (uint32_t)a.ncol;
    const uint32_t j = (uint32_t)(col + a.rec_col0);
    float4 *env = reinterpret_cast<float4 *>(a.env);
    float *envp = a.env_p;
    const T pinc = T(a.pinc);
    // the source scan's environments are stored as it computes them; `cover`:
    // the levels 0 .. cover-1 it has stored, p_cov the pressure of the last
    int cover = 0;
    T p_cov = T(NAN);
    if (ran) {
        const RecordSink sink{env, envp, ls, j, pinc, &cover, &p_cov};
        const int source = a.src_col ? a.src_col[col] : a.source;
        ok = parcel_source<T, XC_MATH>(c, nk, source, T(a.ml_depth), pc, sink);
        a.mulvl[out_col(a, col)] = pc.mulvl;
    }
    const bool go = ok && pc.kmax < nk - 1;
    T lp1 = T(0.0);
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    if (go) {
        // the levels from the parcel's up that the scan did not store, by
        // absolute level so that the wave's lanes write the same row
        const int k0 = pc.kmax > cover ? pc.kmax : cover;
        float4 r1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        T p_below = p_cov;
        for (int k = 0; k < nk; ++k) {
            if (k < k0) continue;
            const Env<T> e = env_from<T, XC_MATH>(load_raw(c, k));
            const float4 r = level_rec(level_from<T, XC_MATH>(e, p_below, pinc));
            env[(uint32_t)k * ls + j] = r;
            if (envp) envp[(uint32_t)k * ls + j] = e.p;
            if (k == pc.kmax + 1) r1 = r;
            p_below = e.p;
        }
        // (the scan stored level kmax + 1 already)
        if (cover > pc.kmax + 1) r1 = env[(uint32_t)(pc.kmax + 1) * ls + j];
        float4 *r4 = reinterpret_cast<float4 *>(reinterpret_cast<ParcelRec *>(a.par) + j);
        r4[0] = make_float4(__int_as_float(pc.kmax), __int_as_float(nk), pc.th2, pc.t2);
        r4[2] = make_float4(pc.e.p, pc.e.pi, pc.e.thv, 0.0f);
        r4[3] = r1;
        if constexpr (SUR) {
            // the surrogate's inputs for the ascent's first sub-step p -> p - dp
            // (p = pc.e.p, :408-429): ln p2, the parcel's theta-e at its start
            // (the reference's parcel formula, getthe(p, t, t, qv), :505) and
            // its total water; lp1 = ln(p2/p1) (0: no such sub-step)
            const Level<T> l1 = level_unpack(level_rec_of(r1), 0.0f);
            const T p2 = pc.e.p - l1.dp;
            if (l1.nloop >= 1 && p2 > T(0.0)) {
                lp1 = F::log(F::div(p2, pc.e.p));
                f0 = (float)F::log(p2);
                f1 = (float)thetae<T, XC_MATH>(pc.e.p, pc.t2, pc.t2, pc.qv2);
                f2 = (float)pc.qv2;
            }
        } else {
            r4[1] = make_float4(pc.qv2, pc.b2, pc.zk, T(NAN));
        }
    } else if (in_grid) {
        // ts <= 0 degC: zeros; no level above the parcel (or an unknown
        // source): CAPE = CIN = 0 and the height's initial -999999
        reinterpret_cast<ParcelRec *>(a.par)[j].kmax = -1;
        T *cape = static_cast<T *>(a.cape), *cin = static_cast<T *>(a.cin),
          *zout = static_cast<T *>(a.zout);
        const int64_t oc = out_col(a, col);
        cape[oc] = T(0.0);
        cin[oc] = T(0.0);
        zout[oc] = ran ? T(-999999.0) : T(0.0);
        if (!ran) a.mulvl[oc] = 0;
        if (a.p_lfc) {
            static_cast<T *>(a.p_lfc)[oc] = T(NAN);
            static_cast<T *>(a.p_el)[oc] = T(NAN);
        }
    }
    if constexpr (SUR) {
        // every lane of the block takes part in its wave's evaluation (lanes
        // without a first sub-step give zero inputs and drop the result)
        extern __shared__ uint32_t dyn_lds[];
        const float r = surrogate_eval_wave(static_cast<const SurrogateDev *>(a.surrogate), f0,
                                            f1, f2,
                                            dyn_lds + (threadIdx.x >> 6) * MLP_LDS_WORDS_PER_WAVE);
        if (go) {
            float4 *r4 = reinterpret_cast<float4 *>(reinterpret_cast<ParcelRec *>(a.par) + j);
            r4[1] = make_float4(pc.qv2, pc.b2, pc.zk, lp1 != T(0.0) ? pc.th2 + r * lp1 : T(NAN));
        }
    }
    if constexpr (STATS) block_stats_add(LaneStats{}, ran, a.stats);
}

// The ascent's columns are split into `parts` contiguous parts with a counter
// each, parts = min(8, grid size), and the grid is a multiple of it.
// Workgroups are dispatched to the eight XCDs in turn, so the waves of part
// blockIdx % 8 all run on one XCD.  (Every part holds an equal share of the
// columns and as many waves, so the parts drain together.)  A wave takes its
// part's columns 64 at a time (one atomic per 64 columns); its lanes draw
// from that block as they finish their columns.
constexpr uint32_t WORK_PARTS = 8;
constexpr uint32_t WORK_BLOCK = 64;

__device__ __forceinline__ uint32_t work_part_begin(uint32_t ncol, uint32_t part,
                                                    uint32_t parts)
{
    return (uint32_t)(((uint64_t)ncol * part) / parts);
}

__host__ __device__ __forceinline__ uint32_t work_parts(uint32_t grid)
{
    return grid < WORK_PARTS ? grid : WORK_PARTS;
}

// PREDICT: the warm start (false: the same loop with every sub-step cold).
// LEVELS: also the LFC and EL.
// Held to 72 VGPRs, 7 waves per SIMD: the ascent is bound by VALU issue (4
// waves per SIMD are within 6 % of 8), and at 64 VGPRs the compiler spilled
// scalar registers to VGPR lanes and shuffled more (7 waves ran 1.5 % faster).
template <bool ICE, bool PSEUDO, bool PREDICT, bool STATS, bool LEVELS = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 7)))
warm_ascent_kernel(CapeGridArgs a)
{
    using T = float;
    const uint32_t ncol = (uint32_t)a.ncol;
    T *cape = static_cast<T *>(a.cape), *cin = static_cast<T *>(a.cin),
      *zout = static_cast<T *>(a.zout);
    const float4 *recs = reinterpret_cast<const float4 *>(a.par);
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    WarmAscent<T, XC_MATH, ICE, PSEUDO, LEVELS> w;
    LaneStats st;
    uint32_t col = 0;
    bool have = false, need = true;
    // wave-uniform: the wave's part (its XCD's), its current block of columns
    // [blk, blk + 64) and the block's slots handed out
    const uint32_t parts = work_parts(gridDim.x);
    const uint32_t part = blockIdx.x % parts;
    const uint32_t part_end = work_part_begin(ncol, part + 1, parts);
    uint32_t blk = 0, used = WORK_BLOCK;
    bool drained = false;
    for (;;) {
        const uint64_t want = __ballot(need);
        if (want != 0 && !drained) {
            if (used >= WORK_BLOCK) {
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(a.work + part, WORK_BLOCK);
                blk = __builtin_amdgcn_readfirstlane(b);
                used = 0;
                drained = blk >= part_end;
            }
            if (!drained) {
                uint32_t avail = WORK_BLOCK - used;
                if (blk + used + avail > part_end) avail = part_end - blk - used;
                const uint32_t rank = (uint32_t)__popcll(want & below);
                if (need && rank < avail) {
                    col = blk + used + rank;
                    need = false;
                    const float4 r0 = recs[4 * col], r1 = recs[4 * col + 1],
                                 r2 = recs[4 * col + 2], r3 = recs[4 * col + 3];
                    const int kmax = __float_as_int(r0.x);
                    if (kmax >= 0) {
                        Parcel<T> pc;
                        pc.kmax = kmax;
                        pc.mulvl = 0;
                        pc.th2 = r0.z;
                        pc.t2 = r0.w;
                        pc.qv2 = r1.x;
                        pc.b2 = r1.y;
                        pc.zk = r1.z;
                        pc.e.p = r2.x;
                        pc.e.pi = r2.y;
                        pc.e.thv = r2.z;
                        // the setup pass queued only columns with a level
                        // above the parcel: init() succeeds
                        have = w.init(pc, __float_as_int(r0.y), r1.w);
                        w.set_next(level_rec_of(r3),
                                   LEVELS ? a.env_p[(size_t)(kmax + 1) * ncol + col] : 0.0f);
                    }
                    need = !have;
                }
                const uint32_t got = (uint32_t)__popcll(want);
                used = avail == 0 ? WORK_BLOCK : used + (got < avail ? got : avail);
            }
        }
        // the wave ends when every lane has run out of columns
        if (drained) need = false;
        if (__ballot(have || need) == 0) break;
        st.blocks = 0;
        if (have) {
            const EnvCol ec{a.env, a.env_p, col, ncol};
            const bool done = w.trip(ec, T(a.pinc), T(a.fallback_tol), st, PREDICT);
            if constexpr (STATS) {
#pragma unroll
                for (int b = 0; b < 6; ++b) st.blk_lanes[b] += (st.blocks >> b) & 1u;
            }
            if (done) {
                ColumnOut<T> o;
                w.result(o);
                const int64_t oc = out_col(a, col);
                cape[oc] = o.cape;
                cin[oc] = o.cin;
                zout[oc] = o.zout;
                if constexpr (LEVELS) {
                    static_cast<float *>(a.p_lfc)[oc] = o.p_lfc;
                    static_cast<float *>(a.p_el)[oc] = o.p_el;
                }
                have = false;
                need = true;
            }
        }
        if constexpr (STATS) {
            // the wave's view of the trip: the blocks some lane ran
            uint32_t any = 0;
#pragma unroll
            for (int b = 0; b < 6; ++b) any |= (__ballot((st.blocks >> b) & 1u) != 0) << b;
            if (lane == 0) {
                st.wave_trips++;
#pragma unroll
                for (int b = 0; b < 6; ++b) st.blk_waves[b] += (any >> b) & 1u;
            }
        }
    }
    if constexpr (STATS) block_stats_add(st, false, a.stats);
}

__global__ void warm_reset_kernel(unsigned int *work, uint32_t ncol, uint32_t grid)
{
    const uint32_t parts = work_parts(grid);
    if (threadIdx.x < parts) work[threadIdx.x] = work_part_begin(ncol, threadIdx.x, parts);
}

// Storm-relative helicity chain, one column per lane.  TB: the storm-motion
// precision (float: the reference's; double: the FP64 re-run).  TI: the
// input element type (double: float64 profiles, which feed the FP64 heights
// and helicity unrounded as the reference's f2py wrappers pass them).
//
// The float chain is held to 6 waves per SIMD (80 VGPRs, 86 unconstrained):
// in the combined call it shares the CUs with the ascent, and each wave per
// SIMD it gained took the step about 1 % faster (the helicity-only call 3 %).
// The FP64 re-run (TB = double) is left unconstrained.
template <typename TB, typename TI>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(TB) == 4 ? 6 : 1, sizeof(TB) == 4 ? 6 : 8)))
srh_grid_kernel(CapeGridArgs a)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    SigmaCol<uint64_t, TI> c{};
    int nk = 0;
    if (!column_at(a, col, c, nk)) return;
    c.us = in_arr<TI>(a.us)[col];
    c.vs = in_arr<TI>(a.vs)[col];
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

template <typename TB>
hipError_t launch_srh(const CapeGridArgs &a, dim3 grid, int block, hipStream_t s)
{
    if (a.in_f64)
        srh_grid_kernel<TB, double><<<grid, block, 0, s>>>(a);
    else
        srh_grid_kernel<TB, float><<<grid, block, 0, s>>>(a);
    return hipGetLastError();
}

template <typename T, typename IDX, bool STATS, bool LEVELS>
hipError_t launch_cape_s(const CapeGridArgs &a, int block, hipStream_t s)
{
    dim3 grid((unsigned)((a.ncol + block - 1) / block));
    switch (a.adiabat) {
    case 1: cape_grid_kernel<T, IDX, false, true, STATS, LEVELS><<<grid, block, 0, s>>>(a); break;
    case 2: cape_grid_kernel<T, IDX, false, false, STATS, LEVELS><<<grid, block, 0, s>>>(a); break;
    case 3: cape_grid_kernel<T, IDX, true, true, STATS, LEVELS><<<grid, block, 0, s>>>(a); break;
    case 4: cape_grid_kernel<T, IDX, true, false, STATS, LEVELS><<<grid, block, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// 32-bit element offsets when every index of the profile arrays fits (the
// stats launches, which only count work, always use 64-bit ones).  The LFC/EL
// outputs come from float, uncounted launches.
template <typename T>
hipError_t launch_cape(const CapeGridArgs &a, int block, hipStream_t s)
{
    if (a.p_lfc && a.stats) return hipErrorInvalidValue;
    if (a.stats) return launch_cape_s<T, uint64_t, true, false>(a, block, s);
    const bool fits = (uint64_t)a.ncol * (uint64_t)a.nk_in < ((uint64_t)1 << 32);
    if (a.p_lfc)
        return fits ? launch_cape_s<T, uint32_t, false, true>(a, block, s)
                    : launch_cape_s<T, uint64_t, false, true>(a, block, s);
    return fits ? launch_cape_s<T, uint32_t, false, false>(a, block, s)
                : launch_cape_s<T, uint64_t, false, false>(a, block, s);
}

// Blocks of the persistent ascent: as many as the device holds at once, never
// more than the columns need (occupancy queried once per instantiation and
// block size)
template <bool ICE, bool PSEUDO, bool PREDICT, bool STATS, bool LEVELS>
int ascent_grid(int block, int64_t ncol)
{
    static int per_cu[5] = {-1, -1, -1, -1, -1};   // by block / 64
    static int ncu = 0;
    const int slot = block / 64;
    if (slot < 1 || slot > 4) return 1;
    if (per_cu[slot] < 0) {
        int dev = 0, pc = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &pc, reinterpret_cast<const void *>(
                         warm_ascent_kernel<ICE, PSEUDO, PREDICT, STATS, LEVELS>),
                block, 0) != hipSuccess ||
            pc < 1)
            pc = 1;
        per_cu[slot] = pc;
    }
    const int64_t need = (ncol + block - 1) / block;
    const int64_t full = (int64_t)(ncu > 0 ? ncu : 1) * per_cu[slot];
    int64_t g = need < full ? need : full;
    // a multiple of the work parts, so that every part has as many waves
    if (g >= (int64_t)WORK_PARTS) g -= g % WORK_PARTS;
    return (int)g;
}

template <bool ICE, bool PSEUDO, bool SUR, bool PREDICT, bool STATS, bool LEVELS>
hipError_t launch_warm_s(const CapeGridArgs &a, int block, hipStream_t s)
{
    const dim3 sgrid((unsigned)((a.ncol + 255) / 256));
    const size_t lds = SUR ? sizeof(uint32_t) * MLP_LDS_WORDS_PER_WAVE * 4 : 0;
    hipError_t e = hipSuccess;
    if (a.warm_phase != 2) {
        warm_setup_kernel<ICE, PSEUDO, SUR, STATS><<<sgrid, 256, lds, s>>>(a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (a.setup_done &&
            (e = hipEventRecord(static_cast<hipEvent_t>(a.setup_done), s)) != hipSuccess)
            return e;
        if (a.warm_phase == 1) return hipSuccess;
    }
    const int grid = ascent_grid<ICE, PSEUDO, PREDICT, STATS, LEVELS>(block, a.ncol);
    // the counters start at their parts' first columns, for the grid's parts
    warm_reset_kernel<<<1, 64, 0, s>>>(a.work, (uint32_t)a.ncol, (uint32_t)grid);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    warm_ascent_kernel<ICE, PSEUDO, PREDICT, STATS, LEVELS><<<grid, block, 0, s>>>(a);
    return hipGetLastError();
}

template <bool ICE, bool PSEUDO>
hipError_t launch_warm_a(const CapeGridArgs &a, int block, hipStream_t s)
{
    const bool sur = a.warm == 1 && a.surrogate;
    const bool predict = a.warm == 1;
    if (a.stats) {
        if (a.p_lfc) return hipErrorInvalidValue;
        if (sur) return launch_warm_s<ICE, PSEUDO, true, true, true, false>(a, block, s);
        return predict ? launch_warm_s<ICE, PSEUDO, false, true, true, false>(a, block, s)
                       : launch_warm_s<ICE, PSEUDO, false, false, true, false>(a, block, s);
    }
    if (a.p_lfc) {
        if (sur) return launch_warm_s<ICE, PSEUDO, true, true, false, true>(a, block, s);
        return predict ? launch_warm_s<ICE, PSEUDO, false, true, false, true>(a, block, s)
                       : launch_warm_s<ICE, PSEUDO, false, false, false, true>(a, block, s);
    }
    if (sur) return launch_warm_s<ICE, PSEUDO, true, true, false, false>(a, block, s);
    return predict ? launch_warm_s<ICE, PSEUDO, false, true, false, false>(a, block, s)
                   : launch_warm_s<ICE, PSEUDO, false, false, false, false>(a, block, s);
}

inline hipError_t launch_warm(const CapeGridArgs &a, int block, hipStream_t s)
{
    if (!a.work || !a.env || !a.par || (a.p_lfc && !a.env_p)) return hipErrorInvalidValue;
    switch (a.adiabat) {
    case 1: return launch_warm_a<false, true>(a, block, s);
    case 2: return launch_warm_a<false, false>(a, block, s);
    case 3: return launch_warm_a<true, true>(a, block, s);
    case 4: return launch_warm_a<true, false>(a, block, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_t(const CapeGridArgs &a, int kind, int block, hipStream_t s)
{
    if (a.ncol <= 0) return hipSuccess;
    dim3 grid((unsigned)((a.ncol + block - 1) / block));
    if constexpr (sizeof(T) == 8) {
        // the FP64 yardstick: CAPE with the reference's cold start, and the
        // helicity chain with FP64 storm motion
        if (kind == GRID_SRH) return launch_srh<double>(a, grid, block, s);
        if (kind != GRID_CAPE || a.warm || a.in_f64) return hipErrorInvalidValue;
        return launch_cape<T>(a, block, s);
    } else {
        switch (kind) {
        case GRID_CAPE:
            if (a.in_f64) return hipErrorInvalidValue;
            return a.warm == 0 ? launch_cape<T>(a, block, s) : launch_warm(a, block, s);
        case GRID_SRH: return launch_srh<float>(a, grid, block, s);
        default: return hipErrorInvalidValue;
        }
    }
}

}  // namespace XC_NS
}  // namespace xcape
