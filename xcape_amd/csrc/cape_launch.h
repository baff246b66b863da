// Host-side launch descriptors shared by the C-ABI and the kernel translation
// units.  All pointers are device pointers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xcape {

// Device counters one launch accumulates (one atomic per block per counter).
enum { ST_COLUMNS = 0, ST_LEVELS, ST_SUBSTEPS, ST_ITERATIONS, ST_NONCONVERGED,
       ST_FALLBACKS, ST_FIRST, ST_WAVE_TRIPS, ST_BLK_LANES, ST_BLK_WAVES = ST_BLK_LANES + 6,
       ST_USED = ST_BLK_WAVES + 6, ST_COUNT = 24 };

struct CapeGridArgs {
    const float *p;        // (ncol, nk_in) sigma grid, or nullptr
    const float *p1d;      // (nk_in,) shared pressure levels, or nullptr
    const int32_t *start;  // (ncol,) 1-based first level (pressure grid), or nullptr
    const float *t, *td;   // (ncol, nk_in)
    const float *ps, *ts, *tds;  // (ncol,)
    int64_t nk_in, ncol;
    // ragged batches staged as a grid (the warm-started kernels only): each
    // column's level count (<= nk_in) and source, and the output slot of each
    // column (its caller's order); nullptr: uniform / identity
    const int32_t *nk_col, *src_col, *out_index;
    int64_t ls;                  // level stride of the (nk_in, ncol) arrays: 1 or ncol
    int in_f64;                  // the profile and surface arrays (p, p1d, t, td, u, v,
                                 // ps, ts, tds, us, vs) hold doubles (helicity only)
    float pinc, ml_depth;
    int source, adiabat;
    void *cape, *cin, *zout;     // float or double (ncol,)
    int32_t *mulvl;              // (ncol,)
    void *p_lfc, *p_el;          // (ncol,) LFC and EL [hPa] (T, as cape), or nullptr
    unsigned long long *stats;   // ST_COUNT counters or nullptr
    // storm-relative helicity in the same pass (fused kernel) or alone
    const float *u, *v;          // (ncol, nk_in)
    const float *us, *vs;        // (ncol,)
    double depth;
    double *srh_rm, *srh_lm;     // (ncol,)
    void *rm, *lm, *mean6;       // (ncol, 2) u, v pairs (float; double in the FP64
                                 // re-run), or nullptr
    // warm start of the moist-adiabat iteration
    int warm;                    // 0: reference loops, cold start; 1: warm-started
                                 // loop; 2: the same loop, every sub-step cold
    float fallback_tol;          // first-iteration residual that rejects a guess [K]
    const void *surrogate;       // device MLP weights (or nullptr)
    // scratch of the warm-started ascent (warm != 0): every level from each
    // parcel's up as a 16-byte record (dp, pi, thv, nloop | stop), (nk_in + 1,
    // ncol) level-major (level 0 the surface; (nk_in + 1) * ncol < 2^32);
    // for the LFC/EL calls the levels' pressures in a plane of the same
    // shape; one 64-byte parcel record per column (ParcelRec); and the
    // ascent's column counters
    float *env;                  // 4 floats per record
    float *env_p;                // 1 float per level (LFC/EL calls), or nullptr
    float *par;                  // 16 floats per column
    unsigned int *work;          // 8 counters
    // host side: an event (hipEvent_t) the warm-started launch records on its
    // stream after the setup pass, or nullptr
    void *setup_done;
    // host side: the warm-started launch's phases, 0 both, 1 the setup pass
    // only, 2 the counter reset and ascent only
    int warm_phase;
    // a setup pass over part of the records' columns: the records' row length
    // (0: ncol) and the part's first column in it (parcel records too)
    int64_t rec_ls, rec_col0;
};

// floats of warm-start scratch per column (env, env_p when `levels`, par)
inline int64_t warm_scratch_floats(int64_t nk_in, bool levels)
{
    return 4 * (nk_in + 1) + (levels ? nk_in + 1 : 0) + 16;
}

// kind of work a grid launch does
enum { GRID_CAPE = 1, GRID_SRH = 2, GRID_CAPE_SRH = 3 };

// (nk, ncol) level-major copies of column-contiguous (ncol, nk) arrays:
// dst[j][i] = src[i][j] for up to 5 arrays, LDS-tiled so that both the
// reads and the writes are coalesced.
hipError_t launch_transpose_levels(const float *const *src, float *const *dst, int narr,
                                   int64_t ncol, int64_t nk, hipStream_t s, int64_t ld = -1);

hipError_t launch_cape_grid_accurate(const CapeGridArgs &a, bool fp64, int kind, int block,
                                     hipStream_t s);
hipError_t launch_cape_grid_fast(const CapeGridArgs &a, int kind, int block, hipStream_t s);

}  // namespace xcape
