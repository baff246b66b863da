/*
 * xcape-hip: C-ABI of the gfx950 column parcel integrator.
 *
 * Plain pointers and sizes; no C++ or torch types.  Every entry point returns
 * 0 on success or a negative XCAPE_E* code; xcape_last_error() then gives a
 * message (thread-local).  Calls are reentrant: concurrent calls from several
 * host threads run on per-thread HIP streams unless a stream is passed.
 *
 * Array layout is the reference's: a profile array of nk levels for ncol
 * columns is the Fortran-order (nk, ncol) array of the reference's f2py
 * routines, i.e. C-order (ncol, nk) with the levels of a column contiguous.
 * Units are the reference's: hPa and degC in, J/kg and m out.
 *
 * Pointers are host pointers unless XCAPE_DEVICE_PTRS is set, in which case
 * every array argument is a device pointer and the call is asynchronous on
 * `stream` (a hipStream_t; NULL is HIP's default stream) unless `stats` is
 * non-NULL (the counters are then read back synchronously).  Host-pointer
 * calls are synchronous; with stream NULL they run on a stream of the calling
 * thread's own.  Some calls also queue work on a second stream of the calling
 * thread (the helicity launches of the combined call, and the staging copies
 * of a warm-started or helicity-only grid of many column-contiguous columns),
 * forked after the work already queued on `stream` and joined back into
 * `stream` before the call returns, so a call stays ordered as one unit.
 */
#ifndef XCAPE_HIP_H
#define XCAPE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define XCAPE_OK 0
#define XCAPE_EINVAL (-1)   /* bad argument (message says which)            */
#define XCAPE_EHIP (-2)     /* HIP runtime error                            */
#define XCAPE_ENOMEM (-3)   /* device allocation failed                     */
#define XCAPE_ENODEV (-4)   /* no usable gfx950 device                      */
#define XCAPE_ESTATE (-5)   /* entry point needs state that is not set up   */

/* ---- flags ---------------------------------------------------------------- */
#define XCAPE_DEVICE_PTRS 0x1u  /* array arguments are device pointers        */
#define XCAPE_FAST_MATH 0x2u    /* hardware exp2/log2/rcp variant (a few ulp)  */
#define XCAPE_WARM_START 0x4u   /* warm-started moist-adiabat solve: each
                                   sub-step's first iterate predicted from the
                                   sub-steps already solved (and, with a loaded
                                   surrogate, from the MLP); same fixed point,
                                   reference tolerance, cold-start fallback   */

#define XCAPE_NO_SURROGATE 0x8u /* with XCAPE_WARM_START: ignore a loaded
                                   surrogate (extrapolation only)              */
#define XCAPE_NO_PREDICTION 0x20u /* with XCAPE_WARM_START: run the warm
                                   start's loop with every sub-step started
                                   cold (no prediction, no surrogate); for
                                   measuring what the prediction itself gains */
#define XCAPE_INPUT_F64 0x40u  /* helicity entry points only: the profile and
                                   surface arrays (p, t, td, u, v, ps, ts, tds,
                                   us, vs) are double.  Heights and helicity
                                   then use them unrounded, Bunkers their
                                   float32 roundings, as the reference's f2py
                                   wrappers pass float64 arrays to its
                                   DOUBLE PRECISION and REAL routines */
#define XCAPE_LEVEL_MAJOR 0x10u /* level-major profiles: the (nk, ncol) arrays
                                   (p on a sigma grid, t, td, u, v) hold level
                                   j of column i at [j * ncol + i] instead of
                                   the default column-contiguous [i * nk + j] */

/* enums as the reference's Python layer maps them (core.py:302-311) */
#define XCAPE_SOURCE_SURFACE 1
#define XCAPE_SOURCE_MOST_UNSTABLE 2
#define XCAPE_SOURCE_MIXED_LAYER 3
#define XCAPE_ADIABAT_PSEUDO_LIQUID 1
#define XCAPE_ADIABAT_REVERSIBLE_LIQUID 2
#define XCAPE_ADIABAT_PSEUDO_ICE 3
#define XCAPE_ADIABAT_REVERSIBLE_ICE 4

/* Counters accumulated over one call. */
typedef struct xcape_stats {
    int64_t columns;       /* columns that ran the ascent (ts > 0 degC)        */
    int64_t levels;        /* ascent levels                                     */
    int64_t substeps;      /* pinc sub-steps                                    */
    int64_t iterations;    /* moist-adiabat fixed-point iterations              */
    int64_t nonconverged;  /* columns that hit the 100-iteration limit          */
    int64_t fallbacks;     /* warm starts rejected for the cold start           */
    int64_t first_substep_iterations; /* iterations of the ascents' first
                                         sub-steps (where no history exists) */
    int64_t reserved[1];
    /* The warm start's loop (XCAPE_WARM_START; zero for the reference's
     * loops).  Every trip of a lane runs some of six blocks: 0 level set-up,
     * 1 sub-step set-up, 2 first and 3 second iteration of the trip,
     * 4 sub-step completion, 5 layer update.  A wavefront runs a block in a
     * trip when any of its lanes needs it, so block_lanes[b] /
     * (64 block_waves[b]) is the block's lane utilisation. */
    int64_t wave_trips;       /* trips of the wavefronts                       */
    int64_t block_lanes[6];   /* lanes that ran block b                        */
    int64_t block_waves[6];   /* wave trips in which some lane ran block b     */
} xcape_stats;

const char *xcape_version(void);
const char *xcape_last_error(void);

/* Number of visible devices (0 when none); never fails. */
int xcape_device_count(void);

/*
 * Host-array calls (no XCAPE_DEVICE_PTRS, column-contiguous arrays) stream
 * their columns in chunks of about 128 MB of input (XCAPE_HOST_CHUNK_MB)
 * through pinned host memory and device buffers that stay allocated for the
 * next call: at most two such staging sets per process (three slots of
 * pinned input, pinned output and device buffers each), so at most two
 * host-array calls stream at once and a third waits for one of them.  This
 * frees the sets no call is using; it returns how many it freed.  No
 * counterpart in the reference (whose f2py routines take host arrays only
 * and hold no memory between calls).
 */
int xcape_release_host_staging(void);

/*
 * Column CAPE/CIN, the most-unstable level and the height of the last ascent
 * level.  Replaces the f2py routines loopcape_ml
 * (reference src/xcape/fortran/CAPE_CODE_model_lev.pyf:6-24 /
 *  CAPE_CODE_model_lev.f90:4-91) and loopcape_pl1d
 * (CAPE_CODE_pressure_lev.pyf:26-45 / CAPE_CODE_pressure_lev.f90:88-169),
 * as called by cape_fortran.cape (src/xcape/cape_fortran.py:49-67).
 *
 *   p_shared == 0: p is (ncol, nk) model-level pressure (loopcape_ml);
 *                  with start_lev != NULL, pressure per column on a
 *                  pressure grid, each column from its 1-based start level
 *                  (loopcape_pl, CAPE_CODE_pressure_lev.f90:4-84, which the
 *                  reference's Python never calls; the helicity entry points
 *                  reject this combination);
 *   p_shared == 1: p is (nk,) pressure levels shared by all columns
 *                  (loopcape_pl1d); start_lev (ncol,) gives each column's
 *                  1-based first level (the reference's pres_lev_pos); when
 *                  NULL it is computed on the device with the same rule
 *                  (core.py:286-289).
 * t, td: (ncol, nk); ps, ts, tds: (ncol,).
 * Outputs (ncol,): cape, cin [J/kg], mulev (1-based, surface = 1, -999999 if
 * not most-unstable), zmulev [m].  Columns with ts <= 0 degC give zeros.
 * Accuracy: CAPE and CIN within max(0.5 J/kg, 0.1 %) of the reference's
 * FP32 results, with one exception: CIN where the reference's own CIN is
 * ill-conditioned.  getcape commits the accumulated CIN when the parcel
 * first enters positive buoyancy (CAPE_CODE_model_lev.f90:509-523); where a
 * layer's buoyancy is within rounding of zero, an ulp of exp/log/pow or the
 * iterate accepted inside the 2e-4 K band decides whether it is committed,
 * and CIN may then differ by up to the whole negative area (CAPE does not).
 */
int xcape_cape_f32(const float *p, int p_shared, const int32_t *start_lev,
                   const float *t, const float *td,
                   const float *ps, const float *ts, const float *tds,
                   int64_t nk, int64_t ncol, float pinc, int source, float ml_depth,
                   int adiabat, float *cape, float *cin, int32_t *mulev, float *zmulev,
                   uint32_t flags, xcape_stats *stats, void *stream);

/*
 * xcape_cape_f32 for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): the
 * same arguments, computation and results, with device pointers only
 * (XCAPE_DEVICE_PTRS required) and no stats (stats must be NULL).  It calls
 * no hipMalloc, hipFree or other allocating routine: the staged copies of
 * column-contiguous profiles and the warm start's records live in a scratch
 * array of 256 MiB that the library holds on each device it runs on, from
 * its first call there until the process ends.  Scratch per column: 12 nk
 * bytes to stage column-contiguous sigma profiles (8 nk on shared pressure
 * levels, none for level-major input), plus 16 nk + 80 bytes of warm-start
 * records with XCAPE_WARM_START; e.g. 3916 bytes for 137 sigma levels with
 * the warm start, so at most 68 548 such columns per call.  A call needing
 * more than 256 MiB fails with XCAPE_EINVAL and a message giving its limit.
 * Calls on one device share the array: calls that overlap in time on one
 * device race and must be ordered by the caller (one stream, or a wait
 * between them).  No counterpart in the reference.
 */
int xcape_cape_noalloc_f32(const float *p, int p_shared, const int32_t *start_lev,
                           const float *t, const float *td,
                           const float *ps, const float *ts, const float *tds,
                           int64_t nk, int64_t ncol, float pinc, int source, float ml_depth,
                           int adiabat, float *cape, float *cin, int32_t *mulev,
                           float *zmulev, uint32_t flags, xcape_stats *stats, void *stream);

/* As xcape_cape_f32, plus each column's level of free convection and
 * equilibrium level, p_lfc and p_el [hPa] (NaN for columns without positive
 * area).  No reference counterpart: getcape computes the crossing pressure
 * ps (CAPE_CODE_model_lev.f90:511,526) and drops it.  Over the layers whose
 * trapezoid adds positive area, the LFC is where the first one's positive
 * area begins (ps where the parcel enters positive buoyancy in that layer,
 * else the layer's bottom) and the EL where the last one's ends (ps where it
 * leaves positive buoyancy, else the layer's top).  Same flags as
 * xcape_cape_f32 except that stats are not collected. */
int xcape_cape_levels_f32(const float *p, int p_shared, const int32_t *start_lev,
                          const float *t, const float *td,
                          const float *ps, const float *ts, const float *tds,
                          int64_t nk, int64_t ncol, float pinc, int source, float ml_depth,
                          int adiabat, float *cape, float *cin, int32_t *mulev, float *zmulev,
                          float *p_lfc, float *p_el, uint32_t flags, void *stream);

/* Same computation evaluated in FP64 on the device from the same FP32 inputs
 * (the accuracy yardstick for the FP32 kernel).  No reference counterpart:
 * the reference computes CAPE in FP32 only. */
int xcape_cape_f64(const float *p, int p_shared, const int32_t *start_lev,
                   const float *t, const float *td,
                   const float *ps, const float *ts, const float *tds,
                   int64_t nk, int64_t ncol, double pinc, int source, double ml_depth,
                   int adiabat, double *cape, double *cin, int32_t *mulev, double *zmulev,
                   uint32_t flags, xcape_stats *stats, void *stream);

/* xcape_cape_levels_f32 evaluated in FP64, as xcape_cape_f64 (the FP64
 * re-run's LFC and EL).  No reference counterpart. */
int xcape_cape_levels_f64(const float *p, int p_shared, const int32_t *start_lev,
                          const float *t, const float *td,
                          const float *ps, const float *ts, const float *tds,
                          int64_t nk, int64_t ncol, double pinc, int source, double ml_depth,
                          int adiabat, double *cape, double *cin, int32_t *mulev,
                          double *zmulev, double *p_lfc, double *p_el, uint32_t flags,
                          void *stream);

/*
 * Storm-relative helicity 0..depth for the Bunkers right and left movers,
 * plus the storm motions and the 0-6 km mean wind.  Replaces the chain
 * calc_srh runs (src/xcape/core.py:516-541): loop_stdheight_ml / _pl1d
 * (stdheight_2D_model_lev.pyf, stdheight_2D_pressure_lev.pyf) with a 2 m
 * surface height, bunkers_loop_ml / _pl (Bunkers_model_lev.pyf:6-21,
 * Bunkers_pressure_lev.pyf) and loop_sreh_ml / _pl (SREH_model_lev.pyf,
 * SREH_pressure_lev.pyf), in the reference's precisions (FP64 heights and
 * helicity, FP32 storm motion).
 *
 * Arrays as for xcape_cape_f32, plus u, v (ncol, nk) and us, vs (ncol,) in
 * m/s; with XCAPE_INPUT_F64 the ten input pointers point at double arrays.
 * Outputs: srh_rm, srh_lm (ncol,) FP64; rm, lm, mean6 (ncol, 2) as
 * (u, v) pairs (the reference's (2, n2) Fortran arrays), all three NULL to
 * skip them.  Grids of 2^18 column-contiguous columns or more are staged in
 * four parts, each part's launch on a second stream of the calling thread
 * while the next part is copied (joined back as above).
 */
int xcape_srh_f32(const float *p, int p_shared, const int32_t *start_lev,
                  const float *t, const float *td, const float *u, const float *v,
                  const float *ps, const float *ts, const float *tds,
                  const float *us, const float *vs, int64_t nk, int64_t ncol, double depth,
                  double *srh_rm, double *srh_lm, float *rm, float *lm, float *mean6,
                  uint32_t flags, void *stream);

/* Same helicity chain with the Bunkers storm motion (and the mean wind) in
 * FP64 instead of the reference's FP32; heights and SREH are FP64 in both.
 * The accuracy yardstick for xcape_srh_f32.  No reference counterpart.
 * FAST_MATH is ignored, as in xcape_srh_f32. */
int xcape_srh_f64(const float *p, int p_shared, const int32_t *start_lev,
                  const float *t, const float *td, const float *u, const float *v,
                  const float *ps, const float *ts, const float *tds,
                  const float *us, const float *vs, int64_t nk, int64_t ncol, double depth,
                  double *srh_rm, double *srh_lm, double *rm, double *lm, double *mean6,
                  uint32_t flags, void *stream);

/*
 * CAPE/CIN (as xcape_cape_f32) and the helicity chain (as xcape_srh_f32) of
 * the same columns in one call: column-contiguous profiles are staged
 * level-major once for both, and the helicity launch runs on a second stream
 * beside the ascent (forked after the work queued on `stream`, with the warm
 * start after the setup pass, with the copies of the winds; joined back
 * before the call returns or anything later on `stream` runs).  No single
 * reference counterpart: it serves callers that need both calls above.
 */
int xcape_cape_srh_f32(const float *p, int p_shared, const int32_t *start_lev,
                       const float *t, const float *td, const float *u, const float *v,
                       const float *ps, const float *ts, const float *tds,
                       const float *us, const float *vs, int64_t nk, int64_t ncol,
                       float pinc, int source, float ml_depth, int adiabat, double depth,
                       float *cape, float *cin, int32_t *mulev, float *zmulev,
                       double *srh_rm, double *srh_lm, float *rm, float *lm, float *mean6,
                       uint32_t flags, xcape_stats *stats, void *stream);

/*
 * Ragged batches: columns with their own level count, parcel source and
 * adiabat in one persistent launch, replayed from a captured hipGraph.
 * No single reference counterpart: the reference handles one (nk, source,
 * adiabat) per loopcape_ml call (CAPE_CODE_model_lev.f90:4-91), so a mixed
 * batch is a group of such calls; per column the results are those of
 * xcape_cape_f32 on that column.
 *
 * Layout (CSR): column i's levels are offsets[i] .. offsets[i+1]-1 of the
 * concatenated p, t, td; ps, ts, tds, source, adiabat are (ncol,).  The *_host
 * arrays are host copies of offsets/source/adiabat used to validate the batch
 * and build the work order (sorted by adiabat, source and descending level
 * count); every other array is a device pointer the plan keeps using.
 *
 * chunk: columns a wavefront takes from the work queue per grab (64, 128 or
 * 256); tile: levels per column staged in LDS (0..64); cols_per_block:
 * columns per workgroup, 16, 32, 64, 128 or 256 (one column per lane; 16 and
 * 32 are one wavefront with the other lanes idle, for the tuner's sweep).
 * Warm-started plans use neither chunk nor tile, and take 64, 128 or 256
 * columns per workgroup of their persistent ascent.
 * flags: XCAPE_FAST_MATH and/or XCAPE_WARM_START, or 0.
 */
typedef struct xcape_ragged_plan xcape_ragged_plan;

int xcape_ragged_plan_create(const int64_t *offsets_host, const int32_t *source_host,
                             const int32_t *adiabat_host, int64_t ncol,
                             const int64_t *offsets, const int32_t *source,
                             const int32_t *adiabat, const float *p, const float *t,
                             const float *td, const float *ps, const float *ts,
                             const float *tds, float pinc, float ml_depth, float *cape,
                             float *cin, int32_t *mulev, float *zmulev, int chunk, int tile,
                             int cols_per_block, uint32_t flags, xcape_ragged_plan **out);

/* Run the plan on `stream`: the first call captures the launch into a
 * hipGraph, later calls replay it.  With stats != NULL the run is a direct
 * launch with counters and returns synchronously. */
int xcape_ragged_plan_run(xcape_ragged_plan *plan, xcape_stats *stats, void *stream);

/* Launch geometry the plan chose (grid blocks, threads per block). */
int xcape_ragged_plan_info(const xcape_ragged_plan *plan, int *grid, int *block);

int xcape_ragged_plan_destroy(xcape_ragged_plan *plan);

/*
 * The warm-start surrogate: a 3 -> 64 -> 64 -> 1 MLP (tanh-form GELU) with
 * BF16 weights and FP32 accumulation.  No reference counterpart (the
 * reference always starts the moist-adiabat iteration cold,
 * CAPE_CODE_model_lev.f90:431).
 *
 * Inputs, normalised by norm = {mean[3], scale[3], out_scale, out_offset}:
 * ln p2 [ln Pa], the parcel's theta-e at the sub-step's start [K] (the
 * reference's parcel formula getthe(p, t, t, qv)), total water [kg/kg].
 * Output: the sub-step rate (theta2 - theta1) / ln(p2/p1) [K].
 *
 * Parameters: an FP32 master vector of xcape_mlp_nparams() floats (W1 64x3,
 * b1, W2 64x64, b2, w3, b3).  Every pointer below is a device pointer; calls
 * are asynchronous on `stream` (NULL: default stream) unless stated.
 */
int xcape_mlp_nparams(void);

/* Network output for n normalised inputs x (n, 3) -> y (n,), VALU path. */
int xcape_mlp_forward(const float *x, int64_t n, const float *theta, float *y, void *stream);

/* Mean-squared-error loss over (x, t) and its gradient with respect to the
 * parameter vector (straight-through through the BF16 rounding). */
int xcape_mlp_loss_grad(const float *x, const float *t, int64_t n, const float *theta,
                        float *grad, float *loss, void *stream);

/* One Adam step on theta (m, v: moment vectors, step >= 1). */
int xcape_mlp_adam(float *theta, const float *grad, float *m, float *v, int64_t step, float lr,
                   float beta1, float beta2, float eps, float weight_decay, void *stream);

/* Make theta (and norm, 8 floats) the surrogate the warm-started kernels
 * use; synchronous.  xcape_surrogate_loaded() says whether one is loaded.
 * Loads are serialised, and each packs into a buffer of its own before it is
 * published: a warm-started call on another thread uses either the previous
 * network or the new one, whole (which one is unspecified while a load is in
 * progress, and iteration counts may differ between them; results stay
 * within the solver tolerance either way).  Replaced networks stay allocated
 * until the process ends; a ragged plan keeps the network that was loaded
 * when it was created. */
int xcape_surrogate_load(const float *theta, const float *norm, void *stream);
int xcape_surrogate_loaded(void);

/* The loaded surrogate on raw inputs x (n, 3) through the in-kernel path
 * (wave-wide BF16 MFMA) -> rates y (n,). */
int xcape_surrogate_eval(const float *x, int64_t n, float *y, void *stream);

/* Training data from the cold-start ascent of a sigma grid (device arrays
 * as xcape_cape_f32): for a hashed `keep_fraction` of the sub-steps (with
 * first_only != 0: of each ascent's first sub-step only), append
 * (ln p2, parcel theta-e, qt, rate, ln(p2/p1)) to samples (capacity, 5); *count
 * (device) is advanced atomically and may exceed capacity (extra records are
 * dropped). */
int xcape_surrogate_log_f32(const float *p, const float *t, const float *td,
                            const float *ps, const float *ts, const float *tds,
                            int64_t nk, int64_t ncol, float pinc, int source, float ml_depth,
                            int adiabat, float keep_fraction, uint32_t seed, int first_only,
                            float *samples, int64_t capacity, unsigned long long *count,
                            void *stream);

/*
 * The accurate mode's device arithmetic, elementwise on device arrays a, b
 * (n,) -> y (n,), for checking its rounding on the device at hand: fn 0 the
 * raw IEEE division kernel, 1 the compiler's IEEE division, 2 division by a
 * per-call constant (its reciprocal rounded once), 3 the exp kernel, 4 the
 * device library's expf, 5 the pow kernel, 6 the device library's powf.
 * b is unused (may be NULL) for fn 3 and 4.  Asynchronous on `stream`.
 * No reference counterpart.
 */
int xcape_math_check(int fn, const float *a, const float *b, int64_t n, float *y,
                     void *stream);

/*
 * The helicity chain's FP64 logarithm, exponential and division,
 * elementwise on device arrays a (n,) -> y (n,): fn 0 the log kernel of the
 * heights, 1 the device library's log, 2 the exp kernel of the heights, 3 the
 * device library's exp; fn 4 the raw division kernel and 5 IEEE division
 * take a (2n,) as the operand pairs (a[2i], a[2i+1]).
 * Asynchronous on `stream`.  No reference counterpart.
 */
int xcape_math_check_f64(int fn, const double *a, int64_t n, double *y, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* XCAPE_HIP_H */
