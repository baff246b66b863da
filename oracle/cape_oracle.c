/*
 * CPU oracle: column CAPE/CIN loops -- TEST INFRASTRUCTURE ONLY.
 *
 * xo_cape_f32 / xo_cape_f64 restate loopcape_ml (reference
 * src/xcape/fortran/CAPE_CODE_model_lev.f90:76-89) and loopcape_pl1d
 * (CAPE_CODE_pressure_lev.f90:151-167): a column with surface temperature
 * <= 0 degC returns zeros, every other column runs getcape on its levels
 * (for pressure grids, levels start_lev..nk_in of the shared profile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---- FP32 instantiation: float literals, C library float functions ---- */
#define REAL float
#define SUF _f32
#define EXPF expf
#define LOGF logf
#define POWF powf
#define L(x) x##f
#define LOGREC 0
#include "cape_column.inc"
#undef REAL
#undef SUF
#undef EXPF
#undef LOGF
#undef POWF
#undef L
#undef LOGREC

/* ---- FP64 instantiation ---- */
#define REAL double
#define SUF _f64
#define EXPF exp
#define LOGF log
#define POWF pow
#define L(x) x
#define LOGREC 0
#include "cape_column.inc"
#undef REAL
#undef SUF
#undef EXPF
#undef LOGF
#undef POWF
#undef L
#undef LOGREC

/* ---- FP32 with solver-sample logging ---- */
#define REAL float
#define SUF _log32
#define EXPF expf
#define LOGF logf
#define POWF powf
#define L(x) x##f
#define LOGREC 1
#include "cape_column.inc"
#undef REAL
#undef SUF
#undef EXPF
#undef LOGF
#undef POWF
#undef L
#undef LOGREC

/* ---- FP32 with each exp/log/pow result moved by up to one ulp ----
 * A stand-in for another implementation of the C library's functions (the
 * GPU's differ from glibc's by up to an ulp): each result is kept, or moved
 * one ulp up or down, by a hash of its argument(s) and the seed set with
 * xo_set_perturbation (process-wide; not for concurrent use).  Columns whose
 * outcome changes between seeds are those the reference decides by its
 * rounding alone. */
static uint32_t xo_pert_seed;
static float xo_pert_iter_amp;

void xo_set_perturbation(uint32_t seed) { xo_pert_seed = seed; }
void xo_set_iterate_perturbation(float amp) { xo_pert_iter_amp = amp; }

static float pert_f(float r, uint32_t a, uint32_t b, uint32_t tag)
{
    uint32_t h = (a * 2654435761u) ^ (b * 2246822519u) ^ (xo_pert_seed * 3266489917u + tag);
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    switch (h % 3u) {
    case 1: return nextafterf(r, INFINITY);
    case 2: return nextafterf(r, -INFINITY);
    default: return r;
    }
}
static uint32_t bits_f(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float pexpf(float x) { return pert_f(expf(x), bits_f(x), 0u, 1u); }
static float plogf(float x) { return pert_f(logf(x), bits_f(x), 0u, 2u); }
static float ppowf(float x, float y) { return pert_f(powf(x, y), bits_f(x), bits_f(y), 3u); }
/* a shift in [-amp, amp] by a hash of the iterate and the seed */
static float iter_delta(float x)
{
    uint32_t h = bits_f(x) * 2654435761u ^ (xo_pert_seed * 3266489917u + 4u);
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    return xo_pert_iter_amp * ((float)(h & 0xffffu) / 32767.5f - 1.0f);
}

#define REAL float
#define SUF _pert32
#define EXPF pexpf
#define LOGF plogf
#define POWF ppowf
#define L(x) x##f
#define LOGREC 0
#define ITERPERT 1
#define ITERPERT_AMP xo_pert_iter_amp
#define ITERPERT_DELTA iter_delta
#include "cape_column.inc"
#undef ITERPERT
#undef ITERPERT_AMP
#undef ITERPERT_DELTA
#undef REAL
#undef SUF
#undef EXPF
#undef LOGF
#undef POWF
#undef L
#undef LOGREC

float xo_getqvs_f32(float p, float t) { return qvs_f32(p, t); }
float xo_getqvi_f32(float p, float t) { return qvi_f32(p, t); }
float xo_getthe_f32(float p, float t, float td, float q) { return the_f32(p, t, td, q); }

#define DEFINE_LOOP(SUFX, R)                                                         \
void xo_cape_levels##SUFX(const R *p, const R *p1d, const R *t, const R *td,         \
                          const R *ps, const R *ts, const R *tds,                    \
                          const int32_t *start_lev, int64_t nk_in, int64_t ncol,     \
                          R pinc, int source, R ml_depth, int adiabat,               \
                          R *cape, R *cin, int32_t *mulvl, R *zout, R *plfc, R *pel, \
                          xo_stats *st)                                              \
{                                                                                    \
    R *work = (R *)malloc(sizeof(R) * 8 * (size_t)(nk_in + 1));                      \
    for (int64_t i = 0; i < ncol; ++i) {                                             \
        if (ts[i] > (R)0) {                                                          \
            int32_t s0 = 0;                                                          \
            const R *pc, *tc = t + i * nk_in, *tdc = td + i * nk_in;                 \
            if (p1d) {                                                               \
                s0 = start_lev[i] - 1;                                               \
                pc = p1d + s0;                                                       \
            } else {                                                                 \
                /* p per column from its start level: pressure_lev.f90:69-75   */ \
                if (start_lev) s0 = start_lev[i] - 1;                                \
                pc = p + i * nk_in + s0;                                             \
            }                                                                        \
            if (st) st->columns++;                                                   \
            getcape##SUFX((int)(nk_in - s0), pc, tc + s0, tdc + s0, ps[i], ts[i],    \
                          tds[i], pinc, source, ml_depth, adiabat, &cape[i], &cin[i], \
                          &mulvl[i], &zout[i], plfc ? &plfc[i] : NULL,               \
                          pel ? &pel[i] : NULL, work, st);                           \
        } else {                                                                     \
            cape[i] = 0; cin[i] = 0; zout[i] = 0; mulvl[i] = 0;                      \
            if (plfc) { plfc[i] = (R)NAN; pel[i] = (R)NAN; }                         \
        }                                                                            \
    }                                                                                \
    free(work);                                                                      \
}                                                                                    \
void xo_cape##SUFX(const R *p, const R *p1d, const R *t, const R *td,                \
                   const R *ps, const R *ts, const R *tds,                           \
                   const int32_t *start_lev, int64_t nk_in, int64_t ncol,            \
                   R pinc, int source, R ml_depth, int adiabat,                      \
                   R *cape, R *cin, int32_t *mulvl, R *zout, xo_stats *st)           \
{                                                                                    \
    xo_cape_levels##SUFX(p, p1d, t, td, ps, ts, tds, start_lev, nk_in, ncol, pinc,   \
                         source, ml_depth, adiabat, cape, cin, mulvl, zout, NULL,    \
                         NULL, st);                                                  \
}

DEFINE_LOOP(_f32, float)
DEFINE_LOOP(_f64, double)
DEFINE_LOOP(_pert32, float)

int64_t xo_cape_log_f32(const float *p, const float *t, const float *td,
                        const float *ps, const float *ts, const float *tds,
                        int64_t nk_in, int64_t ncol, float pinc, int source,
                        float ml_depth, int adiabat, float *rec, int64_t cap)
{
    float *work = (float *)malloc(sizeof(float) * 8 * (size_t)(nk_in + 1));
    int64_t nrec = 0;
    for (int64_t i = 0; i < ncol && nrec < cap; ++i) {
        if (!(ts[i] > 0.0f))
            continue;
        float c, ci, z;
        int32_t m;
        getcape_log32((int)nk_in, p + i * nk_in, t + i * nk_in, td + i * nk_in, ps[i],
                      ts[i], tds[i], pinc, source, ml_depth, adiabat, &c, &ci, &m, &z,
                      NULL, NULL, work, NULL, rec, &nrec, cap);
    }
    free(work);
    return nrec;
}
