/*
 * CPU oracle for the column parcel integration -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's per-column algorithms, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product library (libxcape_hip.so) never links or calls it.
 *
 * Every routine follows the reference's operation order so that the FP32
 * build agrees with the reference Fortran (compiled from source into
 * oracle/_ref/) bit for bit wherever the C library's expf/logf/powf agree
 * with the Fortran runtime's.  The FP64 variants evaluate the same algorithm
 * in double precision with double constants; they serve as the accuracy
 * yardstick for the FP32 results.
 *
 * Layout: column-major profiles as in the reference (level fastest):
 *   p/t/td[col * nk_in + k], k = 0 .. nk_in-1, surface values separate.
 * Units at this boundary are the reference's: hPa and degC.
 */
#ifndef XCAPE_ORACLE_H
#define XCAPE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-column instrumentation (accumulated over the columns of one call). */
typedef struct xo_stats {
    int64_t columns;      /* columns that ran the ascent (ts > 0)          */
    int64_t levels;       /* ascent levels processed                         */
    int64_t substeps;     /* pinc sub-steps                                   */
    int64_t iterations;   /* fixed-point iterations                           */
    int64_t iterations_ice; /* of which with the ice fraction evaluated       */
    int64_t nonconverged; /* columns that hit the 100-iteration limit         */
    int64_t mu_levels;    /* levels scanned by the most-unstable search       */
    int64_t env_levels;   /* levels of environment setup                      */
} xo_stats;

/* CAPE_CODE_model_lev.f90:4-91 (loopcape_ml).  start_lev == NULL for sigma
 * grids.  For the pressure grid (loopcape_pl1d, CAPE_CODE_pressure_lev.f90:
 * 88-169) pass p1d != NULL (shared nk_in levels) and the 1-based start level
 * per column; p is then ignored.  p per column with start_lev != NULL is
 * loopcape_pl (CAPE_CODE_pressure_lev.f90:4-84): each column from its start
 * level of its own p. */
void xo_cape_f32(const float *p, const float *p1d, const float *t, const float *td,
                 const float *ps, const float *ts, const float *tds,
                 const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                 float pinc, int source, float ml_depth, int adiabat,
                 float *cape, float *cin, int32_t *mulvl, float *zout,
                 xo_stats *stats);

/* The same with the LFC and EL pressures [hPa] of each column (NaN without
 * positive area; no reference counterpart: defined from the reference's own
 * crossing pressure, CAPE_CODE_model_lev.f90:511,526, see cape_column.inc). */
void xo_cape_levels_f32(const float *p, const float *p1d, const float *t, const float *td,
                        const float *ps, const float *ts, const float *tds,
                        const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                        float pinc, int source, float ml_depth, int adiabat,
                        float *cape, float *cin, int32_t *mulvl, float *zout,
                        float *plfc, float *pel, xo_stats *stats);
void xo_cape_levels_f64(const double *p, const double *p1d, const double *t, const double *td,
                        const double *ps, const double *ts, const double *tds,
                        const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                        double pinc, int source, double ml_depth, int adiabat,
                        double *cape, double *cin, int32_t *mulvl, double *zout,
                        double *plfc, double *pel, xo_stats *stats);

void xo_cape_f64(const double *p, const double *p1d, const double *t, const double *td,
                 const double *ps, const double *ts, const double *tds,
                 const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                 double pinc, int source, double ml_depth, int adiabat,
                 double *cape, double *cin, int32_t *mulvl, double *zout,
                 xo_stats *stats);

/* xo_cape_f32 with every exp/log/pow result kept or moved one ulp up or
 * down, pseudo-randomly per argument and seed (xo_set_perturbation; one
 * process-wide seed, not for concurrent use): a stand-in for a C library
 * whose functions differ from this one's by up to an ulp.  With
 * xo_set_iterate_perturbation(amp > 0), also every converged sub-step
 * accepts another point of its tolerance band: the last iterate moved by up
 * to amp [K] (pseudo-randomly per iterate and seed), evaluated, and kept
 * when its residual passes the same 2e-4 K test -- the freedom any solver
 * of the reference's test has, a warm-started one included. */
void xo_set_perturbation(uint32_t seed);
void xo_set_iterate_perturbation(float amp);
void xo_cape_pert32(const float *p, const float *p1d, const float *t, const float *td,
                    const float *ps, const float *ts, const float *tds,
                    const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                    float pinc, int source, float ml_depth, int adiabat,
                    float *cape, float *cin, int32_t *mulvl, float *zout,
                    xo_stats *stats);

/* Per-sample log of the moist-adiabat solve: one record per sub-step,
 * (p2 [Pa], theta-e of the parcel at the start of the sub-step, qt, th1,
 *  converged th2, iterations).  Returns the number of records written
 * (<= cap).  Same column semantics as xo_cape_f32 (sigma grid only). */
int64_t xo_cape_log_f32(const float *p, const float *t, const float *td,
                        const float *ps, const float *ts, const float *tds,
                        int64_t nk_in, int64_t ncol, float pinc, int source,
                        float ml_depth, int adiabat, float *rec, int64_t cap);

/* Bolton helpers exactly as the reference computes them
 * (CAPE_CODE_model_lev.f90:570-620). */
float xo_getqvs_f32(float p, float t);
float xo_getqvi_f32(float p, float t);
float xo_getthe_f32(float p, float t, float td, float q);

/* stdheight_2D_model_lev.f90:74-160 / stdheight_2D_pressure_lev.f90:133-219.
 * FP64.  h[col*nk_in + k]; for the pressure grid levels below start are set
 * to -999999. */
void xo_stdheight(const double *p, const double *p1d, const double *t, const double *td,
                  const double *ps, const double *ts, const double *tds,
                  const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                  double aglh0, double *h, double *hs);

/* Bunkers_model_lev.f90:4-236 (FP32), with zero-initialised shear
 * accumulators.  aglh is the FP64 height cast to FP32 as f2py does.
 * rm/lm/mean are (ncol, 2) row-major: [col*2+0] = u, [col*2+1] = v. */
void xo_bunkers(const float *u, const float *v, const float *aglh,
                const float *us, const float *vs, const float *aglhs,
                const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                float *rm, float *lm, float *mean6);

/* SREH_model_lev.f90:4-129 (FP64). */
void xo_sreh(const double *u, const double *v, const double *aglh,
             const double *us, const double *vs, const double *aglhs,
             const double *cu_rm, const double *cv_rm,
             const double *cu_lm, const double *cv_lm, double top,
             const int32_t *start_lev, int64_t nk_in, int64_t ncol,
             double *srh_rm, double *srh_lm);

#ifdef __cplusplus
}
#endif
#endif
