/*
 * CPU oracle: standard height, Bunkers storm motion, storm-relative helicity
 * -- TEST INFRASTRUCTURE ONLY.
 *
 * Literal constants in the reference's DOUBLE PRECISION routines are default
 * (single-precision) REAL literals, so their double values are those of the
 * nearest float; SPD(x) reproduces that promotion.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

#define SPD(x) ((double)(float)(x))

/* stdheight_ml / stdheight_pl body (stdheight_2D_model_lev.f90:74-160,
 * stdheight_2D_pressure_lev.f90:297-383; identical). n levels. */
static void stdheight_col(int n, const double *P, const double *T, const double *Td,
                          double Ps, double Ts, double Tds, double Hin,
                          double *H, double *Hs, double *Tv)
{
    const double R = SPD(287.04);
    const double g = SPD(-9.80665);
    const double eps = SPD(0.6219800858985514);
    double Tin = Ts + SPD(273.15);
    double Tdin = Tds + SPD(273.15);
    double E = SPD(6.112) * exp((SPD(53.49) - (6808.0 / Tdin)) - SPD(5.09) * log(Tdin));
    double w = eps * (E / (Ps - E));
    double Tvs = Tin * ((w + eps) / (eps * (1.0 + w)));
    for (int k = 0; k < n; ++k) {
        Tin = T[k] + SPD(273.15);
        Tdin = Td[k] + SPD(273.15);
        E = SPD(6.112) * exp((SPD(53.49) - (6808.0 / Tdin)) - SPD(5.09) * log(Tdin));
        w = eps * (E / (P[k] - E));
        Tv[k] = Tin * ((w + eps) / (eps * (1.0 + w)));
    }
    *Hs = Hin;
    for (int k = 0; k < n; ++k) {
        if (k == 0)
            H[k] = *Hs + ((R * ((Tv[k] + Tvs) / 2.0)) / g) * log(P[k] / Ps);
        else
            H[k] = H[k - 1] + ((R * ((Tv[k] + Tv[k - 1]) / 2.0)) / g) * log(P[k] / P[k - 1]);
    }
}

void xo_stdheight(const double *p, const double *p1d, const double *t, const double *td,
                  const double *ps, const double *ts, const double *tds,
                  const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                  double aglh0, double *h, double *hs)
{
    double *tv = (double *)malloc(sizeof(double) * (size_t)(nk_in + 1));
    for (int64_t i = 0; i < ncol; ++i) {
        double *hc = h + i * nk_in;
        if (p1d) {
            int s0 = start_lev[i] - 1;
            for (int k = 0; k < s0; ++k)
                hc[k] = -999999.0;
            stdheight_col((int)(nk_in - s0), p1d + s0, t + i * nk_in + s0, td + i * nk_in + s0,
                          ps[i], ts[i], tds[i], aglh0, hc + s0, &hs[i], tv);
        } else {
            stdheight_col((int)nk_in, p + i * nk_in, t + i * nk_in, td + i * nk_in,
                          ps[i], ts[i], tds[i], aglh0, hc, &hs[i], tv);
        }
    }
    free(tv);
}

/* DINTERP2DZ for one column (Bunkers_model_lev.f90:188-236): top-down search
 * for the first bracket z[kp-1] <= h < z[kp] (orientation from z[0] vs
 * z[n-1]); -999999 when none.  W1 = 1 - W2 is formed in double. */
static float interp_z(const float *v, const float *z, int n, float h)
{
    int ip = 0, im = 1;
    if (z[0] > z[n - 1]) { ip = 1; im = 0; }
    for (int kp = n; kp >= 2; --kp) {      /* 1-based like the reference */
        float zlo = z[kp - im - 1], zhi = z[kp - ip - 1];
        if (zlo <= h && zhi > h) {
            float w2 = (h - zlo) / (zhi - zlo);
            float w1 = (float)(1.0 - (double)w2);
            return w1 * v[kp - im - 1] + w2 * v[kp - ip - 1];
        }
    }
    return -999999.0f;
}

/* bunkers_calc_ml (Bunkers_model_lev.f90:75-180), accumulators from zero. */
static void bunkers_col(int n, const float *U, const float *V, const float *Z,
                        float *rm, float *lm, float *mw)
{
    static const float lv[12] = {500.f, 1000.f, 1500.f, 2000.f, 2500.f, 3000.f,
                                 3500.f, 4000.f, 4500.f, 5000.f, 5500.f, 6000.f};
    float us6[13], vs6[13];
    us6[0] = U[0];
    vs6[0] = V[0];
    for (int i = 1; i < 13; ++i) {
        us6[i] = interp_z(U, Z, n, lv[i - 1]);
        vs6[i] = interp_z(V, Z, n, lv[i - 1]);
    }
    float mu = 0.0f, mv = 0.0f;
    for (int i = 0; i < 13; ++i) {
        mu = mu + us6[i];
        mv = mv + vs6[i];
    }
    mu = mu / 13.0f;
    mv = mv / 13.0f;
    float ushr_u = ((0.0f + us6[11]) + us6[12]) / 2.0f;
    float vshr_u = ((0.0f + vs6[11]) + vs6[12]) / 2.0f;
    float ushr_d = ((0.0f + us6[0]) + us6[1]) / 2.0f;
    float vshr_d = ((0.0f + vs6[0]) + vs6[1]) / 2.0f;
    float ushr = ushr_u - ushr_d;
    float vshr = vshr_u - vshr_d;
    float mag = sqrtf(ushr * ushr + vshr * vshr);
    rm[0] = mu + 7.5f * vshr / mag;
    rm[1] = mv - 7.5f * ushr / mag;
    lm[0] = mu - 7.5f * vshr / mag;
    lm[1] = mv + 7.5f * ushr / mag;
    mw[0] = mu;
    mw[1] = mv;
}

void xo_bunkers(const float *u, const float *v, const float *aglh,
                const float *us, const float *vs, const float *aglhs,
                const int32_t *start_lev, int64_t nk_in, int64_t ncol,
                float *rm, float *lm, float *mean6)
{
    size_t nall = (size_t)nk_in + 1;
    float *U = (float *)malloc(sizeof(float) * 3 * nall);
    float *V = U + nall, *Z = V + nall;
    for (int64_t i = 0; i < ncol; ++i) {
        int s0 = start_lev ? start_lev[i] - 1 : 0;
        int n = (int)(nk_in - s0) + 1;
        U[0] = us[i]; V[0] = vs[i]; Z[0] = aglhs[i];
        for (int k = 1; k < n; ++k) {
            U[k] = u[i * nk_in + s0 + k - 1];
            V[k] = v[i * nk_in + s0 + k - 1];
            Z[k] = aglh[i * nk_in + s0 + k - 1];
        }
        bunkers_col(n, U, V, Z, rm + 2 * i, lm + 2 * i, mean6 + 2 * i);
    }
    free(U);
}

/* INTERP1 (SREH_model_lev.f90:123-129); *x1 is updated like the reference's
 * by-reference argument. */
static double interp1(double y1, double y3, double *x1, double x2, double x3)
{
    if (x3 == *x1) *x1 = *x1 - SPD(0.01);
    return y1 + ((y3 - y1) * ((x2 - *x1) / (x3 - *x1)));
}

/* DCALRELHL_ml (SREH_model_lev.f90:61-118). */
static void sreh_col(int n, const double *u, double *v, double *z,
                     double cu_rm, double cv_rm, double cu_lm, double cv_lm, double top,
                     double *utop, double *vtop, double *out_rm, double *out_lm)
{
    for (int k = 0; k < n; ++k) { utop[k] = u[k]; vtop[k] = v[k]; }
    int ktop = 0;      /* 1-based, 0 = none */
    for (int k = 2; k <= n; ++k) {
        if (z[k - 1] > top && ktop == 0) {
            ktop = k;
            utop[ktop - 1] = interp1(u[ktop - 1], u[ktop - 2], &z[ktop - 1], top, z[ktop - 2]);
            vtop[ktop - 1] = interp1(v[ktop - 1], v[ktop - 2], &z[ktop - 1], top, z[ktop - 2]);
        }
    }
    double sum = 0.0;
    for (int k = 2; k <= ktop; ++k) {
        double x = ((utop[k - 1] - cu_rm) * (vtop[k - 1] - vtop[k - 2]))
                 - ((vtop[k - 1] - cv_rm) * (utop[k - 1] - utop[k - 2]));
        sum = sum + x;
    }
    *out_rm = -sum;
    sum = 0.0;
    for (int k = 2; k <= ktop; ++k) {
        double x = ((utop[k - 1] - cu_lm) * (vtop[k - 1] - vtop[k - 2]))
                 - ((vtop[k - 1] - cv_lm) * (utop[k - 1] - utop[k - 2]));
        sum = sum + x;
    }
    *out_lm = -sum;
}

void xo_sreh(const double *u, const double *v, const double *aglh,
             const double *us, const double *vs, const double *aglhs,
             const double *cu_rm, const double *cv_rm,
             const double *cu_lm, const double *cv_lm, double top,
             const int32_t *start_lev, int64_t nk_in, int64_t ncol,
             double *srh_rm, double *srh_lm)
{
    size_t nall = (size_t)nk_in + 1;
    double *U = (double *)malloc(sizeof(double) * 5 * nall);
    double *V = U + nall, *Z = V + nall, *UT = Z + nall, *VT = UT + nall;
    for (int64_t i = 0; i < ncol; ++i) {
        int s0 = start_lev ? start_lev[i] - 1 : 0;
        int n = (int)(nk_in - s0) + 1;
        U[0] = us[i]; V[0] = vs[i]; Z[0] = aglhs[i];
        for (int k = 1; k < n; ++k) {
            U[k] = u[i * nk_in + s0 + k - 1];
            V[k] = v[i * nk_in + s0 + k - 1];
            Z[k] = aglh[i * nk_in + s0 + k - 1];
        }
        sreh_col(n, U, V, Z, cu_rm[i], cv_rm[i], cu_lm[i], cv_lm[i], top, UT, VT,
                 &srh_rm[i], &srh_lm[i]);
    }
    free(U);
}
