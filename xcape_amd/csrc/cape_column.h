// One sounding column on one lane: environment, parcel source, pinc-stepped
// moist-adiabatic ascent and the CAPE/CIN trapezoids.
//
// Follows getcape_ml (reference src/xcape/fortran/CAPE_CODE_model_lev.f90
// :97-564; getcape_pl in CAPE_CODE_pressure_lev.f90:174-642 has the same body)
// with the same operation order, so that in MATH_ACCURATE mode every basic
// operation rounds as the reference's does and only the device library's
// exp/log/pow can differ (by <= 1 ulp).
//
// Differences in organisation, not in results:
//  * the reference first fills nk-long arrays p, t, td, pi, q, th, thv, z; here
//    a level's environment is evaluated when the ascent reaches it and the
//    previous level's is kept in registers (no per-column scratch, so the
//    lane's state is a few dozen VGPRs);
//  * the height z(k) is accumulated with the ascent's own layer thickness,
//    which is the same expression (:246 and :503), hence the same value;
//  * liquid-only adiabats skip the ice saturation function, which the
//    reference evaluates and multiplies by zero (:445); identical unless it
//    overflows to Inf (p equal to the ice saturation pressure).
//
// Two ascent drivers share the pieces below:
//  * cape_column      - the reference's nested level / sub-step / iteration
//                       loops with the cold start thlast = th1 (:431);
//  * cape_column_warm - the same iteration started from a predicted theta and
//                       driven by a flattened loop (one fixed-point iteration
//                       per lane per trip), see its comment.
//
// The column accessor C provides P(k), Tc(k), Tdc(k) in hPa / degC for
// k = 0 .. nk-1, k = 0 being the surface, and raw_above(k), the same three
// values of a level k >= 1 (the ascent's levels, read without the surface
// select).
#pragma once
#include "phys.h"

namespace xcape {

enum { SRC_SURFACE = 1, SRC_MU = 2, SRC_ML = 3 };

// Fixed-point iterations per trip of the warm-started loop.
#ifndef XC_ITERS_PER_TRIP
#define XC_ITERS_PER_TRIP 2
#endif

// Makes v's value unknown to the optimiser at this point (no code).
XC_HD __forceinline__ void xc_opaque(int &v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : "+v"(v));
#else
    __asm__ volatile("" : "+r"(v));
#endif
}

template <typename T> struct ColumnOut {
    T cape, cin, zout;
    int mulvl;
    T p_lfc, p_el;   // [hPa]; only from the drivers' LEVELS variants
};

struct LaneStats {
    uint32_t levels = 0, substeps = 0, iterations = 0, nonconverged = 0, fallbacks = 0;
    uint32_t first = 0;   // iterations of the ascent's first sub-step
    // the warm start's loop: the blocks this lane ran in its last trip (bit
    // b: level set-up, sub-step set-up, first iteration, second iteration,
    // sub-step completion, layer update), and the counts a kernel makes of
    // them -- lanes that ran each block, wave trips in which at least one
    // lane did (the wave pays for the block then), and wave trips
    uint32_t blocks = 0;
    uint32_t blk_lanes[6] = {0, 0, 0, 0, 0, 0};
    uint32_t blk_waves[6] = {0, 0, 0, 0, 0, 0};
    uint32_t wave_trips = 0;
};

template <typename T> struct Env {
    T p, t, td, pi, q, th, thv;
};

// A level's profile values as stored (hPa, degC).  The source scans read
// level k + 1 while they work on level k, so that the loads are in flight
// during a level's arithmetic instead of stalling the wave at its start.
struct Raw {
    float p, t, td;
};

template <class C>
XC_HD __forceinline__ Raw load_raw(const C &col, int k)
{
    return Raw{col.P(k), col.Tc(k), col.Tdc(k)};
}

// :232-240 for one level
template <typename T, int MATH>
XC_HD __forceinline__ Env<T> env_from(const Raw &r)
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;
    Env<T> e;
    e.p = T(100.0) * T(r.p);
    e.t = T(273.15) + T(r.t);
    e.td = T(273.15) + T(r.td);
    e.pi = F::pow(e.p * K::rp00, K::rddcp);
    e.q = qvs<T, MATH>(e.p, e.td);
    e.th = F::div(e.t, e.pi);
    e.thv = F::div(e.th * (T(1.0) + K::reps * e.q), T(1.0) + e.q);
    return e;
}

template <typename T, int MATH, class C>
XC_HD __forceinline__ Env<T> env_at(const C &col, int k)
{
    return env_from<T, MATH>(load_raw(col, k));
}

// layer thickness from the hydrostatic equation in Exner pressure (:246, :503)
template <typename T>
XC_HD __forceinline__ T layer_dz(T thv_k, T thv_km1, T pi_k, T pi_km1)
{
    return -Phys<T>::cpdg * T(0.5) * (thv_k + thv_km1) * (pi_k - pi_km1);
}

// One fixed-point evaluation (:437-462): the parcel state at p2 implied by
// the iterate thlast, and the theta it gives back.  fp_eval_ops: with the
// operations of math mode MATH.
template <typename T, int MATH, bool ICE>
XC_HD __forceinline__ void fp_eval_ops(T p2, T pi2, T logp, T t1, T th1, T qv1, T ql1, T qi1,
                                       T qt, T thlast, T &th2, T &t2, T &qv2, T &ql2, T &qi2)
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;
    t2 = thlast * pi2;
    if (ICE) {
        const T fliq = fmx(fmn(F::div(t2 - T(233.15), T(273.15) - T(233.15)), T(1.0)), T(0.0));
        const T fice = T(1.0) - fliq;
        qv2 = fmn(qt, fliq * qvs<T, MATH>(p2, t2) + fice * qvi<T, MATH>(p2, t2));
        qi2 = fmx(fice * (qt - qv2), T(0.0));
    } else {
        qv2 = fmn(qt, qvs<T, MATH>(p2, t2));
        qi2 = T(0.0);
    }
    ql2 = fmx(qt - qv2 - qi2, T(0.0));
    const T tbar = T(0.5) * (t1 + t2);
    const T qvbar = T(0.5) * (qv1 + qv2);
    const T qlbar = T(0.5) * (ql1 + ql2);
    const T lhv = K::lv1 - K::lv2 * tbar;
    const T rm = K::rd + K::rv * qvbar;
    T arg;
    if (ICE) {
        const T qibar = T(0.5) * (qi1 + qi2);
        const T lhs = K::ls1 - K::ls2 * tbar;
        const T cpm = K::cp + K::cpv * qvbar + K::cpl * qlbar + K::cpi * qibar;
        const T den = cpm * tbar;
        arg = F::div(lhv * (ql2 - ql1), den) + F::div(lhs * (qi2 - qi1), den) +
              (F::div(rm, cpm) - K::rddcp) * logp;
    } else {
        const T cpm = K::cp + K::cpv * qvbar + K::cpl * qlbar;
        arg = F::div(lhv * (ql2 - ql1), cpm * tbar) + (F::div(rm, cpm) - K::rddcp) * logp;
    }
    th2 = th1 * F::exp(arg);
}

#if defined(__HIP_DEVICE_COMPILE__)
// The accurate mode's evaluation on the device: the operations of
// fp_eval_ops in the same order, with the raw IEEE division kernels and
// exp_core (phys.h).  Wherever those give the library's values -- every
// operand in the normal range, exp arguments within (-87, 88) -- the result
// is fp_eval_ops<float, MATH_LIB>'s bit for bit; a zero or infinite divisor
// or other special value turns the raw division into NaN, which reaches th2
// (every quotient feeds qv2 or the exponent), and the evaluation is then
// repeated with the library's operations.  Divisors: the ice ramp's
// 273.15 - 233.15 by a correctly rounded reciprocal; the two latent-heat
// terms share one reciprocal of cpm * tbar.
template <bool ICE>
__device__ __forceinline__ void fp_eval_raw(float p2, float pi2, float logp, float t1, float th1,
                                            float qv1, float ql1, float qi1, float qt,
                                            float thlast, float &th2, float &t2, float &qv2,
                                            float &ql2, float &qi2)
{
    using K = Phys<float>;
    constexpr float RAMP = 273.15f - 233.15f;
    constexpr float RRAMP = 1.0f / RAMP;
    t2 = thlast * pi2;
    const float tc = t2 - 273.15f;
    const float esl = 611.2f * exp_core(div_raw(17.67f * tc, t2 - 29.65f));
    const float qsl = div_raw(K::eps * esl, p2 - esl);
    if (ICE) {
        const float fliq = fmx(fmn(div_const(t2 - 233.15f, RAMP, RRAMP), 1.0f), 0.0f);
        const float fice = 1.0f - fliq;
        const float esi = 611.2f * exp_core(div_raw(21.8745584f * tc, t2 - 7.66f));
        const float qsi = div_raw(K::eps * esi, p2 - esi);
        qv2 = fmn(qt, fliq * qsl + fice * qsi);
        qi2 = fmx(fice * (qt - qv2), 0.0f);
    } else {
        qv2 = fmn(qt, qsl);
        qi2 = 0.0f;
    }
    ql2 = fmx(qt - qv2 - qi2, 0.0f);
    const float tbar = 0.5f * (t1 + t2);
    const float qvbar = 0.5f * (qv1 + qv2);
    const float qlbar = 0.5f * (ql1 + ql2);
    const float lhv = K::lv1 - K::lv2 * tbar;
    const float rm = K::rd + K::rv * qvbar;
    float arg;
    if (ICE) {
        const float qibar = 0.5f * (qi1 + qi2);
        const float lhs = K::ls1 - K::ls2 * tbar;
        const float cpm = K::cp + K::cpv * qvbar + K::cpl * qlbar + K::cpi * qibar;
        float a1, a2;
        div2_raw(lhv * (ql2 - ql1), lhs * (qi2 - qi1), cpm * tbar, a1, a2);
        arg = a1 + a2 + (div_raw(rm, cpm) - K::rddcp) * logp;
    } else {
        const float cpm = K::cp + K::cpv * qvbar + K::cpl * qlbar;
        arg = div_raw(lhv * (ql2 - ql1), cpm * tbar) + (div_raw(rm, cpm) - K::rddcp) * logp;
    }
    th2 = th1 * exp_core(arg);
}
#endif

template <typename T, int MATH, bool ICE>
XC_HD __forceinline__ void fp_eval(T p2, T pi2, T logp, T t1, T th1, T qv1, T ql1, T qi1,
                                   T qt, T thlast, T &th2, T &t2, T &qv2, T &ql2, T &qi2)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4 && MATH == MATH_ACCURATE) {
        fp_eval_raw<ICE>(p2, pi2, logp, t1, th1, qv1, ql1, qi1, qt, thlast, th2, t2, qv2, ql2,
                         qi2);
        if (__builtin_expect(!__builtin_isfinite(th2), 0))
            fp_eval_ops<T, MATH_LIB, ICE>(p2, pi2, logp, t1, th1, qv1, ql1, qi1, qt, thlast,
                                          th2, t2, qv2, ql2, qi2);
        return;
    }
#endif
    fp_eval_ops<T, MATH, ICE>(p2, pi2, logp, t1, th1, qv1, ql1, qi1, qt, thlast, th2, t2, qv2,
                              ql2, qi2);
}

// Moist-adiabat fixed point for one sub-step from the cold start (:428-480).
// Returns false when the iteration limit is exceeded (the reference then
// zeroes CAPE/CIN and returns, :465-474).
template <typename T, int MATH, bool ICE>
XC_HD __forceinline__ bool solve_substep(T p1, T p2, T pi2, T t1, T th1, T qv1, T ql1,
                                              T qi1, T qt, T thlast, T &th2, T &t2, T &qv2,
                                              T &ql2, T &qi2, LaneStats &st)
{
    using F = Fn<T, MATH>;
    const T logp = F::log(F::div(p2, p1));
    int i = 0;
    while (true) {
        i = i + 1;
        fp_eval<T, MATH, ICE>(p2, pi2, logp, t1, th1, qv1, ql1, qi1, qt, thlast, th2, t2, qv2,
                              ql2, qi2);
        st.iterations++;
        if (i > 100) return false;
        const T d = th2 - thlast;
        if ((d < T(0.0) ? -d : d) > Phys<T>::converge) {
            thlast = thlast + T(0.3) * (th2 - thlast);
        } else {
            return true;
        }
    }
}

// Parcel source (:257-379): chooses the starting level kmax, the parcel's
// initial state and the height of kmax.  Returns false for an unknown source
// (unreachable through the Python layer; the reference stops there).
// The parcel starts at the pressure of kmax, so its p2 and pi2 are e.p and
// e.pi (not stored twice).
template <typename T> struct Parcel {
    int kmax, mulvl;
    T th2, t2, qv2, b2, zk;
    Env<T> e;       // environment at kmax
};

// An observer of the level environments the source scans evaluate:
// sink(k, env) for each level k they compute (a caller that also needs the
// environments keeps them instead of computing them again).  With
// every_level, the most-unstable scan also evaluates (and passes on) the
// levels above 500 hPa, which it reads for their pressure anyway: a caller
// that needs every level's environment then reads each level once.
struct NoEnvSink {
    static constexpr bool every_level = false;
    template <typename E> XC_HD __forceinline__ void operator()(int, const E &) const {}
};

template <typename T, int MATH, class C, class Sink = NoEnvSink>
XC_HD __forceinline__ bool parcel_source(const C &col, int nk, int source, T ml_depth,
                                              Parcel<T> &pc, const Sink &sink = Sink{})
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;
    pc.mulvl = -999999;
    int kmax = 0;
    double avgth_d = 0.0, avgqv_d = 0.0;   // real*8 in the reference (:183)
    T zk = T(0.0);
    if (source == SRC_MU) {
        T p0 = T(100.0) * T(col.P(0));
        if (!(p0 < T(50000.0))) {
            T maxthe = T(0.0);
            // z of the levels scanned so far (the reference's z(k), :244-248),
            // summed while every level below has been scanned: with no level
            // skipped below kmax, z(kmax) needs no second pass
            T zrun = T(0.0), zmax = T(0.0);
            int kskip = nk;   // first level skipped (p < 500 hPa)
            Env<T> ep{};
            Raw nx = load_raw(col, 0);
            for (int k = 0; k < nk; ++k) {
                const Raw cur = nx;
                if (k + 1 < nk) nx = load_raw(col, k + 1);
                T pk = T(100.0) * T(cur.p);
                const bool scan = pk >= T(50000.0);
                if (scan || Sink::every_level) {
                    Env<T> e = env_from<T, MATH>(cur);
                    sink(k, e);
                    if (scan) {
                        if (k > 0 && k < kskip)
                            zrun = zrun + layer_dz(e.thv, ep.thv, e.pi, ep.pi);
                        ep = e;
                        T the = thetae<T, MATH>(e.p, e.t, e.td, e.q);
                        if (the > maxthe) {
                            pc.mulvl = k + 1;
                            maxthe = the;
                            kmax = k;
                            zmax = zrun;
                        }
                    }
                }
                if (!scan && kskip == nk) kskip = k;
            }
            if (kmax < kskip) {
                zk = zmax;
            } else {
                Env<T> ep = env_at<T, MATH>(col, 0);
                for (int k = 1; k <= kmax; ++k) {
                    Env<T> e = env_at<T, MATH>(col, k);
                    zk = zk + layer_dz(e.thv, ep.thv, e.pi, ep.pi);
                    ep = e;
                }
            }
        }
    } else if (source == SRC_ML) {
        // z(1), z(2) and z(nk) decide the branch (:287-301)
        Env<T> e0 = env_at<T, MATH>(col, 0);
        sink(0, e0);
        Env<T> ep = e0;
        T z1 = T(0.0), ztop = T(0.0);
        Raw nx = nk > 1 ? load_raw(col, 1) : Raw{};
        for (int k = 1; k < nk; ++k) {
            const Raw cur = nx;
            if (k + 1 < nk) nx = load_raw(col, k + 1);
            Env<T> e = env_from<T, MATH>(cur);
            sink(k, e);
            ztop = ztop + layer_dz(e.thv, ep.thv, e.pi, ep.pi);
            if (k == 1) z1 = ztop;
            ep = e;
        }
        if ((z1 - T(0.0)) > ml_depth) {
            avgth_d = (double)e0.th;
            avgqv_d = (double)e0.q;
            kmax = 0;
        } else if (ztop < ml_depth) {
            avgth_d = (double)ep.th;
            avgqv_d = (double)ep.q;
            kmax = nk - 1;
            zk = ztop;
        } else {
            // trapezoid means up to ml_depth (:306-335), double accumulators
            Env<T> a = e0;
            T za = T(0.0);
            int k = 1;
            Env<T> b = env_at<T, MATH>(col, 1);
            T zb = za + layer_dz(b.thv, a.thv, b.pi, a.pi);
            while (zb <= ml_depth && k < nk) {
                avgth_d = avgth_d + (double)(T(0.5) * (zb - za) * (b.th + a.th));
                avgqv_d = avgqv_d + (double)(T(0.5) * (zb - za) * (b.q + a.q));
                k = k + 1;
                if (k >= nk) break;
                a = b;
                za = zb;
                b = env_at<T, MATH>(col, k);
                zb = za + layer_dz(b.thv, a.thv, b.pi, a.pi);
            }
            // (k == nk needs z(nk) == ml_depth exactly; stay on the last layer)
            T th2i = a.th + F::div((b.th - a.th) * (ml_depth - za), zb - za);
            T qv2i = a.q + F::div((b.q - a.q) * (ml_depth - za), zb - za);
            avgth_d = avgth_d + (double)(T(0.5) * (ml_depth - za) * (th2i + a.th));
            avgqv_d = avgqv_d + (double)(T(0.5) * (ml_depth - za) * (qv2i + a.q));
            avgth_d = avgth_d / (double)ml_depth;
            avgqv_d = avgqv_d / (double)ml_depth;
            kmax = 0;
        }
    } else if (source != SRC_SURFACE) {
        return false;
    }

    // parcel at its initial location (:355-375)
    pc.e = env_at<T, MATH>(col, kmax);
    if (source != SRC_ML) {
        pc.th2 = pc.e.th; pc.t2 = pc.e.t; pc.qv2 = pc.e.q;
        pc.b2 = T(0.0);
    } else {
        pc.th2 = (T)avgth_d;
        pc.qv2 = (T)avgqv_d;
        T thv2 = F::div(pc.th2 * (T(1.0) + K::reps * pc.qv2), T(1.0) + pc.qv2);
        pc.t2 = pc.th2 * pc.e.pi;
        pc.b2 = F::div(K::g * (thv2 - pc.e.thv), pc.e.thv);
    }
    pc.kmax = kmax;
    pc.zk = zk;
    return true;
}

// The level of free convection and the equilibrium level, from the
// reference's own crossing pressure ps (:511, :526), which it computes and
// drops (the reference has no such output).  Over the layers whose trapezoid
// adds positive area: the LFC is where the first one's positive area begins
// (ps when the parcel enters positive buoyancy in that layer, else the
// layer's bottom), the EL where the last one's ends (ps when it leaves
// positive buoyancy, else the layer's top).  Pressures in Pa; NaN while there
// is no positive area.
template <typename T, int MATH> struct LevelTrack {
    T lfc = T(NAN), el = T(NAN);
    XC_HD __forceinline__ void add(bool enter, bool leave, T parea, T b1, T b2, T p_below,
                                   T p_k)
    {
        if (!(parea > T(0.0))) return;
        const T ps = (enter || leave)
                         ? p_below + Fn<T, MATH>::div((p_k - p_below) * (T(0.0) - b1), b2 - b1)
                         : T(0.0);
        if (lfc != lfc) lfc = enter ? ps : p_below;
        el = leave ? ps : p_k;
    }
};

// The default: no levels tracked.
struct NoTrack {
    template <typename T>
    XC_HD __forceinline__ void add(bool, bool, T, T, T, T, T) {}
};

// Buoyancy of the parcel at level k and the CAPE/CIN trapezoid of the layer
// below it (:501-547).  CIN is committed only on entering positive buoyancy.
// p_below: the pressure of level k - 1, for the tracker.
// e_thv, e_pi, e_p: the environment of level k.
template <typename T, int MATH, class TRK = NoTrack>
XC_HD __forceinline__ void layer_update(T th2, T qv2, T ql2, T qi2, T e_thv, T e_pi, T e_p,
                                        T thv_prev, T pi_prev, T b1, T &b2, T &dz, T &cape,
                                        T &cin, T &narea, T p_below = T(0.0),
                                        TRK *trk = nullptr)
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;
    T thv2 = T(0.0);
    bool raw = false;
#if defined(__HIP_DEVICE_COMPILE__)
    // the accurate mode's raw division kernel (phys.h): thv2 and b2 have
    // divisors of at least 1 and the environment's thv, so finite results are
    // the IEEE quotients; otherwise both again with IEEE division
    if constexpr (sizeof(T) == 4 && MATH == MATH_ACCURATE) {
        thv2 = div_raw(th2 * (1.0f + K::reps * qv2), 1.0f + qv2 + ql2 + qi2);
        b2 = div_raw(K::g * (thv2 - e_thv), e_thv);
        raw = __builtin_isfinite(thv2) && __builtin_isfinite(b2);
    }
#endif
    if (!raw) {
        thv2 = F::div(th2 * (T(1.0) + K::reps * qv2), T(1.0) + qv2 + ql2 + qi2);
        b2 = F::div(K::g * (thv2 - e_thv), e_thv);
    }
    dz = layer_dz(e_thv, thv_prev, e_pi, pi_prev);
    // the reference's four cases (:509-545) as selects over one division:
    // each case's expressions are evaluated as written there, and the wave
    // does not run the four branches one after the other
    const bool enter = b2 >= T(0.0) && b1 < T(0.0);   // entering positive area
    const bool leave = b2 < T(0.0) && b1 > T(0.0);    // leaving it
    const bool neg = b2 < T(0.0) && !leave;
    const T fnum = enter ? b2 : b1, fden = enter ? b2 - b1 : b1 - b2;
    // (used only when entering or leaving, where fden = |b1| + |b2| > 0 and
    // the buoyancies are finite: the raw quotient is the IEEE one)
    T frac;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4 && MATH == MATH_ACCURATE)
        frac = raw ? div_raw(fnum, fden) : F::div(fnum, fden);
    else
#endif
        frac = F::div(fnum, fden);
    const T half_dz_sum = T(0.5) * dz * (b1 + b2);
    T parea;
    if (enter) {
        parea = T(0.5) * b2 * dz * frac;
        cin = cin + (narea - T(0.5) * b1 * dz * (T(1.0) - frac));
        narea = T(0.0);
    } else if (leave) {
        parea = T(0.5) * b1 * dz * frac;
        narea = -T(0.5) * b2 * dz * (T(1.0) - frac);
    } else if (neg) {
        parea = T(0.0);
        narea = narea - half_dz_sum;
    } else {
        parea = half_dz_sum;
        narea = T(0.0);
    }
    cape = cape + fmx(T(0.0), parea);
    if (trk) trk->add(enter, leave, parea, b1, b2, p_below, e_p);
}

// sub-step count and size for the layer p_prev -> p_k (:408-415)
template <typename T, int MATH>
XC_HD __forceinline__ int substep_count(T &dp, T pinc)
{
    if (dp < pinc) return 1;
    const int nloop = 1 + ftrunc(Fn<T, MATH>::div(dp, pinc));
    dp = Fn<T, MATH>::div(dp, (T)nloop);
    return nloop;
}

// A level as the ascent enters it (:403-415, and the test of :554): its
// environment's p, pi and thv, the sub-step size dp and count nloop of the
// layer from the level below (substep_count of the pressure drop), and
// whether p <= 100 hPa (where negative buoyancy ends the ascent).
template <typename T> struct Level {
    T p, pi, thv, dp;
    int nloop;
    bool top;
};

// A level as an earlier pass stores it: (dp, pi, thv, nloop | stop << 31),
// nloop clamped to [0, 2^30] (a count below 1 -- NaN or overflowing
// pressures -- runs no sub-step either way).
struct LevelRec {
    float dp, pi, thv;
    uint32_t meta;
};

XC_HD __forceinline__ LevelRec level_pack(const Level<float> &lv)
{
    const int n = lv.nloop < 0 ? 0 : (lv.nloop > (1 << 30) ? (1 << 30) : lv.nloop);
    return LevelRec{lv.dp, lv.pi, lv.thv, (uint32_t)n | (lv.top ? 0x80000000u : 0u)};
}

XC_HD __forceinline__ Level<float> level_unpack(const LevelRec &r, float p)
{
    Level<float> lv;
    lv.dp = r.dp;
    lv.pi = r.pi;
    lv.thv = r.thv;
    lv.nloop = (int)(r.meta & 0x7fffffffu);
    lv.top = (r.meta >> 31) != 0u;
    lv.p = p;
    return lv;
}

template <typename T, int MATH>
XC_HD __forceinline__ Level<T> level_from(const Env<T> &e, T p_below, T pinc)
{
    Level<T> lv;
    lv.p = e.p;
    lv.pi = e.pi;
    lv.thv = e.thv;
    lv.dp = p_below - e.p;
    lv.nloop = substep_count<T, MATH>(lv.dp, pinc);
    lv.top = e.p <= T(10000.0);
    return lv;
}

// ---------------------------------------------------------------------------
// Sub-step observer of the reference driver: called after every converged
// sub-step with its start state (p1, t1, th1, qv1), end pressure p2, total
// water qt and converged th2.  NoLog is the default; the training-data logger
// records the surrogate's inputs and target.
struct NoLog {
    template <typename T>
    XC_HD __forceinline__ void operator()(T, T, T, T, T, T, T) const {}
};

// Reference driver: nested loops, cold start.  LEVELS: also the LFC and EL
// (LevelTrack).
template <typename T, int MATH, bool ICE, bool PSEUDO, class C, class LOG = NoLog,
          bool LEVELS = false>
XC_HD void cape_column(const C &col, int nk, T pinc, int source, T ml_depth,
                       ColumnOut<T> &out, LaneStats &st, const LOG &log = LOG())
{
    using K = Phys<T>;
    Parcel<T> pc;
    out.p_lfc = T(NAN);
    out.p_el = T(NAN);
    if (!parcel_source<T, MATH>(col, nk, source, ml_depth, pc)) {
        out.cape = T(0.0); out.cin = T(0.0); out.mulvl = pc.mulvl; out.zout = T(-999999.0);
        return;
    }
    LevelTrack<T, MATH> trk;
    const int mulvl = pc.mulvl;
    T zout = T(-999999.0);
    T th2 = pc.th2, pi2 = pc.e.pi, p2 = pc.e.p, t2 = pc.t2, qv2 = pc.qv2, b2 = pc.b2, zk = pc.zk;
    T ql2 = T(0.0), qi2 = T(0.0), qt = qv2;
    T cape = T(0.0), cin = T(0.0), narea = T(0.0);
    T p_prev = pc.e.p, pi_prev = pc.e.pi, thv_prev = pc.e.thv;
    int k = pc.kmax;
    bool doit = true;
    bool first_sub = true;

    // ascent (:403-559)
    while (doit && k < nk - 1) {
        k = k + 1;
        Env<T> e = env_from<T, MATH>(col.raw_above(k));
        const T b1 = b2;
        T dp = p_prev - e.p;
        const int nloop = substep_count<T, MATH>(dp, pinc);
        st.levels++;
        for (int n = 1; n <= nloop; ++n) {
            // a pseudo-adiabat carries no condensate from one sub-step to
            // the next (:487-491), so its ql1, qi1 are the constant zero
            const T p1 = p2, t1 = t2, th1 = th2, qv1 = qv2;
            const T ql1 = PSEUDO ? T(0.0) : ql2, qi1 = PSEUDO ? T(0.0) : qi2;
            p2 = p2 - dp;
            pi2 = Fn<T, MATH>::pow(p2 * K::rp00, K::rddcp);
            st.substeps++;
            const uint32_t it0 = st.iterations;
            const bool ok = solve_substep<T, MATH, ICE>(p1, p2, pi2, t1, th1, qv1, ql1, qi1, qt,
                                                        th1, th2, t2, qv2, ql2, qi2, st);
            if (first_sub) st.first += st.iterations - it0;
            first_sub = false;
            if (!ok) {
                st.nonconverged++;
                out.cape = T(0.0); out.cin = T(0.0); out.mulvl = mulvl; out.zout = zout;
                return;
            }
            log(p1, t1, th1, qv1, p2, qt, th2);
            if (PSEUDO) {
                qt = qv2;
                ql2 = T(0.0);
                qi2 = T(0.0);
            }
        }
        T dz;
        if constexpr (LEVELS)
            layer_update<T, MATH>(th2, qv2, ql2, qi2, e.thv, e.pi, e.p, thv_prev, pi_prev, b1,
                                  b2, dz, cape, cin, narea, p_prev, &trk);
        else
            layer_update<T, MATH>(th2, qv2, ql2, qi2, e.thv, e.pi, e.p, thv_prev, pi_prev, b1,
                                  b2, dz, cape, cin, narea);
        if (e.p <= T(10000.0) && b2 < T(0.0)) doit = false;
        zk = zk + dz;
        zout = zk;
        p_prev = e.p; pi_prev = e.pi; thv_prev = e.thv;
    }
    out.cape = cape; out.cin = cin; out.mulvl = mulvl; out.zout = zout;
    if constexpr (LEVELS) {
        out.p_lfc = Fn<T, MATH>::div(trk.lfc, T(100.0));
        out.p_el = Fn<T, MATH>::div(trk.el, T(100.0));
    }
}

// ---------------------------------------------------------------------------
// Warm-started driver.
//
// The fixed point of each sub-step is the reference's (same evaluation, same
// 2e-4 K test, same 100-iteration limit, relaxation 0.3 from the third
// iterate on); the first two iterates differ.  When the predicted first
// iterate fails the test, the second is the secant step
// x + (f(x) - x) / (1 - f') with the map's slope f' measured on the
// sub-steps before (most of those sub-steps then pass at the second).  The
// first iterate is predicted from the sub-steps already converged in
// this column: the rate r = (x* - th1) / ln(p2/p1) of the last two sub-steps
// extrapolated linearly in ln p to the midpoint of the new sub-step (second
// order; first order with one sub-step of history).  x* is the secant
// (Aitken) estimate of the sub-step's exact fixed point,
// x* = x + (f(x) - x) / (1 - f'), with f' from the last two iterates: the
// converged th2 itself is only within ~1e-4 K of the fixed point, and
// differencing that noise into rates would make the guesses no better than
// a few 1e-4 K.  x* only feeds the predictor; the state carried to the next
// sub-step is the reference iteration's th2.  With no history (the
// ascent's first sub-step) the caller's guess is used -- the learned surrogate
// when one is loaded, else the cold start.  Fallback: when the first
// iteration's residual |f(g) - g| exceeds `fallback_tol` (1e-3 K) the guess
// is discarded and the sub-step restarts from the reference's cold start
// th1; a failed second-order guess also restarts the history at first order.
//
// Loop shape: with one or two iterations per sub-step the wavefront would
// spend most of its time waiting for the lane with the most iterations if
// sub-steps were lock-stepped.  Here every trip does up to XC_ITERS_PER_TRIP
// iterations for every running lane, and each lane moves to its next sub-step
// or level on its own (the set-up blocks execute under the mask of the lanes
// that need them).
//
// The driver's state is a WarmAscent: init() loads a parcel, each trip()
// runs one trip and reports when the column is done, and result() gives the
// outputs.  cape_column_warm runs one column to completion; the grid kernel
// instead hands a lane its next column when trip() reports done, so that a
// wavefront's lanes stay busy until the whole grid is drained (a wave that
// ran 64 fixed columns would idle every lane until its longest column ended).
//
// Levels come from the accessor: level_from() of env_from() of its raw
// values, or, for an accessor that declares has_env, its lev(k) (levels
// evaluated by an earlier pass with the same functions), loaded one level
// ahead.
template <class C, class = void> struct HasEnv {
    static constexpr bool value = false;
};
template <class C> struct HasEnv<C, decltype((void)C::has_env)> {
    static constexpr bool value = C::has_env;
};


// Shape of d(theta)/d(ln p) along a saturated pseudo-adiabat at (t, p, th),
// up to a constant factor: th qs (a1 - t) / (t^2 + a2 qs), a1 = L kappa / Rv,
// a2 = L^2 / (cp Rv), with qs and L of liquid water, or of ice (`ice`).
// From dtheta / theta = -L dqs / (cp t) with Clausius-Clapeyron for dqs and
// t = theta pi.  The predictor uses only ratios of it, to scale the last
// measured rate from one sub-step to the next; its arithmetic is the
// predictor's (hardware exp and reciprocal on the device).  One code path
// for both phases, the constants selected per lane.
template <typename T>
XC_HD __forceinline__ T adiabat_shape(T t, T p, T th, bool ice)
{
    using K = Phys<T>;
    constexpr T a1l = K::xlv * K::rddcp / K::rv, a2l = K::xlv * K::xlv / (K::cp * K::rv);
    constexpr T a1i = K::xls * K::rddcp / K::rv, a2i = K::xls * K::xls / (K::cp * K::rv);
    const T a = ice ? T(21.8745584) : T(17.67);
    const T b = ice ? T(7.66) : T(29.65);
    const T a1 = ice ? a1i : a1l;     // L kappa / Rv
    const T a2 = ice ? a2i : a2l;     // L^2 / (cp Rv)
    const T es = T(611.2) * pred_exp(pred_div(a * (t - T(273.15)), t - b));
    const T qs = K::eps * pred_div(es, p - es);
    return pred_div(th * qs * (a1 - t), t * t + a2 * qs);
}

template <typename T, int MATH, bool ICE, bool PSEUDO, bool LEVELS = false>
struct WarmAscent {
    enum { M_LEVEL = 0, M_SUB = 1, M_ITER = 2, M_CONV = 3, M_LAYER = 4 };
    int mode, k, nk, left, it, nh;   // left: sub-steps of the level not yet begun
    bool warm, fell_back, first_sub;
    bool layered;   // a layer is complete: the height output is zk (else -999999)
    T th2, pi2, p2, t2, qv2, b2, zk, ql2, qi2, qt;
    T cape, cin, narea;
    // the previous level's pi and thv (the layer update needs both levels),
    // and the current level's environment and stop test
    T pi_prev, thv_prev;
    T e_p, e_pi, e_thv;
    bool e_top;
    // an accessor with evaluated levels: the next one as stored (and its
    // pressure, for the LFC/EL), loaded a level ahead so that the load is in
    // flight while this level ascends
    LevelRec nx;
    T nx_p;
    T dp;
    // sub-step state
    T p1, t1, th1, qv1, ql1, qi1, logp, thlast;
    // previous iterate (for the secant slope) and the last slope estimate
    // the last relaxed iterate (the one before thlast) and its image, for the
    // secant slope, and the last estimate of the map's slope f'
    T x_prev, f_prev, fpr;
    // extrapolation history: last two rates and their ln-p steps (while
    // there is no history, r1 holds the caller's first guess, NaN for none)
    T r1, r2, dl1, dl2;
    // adiabat_shape at the current sub-step's start (NaN: not a pure liquid
    // or pure ice regime), and whether the last sub-step condensed
    T gs;
    bool sat;
    // LEVELS: the LFC/EL tracker and the previous level's pressure
    LevelTrack<T, MATH> trk;
    T pb;

    // Starts the ascent of parcel pc in a column of nk levels.  Returns false
    // when there is no level above the parcel (result() is then final).
    XC_HD __forceinline__ bool init(const Parcel<T> &pc, int nk_, T guess)
    {
        nk = nk_;
        layered = false;
        th2 = pc.th2; pi2 = pc.e.pi; p2 = pc.e.p; t2 = pc.t2; qv2 = pc.qv2; b2 = pc.b2;
        zk = pc.zk;
        ql2 = T(0.0); qi2 = T(0.0); qt = qv2;
        cape = T(0.0); cin = T(0.0); narea = T(0.0);
        pi_prev = pc.e.pi; thv_prev = pc.e.thv;
        k = pc.kmax;
        e_p = pc.e.p; e_pi = pc.e.pi; e_thv = pc.e.thv; e_top = false;
        dp = T(0.0);
        left = 0; it = 0; nh = 0;
        p1 = T(0.0); t1 = T(0.0); th1 = T(0.0); qv1 = T(0.0); ql1 = T(0.0); qi1 = T(0.0);
        logp = T(0.0); thlast = T(0.0);
        warm = false; fell_back = false; first_sub = true;
        x_prev = T(0.0); f_prev = T(0.0); fpr = T(0.0);
        r1 = guess; r2 = T(0.0); dl1 = T(0.0); dl2 = T(0.0);
        gs = T(NAN);
        sat = false;
        if constexpr (LEVELS) {
            trk = LevelTrack<T, MATH>{};
            pb = pc.e.p;
        }
        mode = M_LEVEL;
        return k < nk - 1;
    }

    // The first level above the parcel, for accessors with evaluated levels
    // (call after init()).
    XC_HD __forceinline__ void set_next(const LevelRec &l1, T p1)
    {
        nx = l1;
        nx_p = p1;
    }

    XC_HD __forceinline__ void result(ColumnOut<T> &out) const
    {
        out.cape = cape; out.cin = cin; out.zout = layered ? zk : T(-999999.0);
        if constexpr (LEVELS) {
            out.p_lfc = Fn<T, MATH>::div(trk.lfc, T(100.0));
            out.p_el = Fn<T, MATH>::div(trk.el, T(100.0));
        }
    }

    // One trip.  Returns true when the column is finished (terminated, top
    // reached, or an iteration limit exceeded: then CAPE = CIN = 0 as in the
    // reference, :464-474).
    template <class C>
    XC_HD __forceinline__ bool trip(const C &col, T pinc, T fallback_tol, LaneStats &st,
                                    bool predict)
    {
        using F = Fn<T, MATH>;
        using K = Phys<T>;
        // the predictor's own arithmetic (slopes, rates, extrapolation
        // weights) only shapes the first iterate, never the fixed point: it
        // divides with the hardware reciprocal in every math mode
        const auto pdiv = [](T a, T b) { return pred_div(a, b); };
        // Every trip runs the same straight-line sequence (level set-up,
        // sub-step set-up, the iterations, level completion), each part under
        // the mask of the lanes whose mode asks for it; `mode` is made opaque
        // at the top of the trip so that the compiler cannot thread the
        // iterate-again path into an inner loop (which would make the lanes
        // that converged wait for the slowest one again).
        xc_opaque(mode);
        st.blocks = 0;
        if (mode == M_LEVEL) {
            st.blocks |= 1u;
            k = k + 1;
            if constexpr (LEVELS) pb = e_p;
            Level<T> lv;
            if constexpr (HasEnv<C>::value) {
                lv = level_unpack(nx, nx_p);
                if (k + 1 < nk) {
                    nx = col.rec_at(k + 1);
                    if constexpr (LEVELS) nx_p = col.p_at(k + 1);
                }
            } else {
                lv = level_from<T, MATH>(env_from<T, MATH>(col.raw_above(k)), e_p, pinc);
            }
            e_p = lv.p; e_pi = lv.pi; e_thv = lv.thv; e_top = lv.top;
            dp = lv.dp;
            left = lv.nloop;
            st.levels++;
            // zero sub-steps (NaN/overflow in the step count): the reference's
            // loop body does not run; straight to the buoyancy update
            mode = left >= 1 ? M_SUB : M_LAYER;
        }
        if (mode == M_SUB) {
            st.blocks |= 2u;
            left = left - 1;
            p1 = p2; t1 = t2; th1 = th2; qv1 = qv2;
            ql1 = PSEUDO ? T(0.0) : ql2;
            qi1 = PSEUDO ? T(0.0) : qi2;
            p2 = p2 - dp;
            pi2 = F::pow(p2 * K::rp00, K::rddcp);
            logp = F::log(F::div(p2, p1));
            T g = th1;
            if (!predict) {
                // every sub-step cold: the flattened loop without prediction
            } else if (nh >= 2) {
                const T r = r1 + (r1 - r2) * pdiv(logp + dl1, dl1 + dl2);
                g = th1 + r * logp;
            } else if (nh == 1) {
                g = th1 + r1 * logp;
            } else if (r1 == r1) {
                // the caller's guess, used once
                g = r1;
                r1 = T(NAN);
            }
            // a guess that is not finite or absurdly far is no guess
            const T dg = g - th1;
            if (!((dg < T(0.0) ? -dg : dg) < T(50.0))) g = th1;
            thlast = g;
            warm = g != th1;
            fell_back = false;
            it = 0;
            st.substeps++;
            mode = M_ITER;
        }
        // up to XC_ITERS_PER_TRIP fixed-point iterations (:436-479): most
        // warm-started sub-steps converge at the first or second, so two per
        // trip halve the trips that pay for the set-up blocks around them.
        // Not unrolled: the unrolled pair, each with the evaluation's IEEE
        // re-evaluation path, no longer fits the 64 registers
#pragma unroll 1
        for (int rep = 0; rep < XC_ITERS_PER_TRIP; ++rep) {
            if (mode != M_ITER) continue;
            st.blocks |= 4u << rep;
            it = it + 1;
            const T x = thlast;
            fp_eval<T, MATH, ICE>(p2, pi2, logp, t1, th1, qv1, ql1, qi1, qt, x, th2, t2, qv2,
                                  ql2, qi2);
            st.iterations++;
            if (first_sub) st.first++;
            const T d = th2 - x;
            const T ad = d < T(0.0) ? -d : d;
            if (it > 100) {
                st.nonconverged++;
                cape = T(0.0);
                cin = T(0.0);
                // a column stopped by the iteration limit reports no levels
                if constexpr (LEVELS) trk = LevelTrack<T, MATH>{};
                return true;
            }
            if (warm && it == 1 && ad > fallback_tol) {
                // fallback: the reference's cold start, full iteration budget
                thlast = th1;
                warm = false;
                fell_back = true;
                it = 0;
                st.fallbacks++;
            } else if (ad > K::converge) {
                // the reference's relaxed step; a warm-started sub-step whose
                // first iterate missed takes the secant step instead, with the
                // map's slope f' from the sub-steps before
                T nxt = x + T(0.3) * (th2 - x);
                if (warm && it == 1) nxt = x + pdiv(th2 - x, T(1.0) - fpr);
                thlast = nxt;
                x_prev = x;
                f_prev = th2;
            } else {
                // converged (:475-479): finished below, once per trip
                mode = M_CONV;
            }
        }
        if (mode == M_CONV) {
            // record the rate, finish the sub-step (:485-497).  x = thlast is
            // the iterate that converged.  The slope f' of the map from this
            // sub-step's last two iterates (kept from an earlier sub-step
            // when it converged at its first)
            const T x = thlast;
            if (it >= 2) {
                const T dx = x - x_prev;
                if (dx != T(0.0)) {
                    const T sl = pdiv(th2 - f_prev, dx);
                    fpr = sl < T(-10.0) ? T(-10.0) : (sl > T(0.9) ? T(0.9) : sl);
                }
            }
        }
        if (mode == M_CONV) st.blocks |= 16u;
        if (mode == M_CONV && warm) {
            // Would the reference's own solve, from the cold start th1 with
            // its relaxed step (contraction c = 1 - 0.3 (1 - f') per
            // iteration), take more than about 50 iterations here?  Then it
            // may hit its 100-iteration limit (which zeroes the column),
            // while a predicted iterate near the fixed point need not: the
            // plain solve decides, restarted cold.  Linearised count:
            // |th1 - x| |c|^50 > the tolerance.  Never on ordinary soundings
            // (the cold solve takes about 10 iterations there).
            const T c = T(1.0) - T(0.3) * (T(1.0) - fpr);
            const T c2 = c * c, c4 = c2 * c2, c8 = c4 * c4, c16 = c8 * c8;
            const T c50 = c16 * c16 * c16 * c2;
            const T e0 = thlast - th1;
            if ((e0 < T(0.0) ? -e0 : e0) * c50 > K::converge) {
                thlast = th1;
                warm = false;
                fell_back = true;
                it = 0;
                st.fallbacks++;
                mode = M_ITER;
            }
        }
        if (mode == M_CONV) {
            const T x = thlast;
            first_sub = false;
            // the rate's shape at the next sub-step's start, for the
            // pseudo-adiabats (the shape assumes the condensate falls out),
            // where the condensate is all liquid (above 0 degC; everywhere
            // for the liquid adiabat) or all ice (below -40 degC); none in
            // the ice adiabat's mixed range
            const bool sat_now = qv2 < qt;
            T ge = T(NAN);
            if (PSEUDO && predict) {
                const bool ice = ICE && t2 < T(233.15);
                if (ice || !ICE || t2 > T(273.15)) ge = adiabat_shape<T>(t2, p2, th2, ice);
            }
            if (logp != T(0.0)) {
                const T xs = x + pdiv(th2 - x, T(1.0) - fpr);
                const T rn = pdiv(xs - th1, logp);
                // saturated ascent within one phase: a history whose
                // second-order extrapolation is this rate scaled by the
                // shape's change between the sub-steps' midpoints (log-linear
                // in ln p, linearised):
                // r = rn (1 + (ge / gs - 1) (logp' + dl1) / (2 dl1)).
                // Otherwise the measured history; a second-order guess that
                // failed marks a change of regime (saturation, ice onset):
                // the older rate belongs to the other side, so restart at
                // first order
                const bool shaped = sat && sat_now && gs == gs && ge == ge;
                const T r2n = shaped ? rn * (T(2.0) - pdiv(ge, gs)) : r1;
                const T dl2n = shaped ? logp : dl1;
                nh = shaped ? 2 : (fell_back && nh >= 2 ? 1 : (nh < 2 ? nh + 1 : 2));
                r2 = r2n;
                dl2 = dl2n;
                r1 = rn;
                dl1 = logp;
            } else {
                nh = 0;
                r1 = T(NAN);
            }
            gs = ge;
            sat = sat_now;
            if (PSEUDO) {
                qt = qv2;
                ql2 = T(0.0);
                qi2 = T(0.0);
            }
            mode = left > 0 ? M_SUB : M_LAYER;
        }
        if (mode == M_LAYER) {
            st.blocks |= 32u;
            // level complete (:501-558)
            T dz;
            const T b1 = b2;
            if constexpr (LEVELS)
                layer_update<T, MATH>(th2, qv2, ql2, qi2, e_thv, e_pi, e_p, thv_prev, pi_prev,
                                      b1, b2, dz, cape, cin, narea, pb, &trk);
            else
                layer_update<T, MATH>(th2, qv2, ql2, qi2, e_thv, e_pi, e_p, thv_prev, pi_prev,
                                      b1, b2, dz, cape, cin, narea);
            zk = zk + dz;
            layered = true;
            pi_prev = e_pi; thv_prev = e_thv;
            if ((e_top && b2 < T(0.0)) || !(k < nk - 1)) return true;
            mode = M_LEVEL;
        }
        return false;
    }
};

// One column to completion.  `pc` comes from parcel_source (the caller may
// have used it to build the first sub-step's surrogate inputs).  predict =
// false starts every sub-step cold: the same loop without prediction, which
// separates the loop's share of the speed-up from the prediction's.
template <typename T, int MATH, bool ICE, bool PSEUDO, bool LEVELS = false, class C>
XC_HD void cape_column_warm(const C &col, int nk, T pinc, const Parcel<T> &pc,
                                 T first_guess, T fallback_tol, ColumnOut<T> &out,
                                 LaneStats &st, bool predict = true)
{
    WarmAscent<T, MATH, ICE, PSEUDO, LEVELS> w;
    out.mulvl = pc.mulvl;
    if (w.init(pc, nk, first_guess)) {
        for (;;)
            if (w.trip(col, pinc, fallback_tol, st, predict)) break;
    }
    w.result(out);
}

}  // namespace xcape
