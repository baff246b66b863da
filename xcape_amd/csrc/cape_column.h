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
// The column accessor C provides P(k), Tc(k), Tdc(k) in hPa / degC for
// k = 0 .. nk-1, k = 0 being the surface.
#pragma once
#include "phys.h"

namespace xcape {

enum { SRC_SURFACE = 1, SRC_MU = 2, SRC_ML = 3 };

template <typename T> struct ColumnOut {
    T cape, cin, zout;
    int mulvl;
};

struct LaneStats {
    uint32_t levels = 0, substeps = 0, iterations = 0, nonconverged = 0;
};

template <typename T> struct Env {
    T p, t, td, pi, q, th, thv;
};

// :232-240 for one level
template <typename T, int MATH, class C>
__device__ __forceinline__ Env<T> env_at(const C& col, int k)
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;
    Env<T> e;
    e.p = T(100.0) * T(col.P(k));
    e.t = T(273.15) + T(col.Tc(k));
    e.td = T(273.15) + T(col.Tdc(k));
    e.pi = F::pow(e.p * K::rp00, K::rddcp);
    e.q = qvs<T, MATH>(e.p, e.td);
    e.th = F::div(e.t, e.pi);
    e.thv = F::div(e.th * (T(1.0) + K::reps * e.q), T(1.0) + e.q);
    return e;
}

// layer thickness from the hydrostatic equation in Exner pressure (:246, :503)
template <typename T>
__device__ __forceinline__ T layer_dz(T thv_k, T thv_km1, T pi_k, T pi_km1)
{
    return -Phys<T>::cpdg * T(0.5) * (thv_k + thv_km1) * (pi_k - pi_km1);
}

// Moist-adiabat fixed point for one sub-step (:428-480).  Returns false when
// the iteration limit is exceeded (the reference then zeroes CAPE/CIN and
// returns, :465-474).  On entry th2 holds the starting guess (the cold start
// is th1, :431); on exit th2/t2/qv2/ql2/qi2 are the converged state.
template <typename T, int MATH, bool ICE>
__device__ __forceinline__ bool solve_substep(T p1, T p2, T pi2, T t1, T th1, T qv1, T ql1,
                                              T qi1, T qt, T thlast, T& th2, T& t2, T& qv2,
                                              T& ql2, T& qi2, LaneStats& st)
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;
    const T logp = F::log(F::div(p2, p1));
    const T ln_coef = K::rddcp;
    int i = 0;
    while (true) {
        i = i + 1;
        t2 = thlast * pi2;
        T fliq, fice;
        if (ICE) {
            fliq = fmx(fmn(F::div(t2 - T(233.15), T(273.15) - T(233.15)), T(1.0)), T(0.0));
            fice = T(1.0) - fliq;
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
                  (F::div(rm, cpm) - ln_coef) * logp;
        } else {
            const T cpm = K::cp + K::cpv * qvbar + K::cpl * qlbar;
            arg = F::div(lhv * (ql2 - ql1), cpm * tbar) + (F::div(rm, cpm) - ln_coef) * logp;
        }
        th2 = th1 * F::exp(arg);
        st.iterations++;
        if (i > 100) return false;
        const T d = th2 - thlast;
        if ((d < T(0.0) ? -d : d) > K::converge) {
            thlast = thlast + T(0.3) * (th2 - thlast);
        } else {
            return true;
        }
    }
}

// Warm-start policy: WarmStart::guess(...) returns the first iterate of the
// fixed point for a sub-step; ColdStart reproduces the reference (:431).
struct ColdStart {
    template <typename T>
    __device__ __forceinline__ T guess(T th1, T, T, T, T) const { return th1; }
};

template <typename T, int MATH, bool ICE, bool PSEUDO, class C, class WS = ColdStart>
__device__ void cape_column(const C& col, int nk, T pinc, int source, T ml_depth,
                            ColumnOut<T>& out, LaneStats& st, const WS& ws = WS())
{
    using F = Fn<T, MATH>;
    using K = Phys<T>;

    int mulvl = -999999;
    T zout = T(-999999.0);

    // ---- parcel source (:257-352) --------------------------------------
    int kmax = 0;
    T avgth = T(0.0), avgqv = T(0.0);   // ML means; double in the reference
    double avgth_d = 0.0, avgqv_d = 0.0;
    T zk = T(0.0);                      // z at kmax
    if (source == SRC_MU) {
        T p0 = T(100.0) * T(col.P(0));
        if (!(p0 < T(50000.0))) {
            T maxthe = T(0.0);
            for (int k = 0; k < nk; ++k) {
                T pk = T(100.0) * T(col.P(k));
                if (pk >= T(50000.0)) {
                    Env<T> e = env_at<T, MATH>(col, k);
                    T the = thetae<T, MATH>(e.p, e.t, e.td, e.q);
                    if (the > maxthe) {
                        mulvl = k + 1;
                        maxthe = the;
                        kmax = k;
                    }
                }
            }
            // height of the source level
            if (kmax > 0) {
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
        Env<T> ep = e0;
        T z1 = T(0.0), ztop = T(0.0);
        for (int k = 1; k < nk; ++k) {
            Env<T> e = env_at<T, MATH>(col, k);
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
        avgth = (T)avgth_d;
        avgqv = (T)avgqv_d;
    } else if (source != SRC_SURFACE) {
        out.cape = T(0.0); out.cin = T(0.0); out.mulvl = mulvl; out.zout = zout;
        return;
    }

    // ---- parcel at its initial location (:355-379) ---------------------
    Env<T> ek = env_at<T, MATH>(col, kmax);
    T th2, pi2, p2, t2, qv2, b2;
    if (source != SRC_ML) {
        th2 = ek.th; pi2 = ek.pi; p2 = ek.p; t2 = ek.t; qv2 = ek.q; b2 = T(0.0);
    } else {
        th2 = avgth;
        qv2 = avgqv;
        T thv2 = F::div(th2 * (T(1.0) + K::reps * qv2), T(1.0) + qv2);
        pi2 = ek.pi; p2 = ek.p;
        t2 = th2 * pi2;
        b2 = F::div(K::g * (thv2 - ek.thv), ek.thv);
    }
    T ql2 = T(0.0), qi2 = T(0.0), qt = qv2;
    T cape = T(0.0), cin = T(0.0), narea = T(0.0);
    T p_prev = ek.p, pi_prev = ek.pi, thv_prev = ek.thv;
    int k = kmax;
    bool doit = true;

    // ---- ascent (:403-559) -----------------------------------------------
    while (doit && k < nk - 1) {
        k = k + 1;
        Env<T> e = env_at<T, MATH>(col, k);
        const T b1 = b2;
        T dp = p_prev - e.p;
        int nloop;
        if (dp < pinc) {
            nloop = 1;
        } else {
            nloop = 1 + ftrunc(F::div(dp, pinc));
            dp = F::div(dp, (T)nloop);
        }
        st.levels++;
        for (int n = 1; n <= nloop; ++n) {
            const T p1 = p2, t1 = t2, th1 = th2, qv1 = qv2, ql1 = ql2, qi1 = qi2;
            p2 = p2 - dp;
            pi2 = F::pow(p2 * K::rp00, K::rddcp);
            st.substeps++;
            const T guess = ws.guess(th1, p2, qt, t1, pi2);
            if (!solve_substep<T, MATH, ICE>(p1, p2, pi2, t1, th1, qv1, ql1, qi1, qt, guess, th2,
                                             t2, qv2, ql2, qi2, st)) {
                st.nonconverged++;
                out.cape = T(0.0); out.cin = T(0.0); out.mulvl = mulvl; out.zout = zout;
                return;
            }
            if (PSEUDO) {
                qt = qv2;
                ql2 = T(0.0);
                qi2 = T(0.0);
            }
        }

        // buoyancy and the CAPE/CIN trapezoids (:501-547)
        const T thv2 = F::div(th2 * (T(1.0) + K::reps * qv2), T(1.0) + qv2 + ql2 + qi2);
        b2 = F::div(K::g * (thv2 - e.thv), e.thv);
        const T dz = layer_dz(e.thv, thv_prev, e.pi, pi_prev);
        T parea;
        if (b2 >= T(0.0) && b1 < T(0.0)) {
            const T frac = F::div(b2, b2 - b1);
            parea = T(0.5) * b2 * dz * frac;
            narea = narea - T(0.5) * b1 * dz * (T(1.0) - frac);
            cin = cin + narea;
            narea = T(0.0);
        } else if (b2 < T(0.0) && b1 > T(0.0)) {
            const T frac = F::div(b1, b1 - b2);
            parea = T(0.5) * b1 * dz * frac;
            narea = -T(0.5) * b2 * dz * (T(1.0) - frac);
        } else if (b2 < T(0.0)) {
            parea = T(0.0);
            narea = narea - T(0.5) * dz * (b1 + b2);
        } else {
            parea = T(0.5) * dz * (b1 + b2);
            narea = T(0.0);
        }
        cape = cape + fmx(T(0.0), parea);
        if (e.p <= T(10000.0) && b2 < T(0.0)) doit = false;
        zk = zk + dz;
        zout = zk;
        p_prev = e.p; pi_prev = e.pi; thv_prev = e.thv;
    }
    out.cape = cape; out.cin = cin; out.mulvl = mulvl; out.zout = zout;
}

}  // namespace xcape
