// One column of the storm-relative-helicity chain on one lane:
//   heights above ground (FP64 hypsometric equation, stdheight_2D_*_lev.f90),
//   Bunkers right/left-mover motion and 0-6 km mean wind (FP32,
//   Bunkers_*_lev.f90), storm-relative helicity 0..depth (FP64, SREH_*_lev.f90),
// with the reference's precisions and operation order.
//
// Organisation (same results as the reference's three passes over stored
// arrays, without any per-column scratch):
//  * pass 1 walks the levels bottom-up, building H(k) and, for each of the 12
//    Bunkers target heights, the interpolant of the *last* bracketing layer
//    seen.  The reference searches top-down and stops at the first bracket
//    (Bunkers_model_lev.f90:218-231), which is the last one bottom-up.  While
//    the heights rise each target is bracketed once, in order, so the walk
//    keeps only the next target and the reference's in-order sums; a column
//    whose heights fall somewhere, or whose heights fall with k throughout
//    (its top below its surface: IP = 1 in the reference), is walked again
//    with the reference's arrays (bunkers_full).
//  * the same walk sums the helicity terms up to the first level above
//    `depth` with the storm motion factored out (it is known only once the
//    walk has ended); it enters the sum at the end.
//
// Literal constants of the reference's DOUBLE PRECISION routines are default
// REAL literals, so their values are those of the nearest float (SPD below).
// The heights' two logarithms and one exponential per level are log_core_f64
// and exp_core_f64 (phys.h), within an ulp like the C library's; the logs
// were half of the chain's instructions with the device library's log.  Their
// divisions are the raw FP64 kernel (IEEE quotients, re-evaluated with IEEE
// division where a level's result is not finite).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phys.h"

namespace xcape {

#define XC_SPD(x) ((double)(float)(x))

// TB: the storm-motion precision, float as in the reference (Bunkers is REAL
// there) or double for the FP64 re-run
template <typename TB> struct SrhOutT {
    double srh_rm, srh_lm;
    TB rm_u, rm_v, lm_u, lm_v, mean_u, mean_v;
};
using SrhOut = SrhOutT<float>;

// FP64 quotient: the raw division kernel (RAW) or IEEE division
template <bool RAW> __device__ __forceinline__ double sdiv(double a, double b)
{
    if constexpr (RAW) return ddiv_raw(a, b);
    else return a / b;
}

// virtual temperature of one level (stdheight_2D_model_lev.f90:102-127), FP64
template <bool RAW> __device__ __forceinline__ double std_tv_ops(double T, double Td, double P)
{
#pragma clang fp contract(off)
    const double eps = XC_SPD(0.6219800858985514);
    double Tin = T + XC_SPD(273.15);
    double Tdin = Td + XC_SPD(273.15);
    double E = XC_SPD(6.112) * exp_core_f64((XC_SPD(53.49) - sdiv<RAW>(6808.0, Tdin)) -
                                            XC_SPD(5.09) * log_core_f64(Tdin));
    double w = eps * sdiv<RAW>(E, P - E);
    return Tin * sdiv<RAW>(w + eps, eps * (1.0 + w));
}

// with the raw divisions, the IEEE ones where the result is not finite (the
// raw kernel's NaN for a zero or infinite divisor): IEEE results throughout
__device__ __forceinline__ double std_tv(double T, double Td, double P)
{
    const double tv = std_tv_ops<true>(T, Td, P);
    if (__builtin_expect(__builtin_isfinite(tv), 1)) return tv;
    return std_tv_ops<false>(T, Td, P);
}

// std_tv's value to within about 1e-3 relative, in FP32 with the hardware
// exp2/log2/reciprocal: only its sign is used (a level's height increment is
// positive when its pressure is below the last one's and both virtual
// temperatures are positive)
__device__ __forceinline__ float std_tv_sign(float T, float Td, float P)
{
    const float eps = 0.6219800858985514f;
    const float Tin = T + 273.15f;
    const float Tdin = Td + 273.15f;
    const float lg = __builtin_amdgcn_logf(Tdin) * 0.6931471805599453f;
    const float ex = (53.49f - 6808.0f * __builtin_amdgcn_rcpf(Tdin)) - 5.09f * lg;
    const float E = 6.112f * __builtin_amdgcn_exp2f(ex * 1.4426950408889634f);
    const float w = eps * E * __builtin_amdgcn_rcpf(P - E);
    return Tin * (w + eps) * __builtin_amdgcn_rcpf(eps * (1.0f + w));
}

// H(k) = H(k-1) + (R ((Tv(k) + Tv(k-1))/2) / g) ln(P(k)/P(k-1)) (:147, :155)
template <bool RAW>
__device__ __forceinline__ double std_dh_ops(double tv, double tv_prev, double P, double P_prev)
{
#pragma clang fp contract(off)
    const double R = XC_SPD(287.04);
    const double g = XC_SPD(-9.80665);
    return sdiv<RAW>(R * ((tv + tv_prev) / 2.0), g) * log_core_f64(sdiv<RAW>(P, P_prev));
}

__device__ __forceinline__ double std_dh(double tv, double tv_prev, double P, double P_prev)
{
    const double dh = std_dh_ops<true>(tv, tv_prev, P, P_prev);
    if (__builtin_expect(__builtin_isfinite(dh), 1)) return dh;
    return std_dh_ops<false>(tv, tv_prev, P, P_prev);
}

// The Bunkers inputs (Bunkers_model_lev.f90:123-177): the winds at the 12
// target heights h_i = 500 (i + 1) m (-999999 where no layer brackets h_i),
// as the sums su = ((0 + u0) + u_0) + ... + u_11 in index order (the
// reference's accumulation) and the three the shear uses.
template <typename TB> struct BunkersIn {
    TB su, sv, u0, v0, u10, v10, u11, v11;
};

template <typename TB> __device__ __forceinline__ TB bunkers_h(int i)
{
    return TB(500 * (i + 1));
}

// The Bunkers inputs by the reference's own method: the winds interpolated
// in every bracketing layer of a walk up the column, the last one bottom-up
// kept (its top-down search stops at the first, :218-231); `decreasing`:
// heights falling with k (IP = 1), the bracket z(k) <= h < z(k-1).  Walks
// the heights again; for the columns whose heights do not rise steadily.
template <typename TB, class C>
__device__ __noinline__ void bunkers_full(const C &col, int nk, double hs, double tv0, double Ps,
                                          bool decreasing, BunkersIn<TB> &b)
{
#pragma clang fp contract(off)
    TB iu[12], iv[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) { iu[i] = TB(-999999.0); iv[i] = TB(-999999.0); }
    double tv_prev = tv0, P_prev = Ps, H = hs;
    TB z_prev = (TB)hs, u_prev = (TB)col.U(0), v_prev = (TB)col.V(0);
    for (int k = 1; k < nk; ++k) {
        const double Pk = (double)col.P(k);
        const double tv = std_tv((double)col.Tc(k), (double)col.Tdc(k), Pk);
        H = H + std_dh(tv, tv_prev, Pk, P_prev);
        const TB z = (TB)H;
        const TB uk = (TB)col.U(k), vk = (TB)col.V(k);
        // the layer's lower and upper ends in the orientation's order
        const TB zl = decreasing ? z : z_prev, zh = decreasing ? z_prev : z;
        const TB ul = decreasing ? uk : u_prev, uh = decreasing ? u_prev : uk;
        const TB vl = decreasing ? vk : v_prev, vh = decreasing ? v_prev : vk;
#pragma unroll
        for (int i = 0; i < 12; ++i) {
            const TB h = bunkers_h<TB>(i);
            if (zl <= h && zh > h) {
                const TB w2 = (h - zl) / (zh - zl);
                const TB w1 = (TB)(1.0 - (double)w2);
                iu[i] = w1 * ul + w2 * uh;
                iv[i] = w1 * vl + w2 * vh;
            }
        }
        z_prev = z; u_prev = uk; v_prev = vk;
        tv_prev = tv; P_prev = Pk;
    }
    b.su = TB(0) + (TB)col.U(0);
    b.sv = TB(0) + (TB)col.V(0);
#pragma unroll
    for (int i = 0; i < 12; ++i) { b.su = b.su + iu[i]; b.sv = b.sv + iv[i]; }
    b.u0 = iu[0]; b.v0 = iv[0];
    b.u10 = iu[10]; b.v10 = iv[10];
    b.u11 = iu[11]; b.v11 = iv[11];
}

// Column accessor C: P(k), Tc(k), Tdc(k), U(k), V(k) for k = 0..nk-1 with
// k = 0 the surface (hPa, degC, m/s).  hs: surface height (2 m in calc_srh).
// TB = float is the reference's arithmetic; TB = double evaluates the storm
// motion (and the helicity's storm-relative winds) in FP64 as well.
template <typename TB, class C>
__device__ void srh_column(const C &col, int nk, double depth, double hs, SrhOutT<TB> &out)
{
#pragma clang fp contract(off)
    // ---- pass 1: heights and Bunkers brackets (increasing orientation),
    // and the helicity sum's storm-motion-free part ----
    const double Ps = (double)col.P(0);
    const double tv0 = std_tv((double)col.Tc(0), (double)col.Tdc(0), Ps);
    const TB zsurf = (TB)hs;
    const TB u0 = (TB)col.U(0), v0 = (TB)col.V(0);
    // SREH 0..depth (SREH_model_lev.f90:82-114) sums, over the levels up to
    // the first one above `depth` (its wind interpolated to `depth`),
    // (u - cu)(v - v') - (v - cv)(u - u') with c the storm motion, known only
    // after this walk.  The sum splits into s0 = sum u(v - v') - v(u - u')
    // and the telescoping sums of (v - v') and (u - u'), v_top - v_0 and
    // u_top - u_0, so c enters at the end and the heights are computed once.
    // FP64; the split moves the sum by rounding only.
    //
    // Bunkers brackets: while the heights do not fall (`mono`), each target
    // height h_i is bracketed once, in the order of i, so the walk keeps the
    // index nb of the next one and the reference's in-order sums as it goes;
    // a column whose heights fall anywhere (or are NaN) gets them from
    // bunkers_full, as does one whose heights decrease with k.
    //
    // No output depends on a level above both the helicity's top and the
    // last Bunkers level (6 km) but through the orientation test after the
    // walk (the height of the column's last level against the surface's):
    // past them the first walk only checks that the heights keep rising
    // (each level's pressure below the last one's, its virtual temperature
    // positive, to a 1 K margin), which settles that test, and skips their
    // FP64 logs and exp.  Where the check fails the column is walked again
    // in full, so the outputs are those of the full walk in every case.
    double tv_prev, P_prev, H, s0, z1prev, ut_prev, vt_prev;
    TB z_prev, u_prev, v_prev;
    bool reached, rising = false, mono;
    int nb;
    BunkersIn<TB> b;
    for (int walk = 0; walk < 2; ++walk) {
        tv_prev = tv0; P_prev = Ps; H = hs;
        z_prev = zsurf; u_prev = u0; v_prev = v0;
        s0 = 0.0; z1prev = hs; ut_prev = (double)col.U(0); vt_prev = (double)col.V(0);
        reached = false;
        rising = false;
        mono = true;
        b.su = TB(0) + u0; b.sv = TB(0) + v0;
        b.u0 = b.v0 = b.u10 = b.v10 = b.u11 = b.v11 = TB(-999999.0);
        // target heights below the surface are never bracketed by rising heights
        nb = 0;
        while (nb < 12 && bunkers_h<TB>(nb) < zsurf) {
            b.su = b.su + TB(-999999.0);
            b.sv = b.sv + TB(-999999.0);
            ++nb;
        }
        bool failed = false;
        float p_last = 0.0f;
        for (int k = 1; k < nk; ++k) {
            if (rising) {
                const float pk = col.P(k);
                if (pk < p_last && std_tv_sign(col.Tc(k), col.Tdc(k), pk) > 1.0f) {
                    p_last = pk;
                    continue;
                }
                failed = true;
                break;
            }
            const double Pk = (double)col.P(k);
            const double tv = std_tv((double)col.Tc(k), (double)col.Tdc(k), Pk);
            H = H + std_dh(tv, tv_prev, Pk, P_prev);
            const TB z = (TB)H;
            // Bunkers reads the winds in its own precision; the helicity in
            // FP64 from the inputs as given (both as the reference's wrappers
            // pass them)
            const TB uk = (TB)col.U(k), vk = (TB)col.V(k);
            if (mono) {
                if (!(z >= z_prev)) {
                    mono = false;
                } else {
                    // the targets h_nb, h_nb+1, ... below z: z_prev <= h_nb holds
                    while (nb < 12) {
                        const TB h = bunkers_h<TB>(nb);
                        if (!(h < z)) break;
                        const TB w2 = (h - z_prev) / (z - z_prev);
                        const TB w1 = (TB)(1.0 - (double)w2);
                        const TB bu = w1 * u_prev + w2 * uk, bv = w1 * v_prev + w2 * vk;
                        b.su = b.su + bu;
                        b.sv = b.sv + bv;
                        if (nb == 0) { b.u0 = bu; b.v0 = bv; }
                        if (nb == 10) { b.u10 = bu; b.v10 = bv; }
                        if (nb == 11) { b.u11 = bu; b.v11 = bv; }
                        ++nb;
                    }
                }
            }
            if (!reached) {
                double ut = (double)col.U(k), vt = (double)col.V(k);
                const bool top = H > depth;
                if (top) {
                    // INTERP1 (SREH_model_lev.f90:123-129): X1 = aglh(ktop) is
                    // lowered by 0.01 (a REAL literal) when it equals aglh(ktop-1)
                    double x1 = H;
                    if (z1prev == x1) x1 = x1 - XC_SPD(0.01);
                    ut = ut + ((ut_prev - ut) * ((depth - x1) / (z1prev - x1)));
                    vt = vt + ((vt_prev - vt) * ((depth - x1) / (z1prev - x1)));
                }
                s0 = s0 + ((ut * (vt - vt_prev)) - (vt * (ut - ut_prev)));
                reached = top;
                // past the top, (ut_prev, vt_prev) keep its interpolated wind
                z1prev = H; ut_prev = ut; vt_prev = vt;
            }
            z_prev = z; u_prev = uk; v_prev = vk;
            tv_prev = tv; P_prev = Pk;
            // every output's levels behind and this level above the surface:
            // from here on the first walk checks the rest only
            if (walk == 0 && reached && z > bunkers_h<TB>(11) && z > zsurf && tv > 1.0) {
                rising = true;
                p_last = (float)Pk;
            }
        }
        if (!failed) break;
    }
    // heights decreasing with k (IP = 1, IM = 0 in the reference); a walk that
    // went on rising ends above the surface, so the test is false for it
    const bool decreasing = !rising && zsurf > z_prev;
    if (decreasing || !mono) {
        bunkers_full<TB>(col, nk, hs, tv0, Ps, decreasing, b);
    } else {
        // the targets no layer reached, in index order
        for (int i = nb; i < 12; ++i) {
            b.su = b.su + TB(-999999.0);
            b.sv = b.sv + TB(-999999.0);
        }
    }

    // ---- Bunkers (Bunkers_model_lev.f90:123-177), accumulators from zero ----
    const TB mu = b.su / TB(13);
    const TB mv = b.sv / TB(13);
    const TB ushr_u = ((TB(0) + b.u10) + b.u11) / TB(2);
    const TB vshr_u = ((TB(0) + b.v10) + b.v11) / TB(2);
    const TB ushr_d = ((TB(0) + u0) + b.u0) / TB(2);
    const TB vshr_d = ((TB(0) + v0) + b.v0) / TB(2);
    const TB ushr = ushr_u - ushr_d;
    const TB vshr = vshr_u - vshr_d;
    TB mag;
    if constexpr (sizeof(TB) == 4) mag = sqrtf(ushr * ushr + vshr * vshr);
    else mag = sqrt(ushr * ushr + vshr * vshr);
    out.rm_u = mu + TB(7.5) * vshr / mag;
    out.rm_v = mv - TB(7.5) * ushr / mag;
    out.lm_u = mu - TB(7.5) * vshr / mag;
    out.lm_v = mv + TB(7.5) * ushr / mag;
    out.mean_u = mu;
    out.mean_v = mv;

    // ---- SREH: the storm motions enter the sums of pass 1 ----
    // no level above depth: the reference's sum is empty (ktop = 0)
    double srm = 0.0, slm = 0.0;
    if (reached) {
        const double du = ut_prev - (double)col.U(0), dv = vt_prev - (double)col.V(0);
        srm = (s0 - (double)out.rm_u * dv) + (double)out.rm_v * du;
        slm = (s0 - (double)out.lm_u * dv) + (double)out.lm_v * du;
    }
    out.srh_rm = -srm;
    out.srh_lm = -slm;
}

}  // namespace xcape
