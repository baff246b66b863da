// Rectangular-grid CAPE kernels (sigma or shared pressure levels), included
// by one translation unit per math mode with XC_MATH defined.
//
// Mapping: one column per lane, 64 columns per wavefront, blockDim columns
// per workgroup.  A column's levels are contiguous in HBM (the reference's
// (nlev, ncol) Fortran-order layout), so a lane walks its own column; the
// lines it touches stay in L1/L2 across the ~16 levels each 64-byte segment
// holds, and HBM traffic stays at the compulsory bytes.
#pragma once
#include "cape_column.h"
#include "cape_launch.h"
#include "mlp_device.h"
#include "srh_column.h"

namespace xcape {
namespace XC_NS {

struct SigmaCol {
    const float *p, *t, *td, *u, *v;
    float ps, ts, tds, us, vs;
    __device__ __forceinline__ float P(int k) const { return k == 0 ? ps : p[k - 1]; }
    __device__ __forceinline__ float Tc(int k) const { return k == 0 ? ts : t[k - 1]; }
    __device__ __forceinline__ float Tdc(int k) const { return k == 0 ? tds : td[k - 1]; }
    __device__ __forceinline__ float U(int k) const { return k == 0 ? us : u[k - 1]; }
    __device__ __forceinline__ float V(int k) const { return k == 0 ? vs : v[k - 1]; }
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
    constexpr int NC = 6;
    __shared__ unsigned long long red[NC][16];
    unsigned long long v[NC] = {(unsigned long long)ran, st.levels, st.substeps, st.iterations,
                                st.nonconverged, st.fallbacks};
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

// KIND: GRID_CAPE, GRID_SRH or GRID_CAPE_SRH (both from one read of the
// column: the helicity chain runs after the ascent on the same lane).
// WARM: warm-started ascent; with a surrogate loaded (a.surrogate), the first
// sub-step's iterate of every lane comes from one wave-wide MLP evaluation
// (MFMA), done at a point every lane of the block reaches.
template <typename T, bool ICE, bool PSEUDO, int KIND, bool WARM>
__global__ void __launch_bounds__(256) cape_grid_kernel(CapeGridArgs a)
{
    extern __shared__ uint32_t dyn_lds[];
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    LaneStats st;
    bool ran = false;
    const bool in_grid = col < a.ncol;
    SigmaCol c{};
    int nk = 0;
    int64_t off = 0;
    if (in_grid) {
        const float ps = a.ps[col];
        off = col * a.nk_in;
        if (a.p1d) {
            int s1 = a.start ? a.start[col] : start_level(a.p1d, (int)a.nk_in, ps);
            // a start level outside 1..nk_in cannot come from the rule above;
            // clamp caller-supplied ones so the column is never read outside
            s1 = s1 < 1 ? 1 : (s1 > (int)a.nk_in ? (int)a.nk_in : s1);
            const int s0 = s1 - 1;
            c.p = a.p1d + s0;
            off += s0;
            nk = (int)a.nk_in - s0 + 1;
        } else {
            c.p = a.p + off;
            nk = (int)a.nk_in + 1;
        }
        c.t = a.t + off;
        c.td = a.td + off;
        c.ps = ps;
        c.ts = a.ts[col];
        c.tds = a.tds[col];
    }
    if constexpr ((KIND & GRID_CAPE) != 0) {
        T *cape = static_cast<T *>(a.cape), *cin = static_cast<T *>(a.cin),
          *zout = static_cast<T *>(a.zout);
        const bool lane_cape = in_grid && c.ts > 0.0f;
        if constexpr (WARM) {
            using F = Fn<T, XC_MATH>;
            using K = Phys<T>;
            Parcel<T> pc;
            bool ok = false;
            T guess = T(NAN), lp1 = T(0.0);
            float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
            if (lane_cape) {
                ok = parcel_source<T, XC_MATH>(c, nk, a.source, T(a.ml_depth), pc);
                if (ok && a.surrogate && pc.kmax < nk - 1) {
                    // the first sub-step: pc.p2 -> pc.p2 - dp (:408-429)
                    const Env<T> e1 = env_at<T, XC_MATH>(c, pc.kmax + 1);
                    T dp = pc.e.p - e1.p;
                    const int nl = substep_count<T, XC_MATH>(dp, T(a.pinc));
                    const T p2 = pc.p2 - dp;
                    if (nl >= 1 && p2 > T(0.0)) {
                        lp1 = F::log(F::div(p2, pc.p2));
                        f0 = (float)F::log(p2);
                        f1 = (float)thetae<T, XC_MATH>(pc.p2, pc.t2, pc.t2, pc.qv2);
                        f2 = (float)pc.qv2;
                    }
                }
            }
            if (a.surrogate) {
                // uniform across the block: every lane takes part
                const float r = surrogate_eval_wave(
                    static_cast<const SurrogateDev *>(a.surrogate), f0, f1, f2,
                    dyn_lds + (threadIdx.x >> 6) * MLP_LDS_WORDS_PER_WAVE);
                if (lp1 != T(0.0)) guess = pc.th2 + T(r) * lp1;
            }
            (void)K::g;
            if (lane_cape) {
                ran = true;
                ColumnOut<T> o;
                if (ok) {
                    cape_column_warm<T, XC_MATH, ICE, PSEUDO>(c, nk, T(a.pinc), pc, guess,
                                                              T(a.fallback_tol), o, st);
                } else {
                    o.cape = T(0); o.cin = T(0); o.mulvl = pc.mulvl; o.zout = T(-999999.0);
                }
                cape[col] = o.cape;
                cin[col] = o.cin;
                a.mulvl[col] = o.mulvl;
                zout[col] = o.zout;
            }
        } else {
            if (lane_cape) {
                ran = true;
                ColumnOut<T> o;
                cape_column<T, XC_MATH, ICE, PSEUDO>(c, nk, T(a.pinc), a.source, T(a.ml_depth),
                                                     o, st);
                cape[col] = o.cape;
                cin[col] = o.cin;
                a.mulvl[col] = o.mulvl;
                zout[col] = o.zout;
            }
        }
        if (in_grid && !lane_cape) {
            cape[col] = T(0);
            cin[col] = T(0);
            zout[col] = T(0);
            a.mulvl[col] = 0;
        }
    }
    if constexpr ((KIND & GRID_SRH) != 0) {
        if (in_grid) {
            c.u = a.u + off;
            c.v = a.v + off;
            c.us = a.us[col];
            c.vs = a.vs[col];
            SrhOut so;
            srh_column(c, nk, a.depth, 2.0, so);
            a.srh_rm[col] = so.srh_rm;
            a.srh_lm[col] = so.srh_lm;
            if (a.rm) {
                a.rm[2 * col] = so.rm_u;
                a.rm[2 * col + 1] = so.rm_v;
                a.lm[2 * col] = so.lm_u;
                a.lm[2 * col + 1] = so.lm_v;
                a.mean6[2 * col] = so.mean_u;
                a.mean6[2 * col + 1] = so.mean_v;
            }
        }
    }
    if ((KIND & GRID_CAPE) && a.stats) block_stats_add(st, ran, a.stats);
}

template <typename T, int KIND, bool WARM>
hipError_t launch_kind(const CapeGridArgs &a, int block, hipStream_t s)
{
    dim3 grid((unsigned)((a.ncol + block - 1) / block));
    const size_t lds = (WARM && a.surrogate)
                           ? sizeof(uint32_t) * MLP_LDS_WORDS_PER_WAVE * (size_t)(block / 64)
                           : 0;
    if (KIND == GRID_SRH) {
        cape_grid_kernel<T, false, true, GRID_SRH, false><<<grid, block, 0, s>>>(a);
        return hipGetLastError();
    }
    switch (a.adiabat) {
    case 1: cape_grid_kernel<T, false, true, KIND, WARM><<<grid, block, lds, s>>>(a); break;
    case 2: cape_grid_kernel<T, false, false, KIND, WARM><<<grid, block, lds, s>>>(a); break;
    case 3: cape_grid_kernel<T, true, true, KIND, WARM><<<grid, block, lds, s>>>(a); break;
    case 4: cape_grid_kernel<T, true, false, KIND, WARM><<<grid, block, lds, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_t(const CapeGridArgs &a, int kind, int block, hipStream_t s)
{
    if (a.ncol <= 0) return hipSuccess;
    if constexpr (sizeof(T) == 8) {
        // the FP64 yardstick covers CAPE with the reference's cold start only
        if (kind != GRID_CAPE || a.warm) return hipErrorInvalidValue;
        return launch_kind<T, GRID_CAPE, false>(a, block, s);
    } else {
        switch (kind) {
        case GRID_CAPE:
            return a.warm ? launch_kind<T, GRID_CAPE, true>(a, block, s)
                          : launch_kind<T, GRID_CAPE, false>(a, block, s);
        case GRID_SRH: return launch_kind<T, GRID_SRH, false>(a, block, s);
        case GRID_CAPE_SRH:
            return a.warm ? launch_kind<T, GRID_CAPE_SRH, true>(a, block, s)
                          : launch_kind<T, GRID_CAPE_SRH, false>(a, block, s);
        default: return hipErrorInvalidValue;
        }
    }
}

}  // namespace XC_NS
}  // namespace xcape
