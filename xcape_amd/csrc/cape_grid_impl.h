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
                                                unsigned long long *stats, int nonconv)
{
    __shared__ unsigned long long red[5][16];
    unsigned long long v[5] = {(unsigned long long)ran, st.levels, st.substeps, st.iterations,
                               (unsigned long long)nonconv};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        unsigned long long x = v[c];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[c][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
        if (s) atomicAdd(&stats[threadIdx.x], s);
    }
}

// KIND: GRID_CAPE, GRID_SRH or GRID_CAPE_SRH (both from one read of the
// column: the helicity chain runs after the ascent on the same lane).
template <typename T, bool ICE, bool PSEUDO, int KIND>
__global__ void __launch_bounds__(256) cape_grid_kernel(CapeGridArgs a)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    LaneStats st;
    bool ran = false;
    if (col < a.ncol) {
        SigmaCol c;
        int nk;
        const float ps = a.ps[col];
        int64_t off = col * a.nk_in;
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
        if (KIND & GRID_CAPE) {
            T *cape = static_cast<T *>(a.cape), *cin = static_cast<T *>(a.cin),
              *zout = static_cast<T *>(a.zout);
            if (c.ts > 0.0f) {
                ran = true;
                ColumnOut<T> o;
                cape_column<T, XC_MATH, ICE, PSEUDO>(c, nk, T(a.pinc), a.source, T(a.ml_depth),
                                                     o, st);
                cape[col] = o.cape;
                cin[col] = o.cin;
                a.mulvl[col] = o.mulvl;
                zout[col] = o.zout;
            } else {
                cape[col] = T(0);
                cin[col] = T(0);
                zout[col] = T(0);
                a.mulvl[col] = 0;
            }
        }
        if (KIND & GRID_SRH) {
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
    if ((KIND & GRID_CAPE) && a.stats) block_stats_add(st, ran, a.stats, (int)st.nonconverged);
}

template <typename T, int KIND>
hipError_t launch_kind(const CapeGridArgs &a, int block, hipStream_t s)
{
    dim3 grid((unsigned)((a.ncol + block - 1) / block));
    if (KIND == GRID_SRH) {
        cape_grid_kernel<T, false, true, GRID_SRH><<<grid, block, 0, s>>>(a);
        return hipGetLastError();
    }
    switch (a.adiabat) {
    case 1: cape_grid_kernel<T, false, true, KIND><<<grid, block, 0, s>>>(a); break;
    case 2: cape_grid_kernel<T, false, false, KIND><<<grid, block, 0, s>>>(a); break;
    case 3: cape_grid_kernel<T, true, true, KIND><<<grid, block, 0, s>>>(a); break;
    case 4: cape_grid_kernel<T, true, false, KIND><<<grid, block, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_t(const CapeGridArgs &a, int kind, int block, hipStream_t s)
{
    if (a.ncol <= 0) return hipSuccess;
    if constexpr (sizeof(T) == 8) {
        // the FP64 yardstick covers CAPE only
        return kind == GRID_CAPE ? launch_kind<T, GRID_CAPE>(a, block, s) : hipErrorInvalidValue;
    }
    switch (kind) {
    case GRID_CAPE: return launch_kind<T, GRID_CAPE>(a, block, s);
    case GRID_SRH: return launch_kind<T, GRID_SRH>(a, block, s);
    case GRID_CAPE_SRH: return launch_kind<T, GRID_CAPE_SRH>(a, block, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace XC_NS
}  // namespace xcape
