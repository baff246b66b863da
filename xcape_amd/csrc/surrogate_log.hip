// Training data for the warm-start surrogate, logged from the reference
// (cold-start) ascent.
//
// For a hashed subset of the converged sub-steps (or of the ascents' first
// sub-steps only, the ones a warm start has no history for) the logger
// writes one record (ln p2 [ln Pa], parcel theta-e at the sub-step's start
// [K], total water qt [kg/kg], rate (th2 - th1) / ln(p2/p1) [K], and the
// step ln(p2/p1)) to a device buffer, reserving slots with one atomic per
// record.  The parcel theta-e is the
// reference's own formula for the parcel, getthe(p, t, t, qv)
// (CAPE_CODE_model_lev.f90:505, which computes and discards it).
#include <hip/hip_runtime.h>

#include "../../include/xcape_hip.h"
#include "capi_common.h"
#include "cape_column.h"

using namespace xcape;

namespace {

struct LogCol {
    const float *p, *t, *td;
    float ps, ts, tds;
    __device__ __forceinline__ float P(int k) const { return k == 0 ? ps : p[k - 1]; }
    __device__ __forceinline__ float Tc(int k) const { return k == 0 ? ts : t[k - 1]; }
    __device__ __forceinline__ float Tdc(int k) const { return k == 0 ? tds : td[k - 1]; }
    __device__ __forceinline__ Raw raw_above(int k) const { return Raw{p[k - 1], t[k - 1], td[k - 1]}; }
};

struct SampleLogger {
    float *out;
    unsigned long long *count;
    unsigned long long cap;
    uint32_t keep;        // keep a sub-step when its hash's low 16 bits < keep
    uint32_t seed;
    int first_only;       // record only the ascent's first sub-step
    int64_t col;
    mutable uint32_t sub = 0;

    __device__ __forceinline__ void operator()(float p1, float t1, float th1, float qv1, float p2,
                                               float qt, float th2) const
    {
        uint32_t h = (uint32_t)col * 2654435761u ^ (sub * 2246822519u + seed);
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        ++sub;
        if (first_only && sub > 1) return;
        if ((h & 0xffffu) >= keep) return;
        const float lp = logf(p2 / p1);
        if (!(lp < 0.0f)) return;
        const unsigned long long i = atomicAdd(count, 1ull);
        if (i >= cap) return;
        const float the = thetae<float, MATH_ACCURATE>(p1, t1, t1, qv1);
        out[5 * i + 0] = logf(p2);
        out[5 * i + 1] = the;
        out[5 * i + 2] = qt;
        out[5 * i + 3] = (th2 - th1) / lp;
        out[5 * i + 4] = lp;
    }
};

template <bool ICE, bool PSEUDO>
__global__ void __launch_bounds__(256) log_kernel(const float *p, const float *t,
                                                  const float *td, const float *ps,
                                                  const float *ts, const float *tds, int nk_in,
                                                  int64_t ncol, float pinc, int source,
                                                  float ml_depth, SampleLogger lg)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol || !(ts[col] > 0.0f)) return;
    LogCol c{p + col * nk_in, t + col * nk_in, td + col * nk_in, ps[col], ts[col], tds[col]};
    LaneStats st;
    ColumnOut<float> o;
    lg.col = col;
    cape_column<float, MATH_ACCURATE, ICE, PSEUDO>(c, nk_in + 1, pinc, source, ml_depth, o, st,
                                                   lg);
}

}  // namespace

extern "C" int xcape_surrogate_log_f32(const float *p, const float *t, const float *td,
                                       const float *ps, const float *ts, const float *tds,
                                       int64_t nk, int64_t ncol, float pinc, int source,
                                       float ml_depth, int adiabat, float keep_fraction,
                                       uint32_t seed, int first_only, float *samples,
                                       int64_t capacity, unsigned long long *count, void *stream)
{
    if (nk < 1 || ncol < 1) return fail(XCAPE_EINVAL, "empty grid");
    if (!p || !t || !td || !ps || !ts || !tds || !samples || !count)
        return fail(XCAPE_EINVAL, "NULL argument");
    if (!(pinc >= 1.0f)) return fail(XCAPE_EINVAL, "pinc must be >= 1 Pa");
    if (source < 1 || source > 3 || adiabat < 1 || adiabat > 4)
        return fail(XCAPE_EINVAL, "source/adiabat out of range");
    if (!(keep_fraction > 0.0f && keep_fraction <= 1.0f))
        return fail(XCAPE_EINVAL, "keep_fraction must be in (0, 1]");
    int rc = ensure_device();
    if (rc) return rc;
    SampleLogger lg{samples, count, (unsigned long long)capacity,
                    (uint32_t)(keep_fraction * 65536.0f), seed, first_only ? 1 : 0, 0};
    if (keep_fraction >= 1.0f) lg.keep = 0x10000u;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((unsigned)((ncol + 255) / 256));
    switch (adiabat) {
    case 1: log_kernel<false, true><<<grid, 256, 0, s>>>(p, t, td, ps, ts, tds, (int)nk, ncol, pinc, source, ml_depth, lg); break;
    case 2: log_kernel<false, false><<<grid, 256, 0, s>>>(p, t, td, ps, ts, tds, (int)nk, ncol, pinc, source, ml_depth, lg); break;
    case 3: log_kernel<true, true><<<grid, 256, 0, s>>>(p, t, td, ps, ts, tds, (int)nk, ncol, pinc, source, ml_depth, lg); break;
    default: log_kernel<true, false><<<grid, 256, 0, s>>>(p, t, td, ps, ts, tds, (int)nk, ncol, pinc, source, ml_depth, lg); break;
    }
    hipError_t e = hipGetLastError();
    return e ? hip_fail(e, "surrogate log") : XCAPE_OK;
}
