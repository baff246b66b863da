// xcape_math_check: the accurate mode's device arithmetic, elementwise, so
// that its rounding can be checked against IEEE division and the C library
// on the device it runs on (the kernels' results depend on these agreeing).
//
// fn: 0 div_raw(a, b)   1 a / b (the compiler's IEEE division)
//     2 div_const(a, b, RN(1/b))   3 exp_core(a)   4 expf(a)
//     5 pow_core(a, b)   6 powf(a, b)
// xcape_math_check_f64: fn 0 log_core_f64(a)   1 log(a)   2 exp_core_f64(a)
//     3 exp(a)   4 ddiv_raw(a0, a1)   5 a0 / a1 (FP64; a holds the pairs)
#include <hip/hip_runtime.h>

#include "../../include/xcape_hip.h"
#include "capi_common.h"
#include "phys.h"

using namespace xcape;

namespace {

__global__ void math_check_kernel(int fn, const float *a, const float *b, int64_t n, float *y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = a[i];
    const float z = b ? b[i] : 0.0f;
    float r;
    switch (fn) {
    case 0: r = div_raw(x, z); break;
    case 1: r = x / z; break;
    case 2: r = div_const(x, z, 1.0f / z); break;
    case 3: r = exp_core(x); break;
    case 4: r = expf(x); break;
    case 5: r = pow_core(x, z); break;
    default: r = powf(x, z); break;
    }
    y[i] = r;
}

__global__ void math_check_f64_kernel(int fn, const double *a, int64_t n, double *y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fn >= 4) {
        // binary: the operands are the pairs (a[2i], a[2i + 1])
        const double u = a[2 * i], v = a[2 * i + 1];
        y[i] = fn == 4 ? ddiv_raw(u, v) : u / v;
        return;
    }
    const double x = a[i];
    y[i] = fn == 0 ? log_core_f64(x) : (fn == 1 ? log(x) : (fn == 2 ? exp_core_f64(x) : exp(x)));
}

}  // namespace

extern "C" int xcape_math_check(int fn, const float *a, const float *b, int64_t n, float *y,
                                void *stream)
{
    if (fn < 0 || fn > 6) return fail(XCAPE_EINVAL, "fn=%d not in 0..6", fn);
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(XCAPE_EINVAL, "n=%lld out of range", (long long)n);
    const bool binary = fn <= 2 || fn >= 5;
    if (!a || !y || (binary && !b)) return fail(XCAPE_EINVAL, "NULL argument");
    int rc = ensure_device();
    if (rc) return rc;
    if (n == 0) return XCAPE_OK;
    math_check_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(fn, a, b, n, y);
    const hipError_t e = hipGetLastError();
    return e ? hip_fail(e, "math_check") : XCAPE_OK;
}

extern "C" int xcape_math_check_f64(int fn, const double *a, int64_t n, double *y, void *stream)
{
    if (fn < 0 || fn > 5) return fail(XCAPE_EINVAL, "fn=%d not in 0..5", fn);
    if (n < 0 || n > ((int64_t)1 << 31)) return fail(XCAPE_EINVAL, "n=%lld out of range", (long long)n);
    if (!a || !y) return fail(XCAPE_EINVAL, "NULL argument");
    int rc = ensure_device();
    if (rc) return rc;
    if (n == 0) return XCAPE_OK;
    math_check_f64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(fn, a, n, y);
    const hipError_t e = hipGetLastError();
    return e ? hip_fail(e, "math_check_f64") : XCAPE_OK;
}
