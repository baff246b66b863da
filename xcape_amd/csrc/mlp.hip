// Training and evaluation kernels of the warm-start surrogate, and the
// library's resident copy of the loaded network.
//
// Training step (all FP32 accumulation; BF16 rounding where the inference
// kernel rounds, so the trained network is the one the ascent evaluates):
//   mlp_forward_store  one thread per sample: z1 = W1 u + b1, h1 = bf16(gelu(z1)),
//                      z2 = W2 h1 + b2, h2 = gelu(z2), y = w3 . h2 + b3
//   mlp_backward_rows  one thread per sample: dy = 2 (y - t) / n,
//                      dz2 = dy w3 gelu'(z2), dz1 = (W2^T dz2) gelu'(z1)
//   mlp_weight_grads   one block per hidden unit: sums over the batch of the
//                      per-sample outer products (fixed order: deterministic)
//   mlp_adam           Adam on the FP32 master vector
// The gradient is the straight-through one: the BF16 rounding is treated as
// the identity in the backward pass.
#include <hip/hip_runtime.h>
#include <math.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/xcape_hip.h"
#include "capi_common.h"
#include "mlp_device.h"

using namespace xcape;

namespace {

constexpr int H = MLP_H;

// weights as the forward uses them: BF16-rounded W1, W2, w3; FP32 biases
struct alignas(16) FwdWeights {
    float w1[H * 3], b1[H], w2[H * H], b2[H], w3[H], b3;
};

__device__ __forceinline__ void load_weights(const float *theta, FwdWeights &w)
{
    for (int i = threadIdx.x; i < H * 3; i += blockDim.x) w.w1[i] = bf16_round(theta[MLP_OFF_W1 + i]);
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) w.w2[i] = bf16_round(theta[MLP_OFF_W2 + i]);
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        w.b1[i] = theta[MLP_OFF_B1 + i];
        w.b2[i] = theta[MLP_OFF_B2 + i];
        w.w3[i] = bf16_round(theta[MLP_OFF_W3 + i]);
    }
    if (threadIdx.x == 0) w.b3 = theta[MLP_OFF_B3];
    __syncthreads();
}

// forward for one sample; optionally stores z1, h1 (BF16-rounded), z2, h2
__device__ __forceinline__ float forward_one(const FwdWeights &w, const float *u, float *z1s,
                                             float *h1s, float *z2s, float *h2s)
{
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const float z = w.b1[j] + w.w1[3 * j] * u[0] + w.w1[3 * j + 1] * u[1] + w.w1[3 * j + 2] * u[2];
        h1[j] = bf16_round(gelu_f(z));
        if (z1s) { z1s[j] = z; h1s[j] = h1[j]; }
    }
    float y = w.b3;
    for (int i = 0; i < H; ++i) {
        float z = w.b2[i];
#pragma unroll
        for (int j = 0; j < H; ++j) z += w.w2[i * H + j] * h1[j];
        const float h = gelu_f(z);
        if (z2s) { z2s[i] = z; h2s[i] = h; }
        y += w.w3[i] * h;
    }
    return y;
}

__global__ void __launch_bounds__(256) mlp_forward_store(const float *x, int64_t n,
                                                         const float *theta, float *y, float *Z1,
                                                         float *H1, float *Z2, float *H2)
{
    __shared__ FwdWeights w;
    load_weights(theta, w);
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float u[3] = {x[3 * s], x[3 * s + 1], x[3 * s + 2]};
    if (Z1)
        y[s] = forward_one(w, u, Z1 + s * H, H1 + s * H, Z2 + s * H, H2 + s * H);
    else
        y[s] = forward_one(w, u, nullptr, nullptr, nullptr, nullptr);
}

__global__ void __launch_bounds__(256) mlp_backward_rows(const float *y, const float *t,
                                                         int64_t n, const float *theta,
                                                         const float *Z1, const float *Z2,
                                                         float *dZ1, float *dZ2, float *dy)
{
    __shared__ FwdWeights w;
    load_weights(theta, w);
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float g = 2.0f * (y[s] - t[s]) / (float)n;
    dy[s] = g;
    float dz2[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        dz2[i] = g * w.w3[i] * gelu_grad_f(Z2[s * H + i]);
        dZ2[s * H + i] = dz2[i];
    }
    for (int j = 0; j < H; ++j) {
        float d = 0.0f;
#pragma unroll
        for (int i = 0; i < H; ++i) d += w.w2[i * H + j] * dz2[i];
        dZ1[s * H + j] = d * gelu_grad_f(Z1[s * H + j]);
    }
}

// Block b < 64: gradients owned by hidden unit b (row b of W2, W1, entries b
// of b1, b2, w3).  Block 64: b3 and the loss.  Each thread sums a strided
// subset of the samples; the block then reduces in a fixed tree.
constexpr int WG_THREADS = 128;

__global__ void __launch_bounds__(WG_THREADS) mlp_weight_grads(const float *x, const float *t,
                                                        const float *y, int64_t n,
                                                        const float *H1, const float *H2,
                                                        const float *dZ1, const float *dZ2,
                                                        const float *dy, float *grad,
                                                        float *loss)
{
    constexpr int NV = H + 3 + 3;   // dW2 row (64), dW1 row (3), db1, db2, dw3
    __shared__ float red[NV][WG_THREADS];
    const int b = blockIdx.x;
    float acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0f;
    if (b < H) {
        for (int64_t s = threadIdx.x; s < n; s += blockDim.x) {
            const float d2 = dZ2[s * H + b], d1 = dZ1[s * H + b];
#pragma unroll
            for (int j = 0; j < H; ++j) acc[j] += d2 * H1[s * H + j];
            acc[H + 0] += d1 * x[3 * s];
            acc[H + 1] += d1 * x[3 * s + 1];
            acc[H + 2] += d1 * x[3 * s + 2];
            acc[H + 3] += d1;
            acc[H + 4] += d2;
            acc[H + 5] += dy[s] * H2[s * H + b];
        }
    } else {
        for (int64_t s = threadIdx.x; s < n; s += blockDim.x) {
            acc[0] += dy[s];
            const float r = y[s] - t[s];
            acc[1] += r * r;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int stride = WG_THREADS / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride) {
#pragma unroll
            for (int v = 0; v < NV; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + stride];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (b < H) {
            for (int j = 0; j < H; ++j) grad[MLP_OFF_W2 + b * H + j] = red[j][0];
            grad[MLP_OFF_W1 + 3 * b + 0] = red[H + 0][0];
            grad[MLP_OFF_W1 + 3 * b + 1] = red[H + 1][0];
            grad[MLP_OFF_W1 + 3 * b + 2] = red[H + 2][0];
            grad[MLP_OFF_B1 + b] = red[H + 3][0];
            grad[MLP_OFF_B2 + b] = red[H + 4][0];
            grad[MLP_OFF_W3 + b] = red[H + 5][0];
        } else {
            grad[MLP_OFF_B3] = red[0][0];
            if (loss) *loss = red[1][0] / (float)n;
        }
    }
}

__global__ void mlp_adam(float *theta, const float *grad, float *m, float *v, int64_t step,
                         float lr, float b1, float b2, float eps, float wd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAMS) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)step));
    const float vh = vi / (1.0f - powf(b2, (float)step));
    theta[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * theta[i]);
}

// Pack the FP32 master vector and normalisation into the device layout the
// ascent kernels read.
__global__ void pack_surrogate(const float *theta, const float *norm, SurrogateDev *S)
{
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) S->w2[i] = bf16_bits(theta[MLP_OFF_W2 + i]);
    for (int i = threadIdx.x; i < H * 3; i += blockDim.x) S->w1[i] = bf16_round(theta[MLP_OFF_W1 + i]);
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        S->b1[i] = theta[MLP_OFF_B1 + i];
        S->b2[i] = theta[MLP_OFF_B2 + i];
        S->w3[i] = bf16_round(theta[MLP_OFF_W3 + i]);
    }
    if (threadIdx.x == 0) {
        S->b3 = theta[MLP_OFF_B3];
        for (int k = 0; k < 3; ++k) {
            S->in_mean[k] = norm[k];
            S->in_scale[k] = norm[3 + k];
        }
        S->out_scale = norm[6];
        S->out_offset = norm[7];
    }
}

// Batch evaluation through the in-kernel (MFMA) path: each wave evaluates 64
// samples with surrogate_eval_wave.  Lanes past n evaluate a dummy sample.
__global__ void __launch_bounds__(256) surrogate_eval_kernel(const SurrogateDev *S,
                                                             const float *x, int64_t n,
                                                             float *y)
{
    __shared__ uint32_t lds[4 * MLP_LDS_WORDS_PER_WAVE];
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = s < n;
    const float x0 = ok ? x[3 * s] : 0.0f, x1 = ok ? x[3 * s + 1] : 0.0f,
                x2 = ok ? x[3 * s + 2] : 0.0f;
    const float r = surrogate_eval_wave(S, x0, x1, x2,
                                        lds + (threadIdx.x >> 6) * MLP_LDS_WORDS_PER_WAVE);
    if (ok) y[s] = r;
}

int check_n(int64_t n)
{
    if (n < 1 || n > (int64_t)1 << 31) return fail(XCAPE_EINVAL, "n=%lld out of range", (long long)n);
    return XCAPE_OK;
}

}  // namespace

namespace xcape {

// The loaded network (device memory, one per process); nullptr until
// xcape_surrogate_load.  A load packs the new weights into a buffer of its
// own, waits for the packing to finish and only then publishes the pointer,
// so a kernel launched by another thread sees either the old network or the
// new one, whole.  Replaced buffers are kept (13.6 KB each) until the process
// ends: kernels already queued, and ragged plans (which keep the pointer
// they were created with), may still read them.
std::atomic<SurrogateDev *> g_surrogate{nullptr};
std::mutex g_surrogate_mu;                   // serialises loads
std::vector<SurrogateDev *> g_surrogate_old;  // replaced buffers, kept alive

const void *surrogate_device() { return g_surrogate.load(std::memory_order_acquire); }

}  // namespace xcape

extern "C" {

int xcape_mlp_nparams(void) { return MLP_NPARAMS; }

int xcape_mlp_forward(const float *x, int64_t n, const float *theta, float *y, void *stream)
{
    int rc = check_n(n);
    if (rc || (rc = ensure_device())) return rc;
    if (!x || !theta || !y) return fail(XCAPE_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    mlp_forward_store<<<blocks, 256, 0, s>>>(x, n, theta, y, nullptr, nullptr, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    return e ? hip_fail(e, "mlp_forward") : XCAPE_OK;
}

int xcape_mlp_loss_grad(const float *x, const float *t, int64_t n, const float *theta,
                        float *grad, float *loss, void *stream)
{
    int rc = check_n(n);
    if (rc || (rc = ensure_device())) return rc;
    if (!x || !t || !theta || !grad) return fail(XCAPE_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t nh = (size_t)n * H;
    DevBuf<float> y, Z1, H1, Z2, H2, dZ1, dZ2, dy;
    hipError_t e;
    if ((e = y.alloc(n, s)) || (e = Z1.alloc(nh, s)) || (e = H1.alloc(nh, s)) ||
        (e = Z2.alloc(nh, s)) || (e = H2.alloc(nh, s)) || (e = dZ1.alloc(nh, s)) ||
        (e = dZ2.alloc(nh, s)) || (e = dy.alloc(n, s)))
        return hip_fail(e, "mlp workspace");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    mlp_forward_store<<<blocks, 256, 0, s>>>(x, n, theta, y.ptr, Z1.ptr, H1.ptr, Z2.ptr, H2.ptr);
    mlp_backward_rows<<<blocks, 256, 0, s>>>(y.ptr, t, n, theta, Z1.ptr, Z2.ptr, dZ1.ptr,
                                             dZ2.ptr, dy.ptr);
    mlp_weight_grads<<<H + 1, WG_THREADS, 0, s>>>(x, t, y.ptr, n, H1.ptr, H2.ptr, dZ1.ptr, dZ2.ptr,
                                           dy.ptr, grad, loss);
    e = hipGetLastError();
    return e ? hip_fail(e, "mlp_loss_grad") : XCAPE_OK;
}

int xcape_mlp_adam(float *theta, const float *grad, float *m, float *v, int64_t step, float lr,
                   float beta1, float beta2, float eps, float weight_decay, void *stream)
{
    if (!theta || !grad || !m || !v) return fail(XCAPE_EINVAL, "NULL argument");
    if (step < 1) return fail(XCAPE_EINVAL, "step must be >= 1");
    int rc = ensure_device();
    if (rc) return rc;
    mlp_adam<<<(MLP_NPARAMS + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        theta, grad, m, v, step, lr, beta1, beta2, eps, weight_decay);
    hipError_t e = hipGetLastError();
    return e ? hip_fail(e, "mlp_adam") : XCAPE_OK;
}

int xcape_surrogate_load(const float *theta, const float *norm, void *stream)
{
    if (!theta || !norm) return fail(XCAPE_EINVAL, "NULL argument");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    std::lock_guard<std::mutex> lk(g_surrogate_mu);
    SurrogateDev *fresh = nullptr;
    if ((e = hipMalloc((void **)&fresh, sizeof(SurrogateDev))))
        return hip_fail(e, "surrogate allocation");
    pack_surrogate<<<1, 256, 0, s>>>(theta, norm, fresh);
    if ((e = hipGetLastError()) || (e = hipStreamSynchronize(s))) {
        (void)hipFree(fresh);
        return hip_fail(e, "surrogate load");
    }
    SurrogateDev *old = g_surrogate.exchange(fresh, std::memory_order_acq_rel);
    if (old) g_surrogate_old.push_back(old);
    return XCAPE_OK;
}

int xcape_surrogate_loaded(void) { return surrogate_device() != nullptr; }

int xcape_surrogate_eval(const float *x, int64_t n, float *y, void *stream)
{
    int rc = check_n(n);
    if (rc || (rc = ensure_device())) return rc;
    const SurrogateDev *sur = static_cast<const SurrogateDev *>(surrogate_device());
    if (!sur) return fail(XCAPE_ESTATE, "no surrogate loaded");
    if (!x || !y) return fail(XCAPE_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    surrogate_eval_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(sur, x, n, y);
    hipError_t e = hipGetLastError();
    return e ? hip_fail(e, "surrogate_eval") : XCAPE_OK;
}

}  // extern "C"
