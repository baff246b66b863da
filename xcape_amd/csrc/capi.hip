// C-ABI entry points of libxcape_hip.so (declared in include/xcape_hip.h).
//
// Host-pointer calls stream their columns through pinned staging slots kept
// between calls (level-major host arrays: stream-ordered device buffers on
// the calling thread's stream); device-pointer calls launch directly.  Global
// state: the per-thread streams and error message, and the idle staging sets.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xcape_hip.h"
#include "capi_common.h"
#include "cape_launch.h"

using namespace xcape;

namespace xcape {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory)
        return fail(XCAPE_ENOMEM, "%s: %s", what, hipGetErrorString(e));
    return fail(XCAPE_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int ensure_device()
{
    static thread_local int checked = 0;
    if (checked) return XCAPE_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(XCAPE_ENODEV, "no HIP device visible");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(XCAPE_ENODEV, "hipGetDevice failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(XCAPE_ENODEV, "hipGetDeviceProperties failed");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(XCAPE_ENODEV, "device %d is %s; this library is built for gfx950", dev,
                    prop.gcnArchName);
    checked = 1;
    return XCAPE_OK;
}

hipStream_t thread_stream()
{
    static thread_local hipStream_t s = nullptr;
    if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    return s;
}

// A second stream of the calling thread (default priority), for work that
// runs beside the call's own stream and joins it again through events.
hipStream_t aux_stream()
{
    static thread_local hipStream_t s = nullptr;
    if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    return s;
}

// The thread's events that order its two streams, created once (recording an
// event again only moves it to the new point of its stream; a wait already
// queued keeps the point it was queued for).
constexpr int WARM_PARTS = 4;   // parts of a large warm-started grid's staging
enum { EV_FORK = 0, EV_JOIN, EV_SETUP, EV_PART, EV_COUNT = EV_PART + WARM_PARTS };

hipEvent_t thread_event(int i)
{
    static thread_local hipEvent_t ev[EV_COUNT] = {};
    if (!ev[i] && hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
        ev[i] = nullptr;
    return ev[i];
}

__global__ void zero_u64(unsigned long long *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

hipError_t zero_counters(unsigned long long *p, int n, hipStream_t s)
{
    zero_u64<<<1, 64, 0, s>>>(p, n);
    return hipGetLastError();
}

void copy_stats(const unsigned long long *h, xcape_stats *stats)
{
    stats->columns = (int64_t)h[ST_COLUMNS];
    stats->levels = (int64_t)h[ST_LEVELS];
    stats->substeps = (int64_t)h[ST_SUBSTEPS];
    stats->iterations = (int64_t)h[ST_ITERATIONS];
    stats->nonconverged = (int64_t)h[ST_NONCONVERGED];
    stats->fallbacks = (int64_t)h[ST_FALLBACKS];
    stats->first_substep_iterations = (int64_t)h[ST_FIRST];
    stats->reserved[0] = 0;
    stats->wave_trips = (int64_t)h[ST_WAVE_TRIPS];
    for (int b = 0; b < 6; ++b) {
        stats->block_lanes[b] = (int64_t)h[ST_BLK_LANES + b];
        stats->block_waves[b] = (int64_t)h[ST_BLK_WAVES + b];
    }
}

}  // namespace xcape

// ---------------------------------------------------------------------------
// Scratch of xcape_cape_noalloc_f32: one file-scope device array, of which
// every device holds its own copy (at its own address) from the first use of
// this library on it until the process ends.  The first NOALLOC_HEAD bytes
// hold the warm ascent's column counters (set by a kernel at the start of
// each launch), the rest the staged profiles and warm-start records of one
// call.  Calls on one device share it: they must not overlap.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
constexpr size_t NOALLOC_HEAD = 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
unsigned char *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch` (looked up once per device, with the
// device current; the host-side name is not a device address).
int noalloc_scratch(unsigned char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(XCAPE_ENODEV, "device %d beyond the %d this entry point serves", dev,
                    NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        if ((e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch))) != hipSuccess)
            return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(XCAPE_EHIP, "hipGetSymbolAddress(scratch) gave NULL");
        g_noalloc_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_noalloc_ptr[dev];
    return XCAPE_OK;
}

int check_common(int64_t nk, int64_t ncol)
{
    if (nk < 1 || nk > (1 << 20)) return fail(XCAPE_EINVAL, "nk=%lld out of range", (long long)nk);
    if (ncol < 0) return fail(XCAPE_EINVAL, "ncol=%lld is negative", (long long)ncol);
    return XCAPE_OK;
}

int check_cape_params(double pinc, int source, int adiabat)
{
    if (!(pinc >= 1.0) || !isfinite(pinc))
        return fail(XCAPE_EINVAL, "pinc=%g Pa: must be finite and >= 1 Pa", pinc);
    if (source < 1 || source > 3) return fail(XCAPE_EINVAL, "source=%d not in 1..3", source);
    if (adiabat < 1 || adiabat > 4) return fail(XCAPE_EINVAL, "adiabat=%d not in 1..4", adiabat);
    return XCAPE_OK;
}

// Threads per block of the ascent launches.  XCAPE_CAPE_BLOCK (64, 128 or
// 256) overrides the default, for tuning.
int cape_block()
{
    static const int env = [] {
        const char *v = getenv("XCAPE_CAPE_BLOCK");
        const int b = v ? atoi(v) : 0;
        return (b == 64 || b == 128 || b == 256) ? b : 0;
    }();
    if (env) return env;
    return 256;
}

// Profile copies only the helicity launch reads (the winds of a combined
// call), made on whichever stream runs that launch.
struct HelicityStage {
    const float *src[2];
    float *dst[2];
    int narr;
    int64_t n, nk;
};

// The grid kernels read level-major profiles: with a column per lane, 64
// lanes reading 64 different columns of column-contiguous arrays fetch 64
// cache lines per level each, and on large grids the L2 cannot hold them
// for the ~32 levels each line serves (measured on 2048x1024x137: 142 GB
// fetched per ascent launch for 3.4 GB of profiles).  Level-major profiles go
// straight to the kernels; column-contiguous ones are staged level-major
// first through the LDS-tiled transpose.  run_grid works in chunks of
// columns whose scratch fits in about 16 GiB: per chunk, one transpose launch
// for the profile arrays the CAPE launches read (shared with the helicity
// launch; none for level-major input), the warm-started ascent's scratch
// (level environments and parcels), then the launches.  A combined call
// copies the winds on the helicity's stream, where they run beside the CAPE
// launches (beside the warm-started ascent, after its setup pass) instead of
// ahead of them.
hipError_t launch_grid(const CapeGridArgs &a, bool cape, bool srh, int math, hipStream_t s,
                       const HelicityStage *hs = nullptr)
{
    hipError_t e = hipSuccess;
    if (cape && srh) {
        // the helicity chain (memory- and FP64-bound) beside the ascent
        // (FP32-bound): forked onto the thread's second stream after the
        // work already queued on s, joined back before anything later on s
        hipStream_t s2 = aux_stream();
        const hipEvent_t fork = thread_event(EV_FORK), join = thread_event(EV_JOIN),
                         setup = thread_event(EV_SETUP);
        if (s2 && s2 != s && fork && join && hs && setup && a.warm != 0 && math != 2) {
            // warm start with the winds to copy: the second stream starts after
            // the setup pass, so that the copies run beside the ascent (bound by
            // VALU issue) rather than beside the setup pass (which moves about
            // as many bytes per microsecond as they do)
            CapeGridArgs ac = a;
            ac.setup_done = setup;
            if ((e = launch_grid(ac, true, false, math, s)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(s2, setup, 0)) != hipSuccess) return e;
            e = launch_transpose_levels(hs->src, hs->dst, hs->narr, hs->n, hs->nk, s2);
            if (e == hipSuccess) e = launch_cape_grid_accurate(a, false, GRID_SRH, 256, s2);
            const hipError_t er = hipEventRecord(join, s2);
            if (e == hipSuccess) e = er;
            // join even when a launch failed: s must not run ahead of s2
            const hipError_t ej = hipStreamWaitEvent(s, join, 0);
            return e != hipSuccess ? e : ej;
        }
        if (s2 && s2 != s && fork && join) {
            if ((e = hipEventRecord(fork, s)) != hipSuccess ||
                (e = hipStreamWaitEvent(s2, fork, 0)) != hipSuccess)
                return e;
            if (hs) e = launch_transpose_levels(hs->src, hs->dst, hs->narr, hs->n, hs->nk, s2);
            if (e == hipSuccess) e = launch_cape_grid_accurate(a, false, GRID_SRH, 256, s2);
            const hipError_t er = hipEventRecord(join, s2);
            if (e == hipSuccess) e = er;
            if (e == hipSuccess) e = launch_grid(a, true, false, math, s);
            // join even when a launch failed: s must not run ahead of s2
            const hipError_t ej = hipStreamWaitEvent(s, join, 0);
            return e != hipSuccess ? e : ej;
        }
    }
    if (hs && (e = launch_transpose_levels(hs->src, hs->dst, hs->narr, hs->n, hs->nk, s)) !=
                  hipSuccess)
        return e;
    if (cape) {
        const int block = cape_block();
        if (math == 2)
            e = launch_cape_grid_accurate(a, true, GRID_CAPE, block, s);
        else if (math == 1)
            e = launch_cape_grid_fast(a, GRID_CAPE, block, s);
        else
            e = launch_cape_grid_accurate(a, false, GRID_CAPE, block, s);
        if (e != hipSuccess) return e;
    }
    // the helicity chain keeps the reference's precisions in every math mode
    // but the FP64 one (storm motion in double)
    if (srh) e = launch_cape_grid_accurate(a, math == 2, GRID_SRH, 256, s);
    return e;
}

// Scratch floats per column of one run_grid chunk (staged profiles plus the
// warm-started ascent's records); 0 when the call needs none.
int64_t grid_scratch_floats(const CapeGridArgs &a0, bool cape, bool srh, int math)
{
    const bool stage = a0.ls == 1 && a0.nk_in > 1 && !a0.in_f64;
    const bool warm = cape && a0.warm != 0 && math != 2;
    const int narr = stage ? 2 + (a0.p != nullptr) + (srh ? 2 : 0) : 0;
    return (int64_t)narr * a0.nk_in +
           (warm ? warm_scratch_floats(a0.nk_in, a0.p_lfc != nullptr) : 0);
}

// A helicity-only call on at least this many column-contiguous columns
// (XCAPE_SRH_PIPE_COLUMNS overrides it, read on every call; for tests) is
// staged and computed in SRH_PARTS parts: each part's helicity launch runs on
// the thread's second stream while the next part's profiles are copied.
constexpr int SRH_PARTS = 4;

int64_t srh_pipe_min_columns()
{
    if (const char *env = getenv("XCAPE_SRH_PIPE_COLUMNS")) {
        const long long v = atoll(env);
        if (v > 0) return v;
    }
    return (int64_t)1 << 18;
}

// Warm-started grids of at least this many column-contiguous columns
// (XCAPE_WARM_PIPE_COLUMNS overrides it, read on every call; for tests) are
// staged in WARM_PARTS parts.
int64_t warm_pipe_min_columns()
{
    if (const char *env = getenv("XCAPE_WARM_PIPE_COLUMNS")) {
        const long long v = atoll(env);
        if (v > 0) return v;
    }
    return (int64_t)1 << 20;
}

hipError_t launch_cape_only(const CapeGridArgs &a, int math, hipStream_t s)
{
    const int block = cape_block();
    if (math == 1) return launch_cape_grid_fast(a, GRID_CAPE, block, s);
    return launch_cape_grid_accurate(a, math == 2, GRID_CAPE, block, s);
}

// A large warm-started grid of column-contiguous profiles, staged in parts:
// part i's setup pass runs on s while part i + 1's profiles are copied on the
// thread's second stream s2, then one counter reset and ascent cover every
// column.  The copies fill the whole grid's level-major arrays dst (rows of
// a.ncol; a's profile pointers are those arrays), and each setup pass its
// part's columns of the records.  In a combined call the winds (hs) are then
// copied and the helicity launched on s2, after the last setup pass, beside
// the ascent.  Columns are independent, so the outputs are those of the
// one-launch call.
hipError_t launch_warm_parts(const CapeGridArgs &a, bool srh, int math, hipStream_t s,
                             hipStream_t s2, const float *const *src, float *const *dst,
                             int narr, const HelicityStage *hs)
{
    const int64_t n = a.ncol, nk = a.nk_in, esz = 4;
    int64_t c = (n + WARM_PARTS - 1) / WARM_PARTS;
    c = (c + 63) & ~(int64_t)63;   // parts of whole waves
    const int parts = (int)((n + c - 1) / c);
    // once anything is queued on s2, every return joins s2 back into s
    bool forked = false;
    const auto joined = [&](hipError_t err) {
        if (forked) {
            const hipEvent_t join = thread_event(EV_JOIN);
            const hipError_t er = hipEventRecord(join, s2);
            const hipError_t ew = hipStreamWaitEvent(s, join, 0);
            if (err == hipSuccess) err = er != hipSuccess ? er : ew;
        }
        return err;
    };
    const auto copies = [&](int64_t c0, hipStream_t st) {
        const int64_t m = c < n - c0 ? c : n - c0;
        const float *ps[5];
        float *pd[5];
        for (int i = 0; i < narr; ++i) {
            ps[i] = src[i] + c0 * nk;
            pd[i] = dst[i] + c0;
        }
        return launch_transpose_levels(ps, pd, narr, m, nk, st, n);
    };
    hipError_t e = copies(0, s);
    if (e != hipSuccess) return e;
    const hipEvent_t fork = thread_event(EV_FORK);
    if ((e = hipEventRecord(fork, s)) != hipSuccess ||
        (e = hipStreamWaitEvent(s2, fork, 0)) != hipSuccess)
        return e;
    forked = true;
    for (int i = 1; i < parts; ++i) {
        if ((e = copies(i * c, s2)) != hipSuccess ||
            (e = hipEventRecord(thread_event(EV_PART + i), s2)) != hipSuccess)
            return joined(e);
    }
    for (int i = 0; i < parts; ++i) {
        const int64_t c0 = i * c, m = c < n - c0 ? c : n - c0;
        if (i > 0 && (e = hipStreamWaitEvent(s, thread_event(EV_PART + i), 0)) != hipSuccess)
            return joined(e);
        CapeGridArgs b = a;
        b.ncol = m;
        // level k of column c at [k * ls + c]: the part's columns start c0 into each row
        b.t += c0;
        b.td += c0;
        if (b.p) b.p += c0;
        b.ps += c0; b.ts += c0; b.tds += c0;
        if (b.start) b.start += c0;
        b.cape = static_cast<char *>(b.cape) + c0 * esz;
        b.cin = static_cast<char *>(b.cin) + c0 * esz;
        b.zout = static_cast<char *>(b.zout) + c0 * esz;
        b.mulvl += c0;
        if (b.p_lfc) {
            b.p_lfc = static_cast<char *>(b.p_lfc) + c0 * esz;
            b.p_el = static_cast<char *>(b.p_el) + c0 * esz;
        }
        b.rec_ls = n;
        b.rec_col0 = c0;
        b.warm_phase = 1;
        b.setup_done = nullptr;
        if ((e = launch_cape_only(b, math, s)) != hipSuccess) return joined(e);
    }
    if (srh) {
        const hipEvent_t setup = thread_event(EV_SETUP);
        if ((e = hipEventRecord(setup, s)) != hipSuccess ||
            (e = hipStreamWaitEvent(s2, setup, 0)) != hipSuccess)
            return joined(e);
        if (hs && (e = launch_transpose_levels(hs->src, hs->dst, hs->narr, hs->n, hs->nk,
                                               s2)) != hipSuccess)
            return joined(e);
        if ((e = launch_cape_grid_accurate(a, false, GRID_SRH, 256, s2)) != hipSuccess)
            return joined(e);
    }
    CapeGridArgs z = a;
    z.warm_phase = 2;
    return joined(launch_cape_only(z, math, s));
}

// ext/ext_work: caller-owned scratch (grid_scratch_floats per column of the
// whole call, which holds any chunk's, and 8 counters) used instead of
// allocating; NULL to allocate.
hipError_t run_grid(const CapeGridArgs &a0, bool cape, bool srh, int math, hipStream_t s,
                    float *ext = nullptr, unsigned int *ext_work = nullptr)
{
    // float64 inputs (the helicity chain only) are read where they lie
    const bool stage = a0.ls == 1 && a0.nk_in > 1 && !a0.in_f64;
    const bool warm = cape && a0.warm != 0 && math != 2;
    if (!stage && !warm) return launch_grid(a0, cape, srh, math, s);
    const bool sigma = a0.p != nullptr;
    const float *src0[5];
    int narr = 0;
    int it = -1, itd = -1, iu = -1, iv = -1, ip = -1;
    if (stage) {
        it = narr; src0[narr++] = a0.t;
        itd = narr; src0[narr++] = a0.td;
        if (sigma) { ip = narr; src0[narr++] = a0.p; }
        if (srh) { iu = narr; src0[narr++] = a0.u; iv = narr; src0[narr++] = a0.v; }
    }
    const int64_t nk = a0.nk_in;
    const int64_t wfl = warm ? warm_scratch_floats(nk, a0.p_lfc != nullptr) : 0;
    const int64_t per_col = ((int64_t)narr * nk + wfl) * (int64_t)sizeof(float);
    int64_t chunk = ((int64_t)16 << 30) / per_col;
    if (chunk < ((int64_t)1 << 20)) chunk = (int64_t)1 << 20;
    // XCAPE_STAGE_COLUMNS: columns per chunk, overriding the 16 GiB sizing
    // (read on every call; for tests of the chunked path)
    if (const char *env = getenv("XCAPE_STAGE_COLUMNS")) {
        const long long v = atoll(env);
        if (v > 0) chunk = v;
    }
    // the warm ascent indexes its level environments with 32-bit offsets
    if (warm) {
        const int64_t cap = (((int64_t)1 << 32) - 1) / (nk + 1);
        if (chunk > cap) chunk = cap;
    }
    if (chunk > a0.ncol) chunk = a0.ncol;
    // the helicity-only call in parts (each part's copies in a region of its
    // own, so that one part's helicity launch can still read them while the
    // next part is copied), when the whole call fits one chunk's scratch
    hipStream_t s2 = nullptr;
    int64_t parts = 1;
    if (!cape && srh && stage && !ext && chunk >= a0.ncol &&
        a0.ncol >= srh_pipe_min_columns() && a0.ncol >= 2 * 64 &&
        (s2 = aux_stream()) != nullptr && s2 != s && thread_event(EV_FORK) &&
        thread_event(EV_JOIN)) {
        int64_t c = (a0.ncol + SRH_PARTS - 1) / SRH_PARTS;
        c = (c + 63) & ~(int64_t)63;   // parts of whole waves
        chunk = c;
        parts = (a0.ncol + c - 1) / c;
    }
    // a large warm-started grid: staged in parts, each beside a setup pass
    bool wparts = false;
    if (warm && stage && !ext && chunk >= a0.ncol && a0.ncol >= warm_pipe_min_columns() &&
        a0.ncol >= 2 * 64 && !a0.nk_col && !a0.src_col && !a0.out_index) {
        s2 = aux_stream();
        wparts = s2 && s2 != s && thread_event(EV_FORK) && thread_event(EV_JOIN) &&
                 thread_event(EV_SETUP) && thread_event(EV_PART + WARM_PARTS - 1);
        for (int i = 1; wparts && i < WARM_PARTS; ++i) wparts = thread_event(EV_PART + i);
    }
    DevBuf<float> scratch_;
    DevBuf<unsigned int> work_;
    float *scratch = ext;
    unsigned int *work = ext_work;
    hipError_t e = hipSuccess;
    if (!ext) {
        if ((e = scratch_.alloc((size_t)(narr * nk + wfl) * (size_t)chunk * (size_t)parts,
                                s)) != hipSuccess)
            return e;
        if (warm && (e = work_.alloc(8, s)) != hipSuccess) return e;   // WORK_PARTS counters
        scratch = scratch_.ptr;
        work = work_.ptr;
    }
    const int64_t esz = math == 2 ? 8 : 4;
    // once a part's helicity launch is on s2, every return joins s2 back into
    // s first: the scratch is freed on s, and s2 may still read it
    bool forked = false;
    const auto joined = [&](hipError_t err) {
        if (forked) {
            const hipEvent_t join = thread_event(EV_JOIN);
            const hipError_t er = hipEventRecord(join, s2);
            const hipError_t ew = hipStreamWaitEvent(s, join, 0);
            forked = false;
            if (err == hipSuccess) err = er != hipSuccess ? er : ew;
        }
        return err;
    };
    for (int64_t c0 = 0; c0 < a0.ncol; c0 += chunk) {
        const int64_t n = chunk < a0.ncol - c0 ? chunk : a0.ncol - c0;
        CapeGridArgs a = a0;
        a.ncol = n;
        HelicityStage hs_, *hs = nullptr;
        const float *src[5];
        float *dst[5];
        int nfirst = 0;
        if (stage) {
            const int64_t o = c0 * nk;
            // (parts > 1: the part's own region)
            float *cs = scratch + (parts > 1 ? (size_t)narr * (size_t)nk * (size_t)c0 : 0);
            for (int i = 0; i < narr; ++i) {
                src[i] = src0[i] + o;
                dst[i] = cs + (size_t)i * (size_t)nk * (size_t)n;
            }
            // in a combined call the winds are copied on the helicity's stream,
            // beside the CAPE launches; a grid in parts is copied part by part
            nfirst = cape && srh ? narr - 2 : narr;
            if (!wparts && (e = launch_transpose_levels(src, dst, nfirst, n, nk, s)) !=
                               hipSuccess)
                return joined(e);
            if (nfirst < narr) {
                hs_ = HelicityStage{{src[iu], src[iv]}, {dst[iu], dst[iv]}, 2, n, nk};
                hs = &hs_;
            }
            a.ls = n;
            a.t = dst[it];
            a.td = dst[itd];
            if (sigma) a.p = dst[ip];
            if (srh) { a.u = dst[iu]; a.v = dst[iv]; }
        } else {
            // level-major input (or a single level): the chunk's columns
            // start c0 into each level's row
            a.t += c0; a.td += c0;
            if (sigma) a.p += c0;
            if (srh) { a.u += c0; a.v += c0; }
        }
        a.ps += c0; a.ts += c0; a.tds += c0;
        if (a.start) a.start += c0;
        if (warm) {
            float *w = scratch + (size_t)narr * (size_t)nk * (size_t)n;
            const size_t lv = (size_t)(nk + 1) * (size_t)n;
            a.env = w;
            a.par = w + 4 * lv;
            a.env_p = a0.p_lfc ? a.par + 16 * (size_t)n : nullptr;
            a.work = work;
        }
        if (cape) {
            a.cape = static_cast<char *>(a.cape) + c0 * esz;
            a.cin = static_cast<char *>(a.cin) + c0 * esz;
            a.zout = static_cast<char *>(a.zout) + c0 * esz;
            a.mulvl += c0;
            if (a.p_lfc) {
                a.p_lfc = static_cast<char *>(a.p_lfc) + c0 * esz;
                a.p_el = static_cast<char *>(a.p_el) + c0 * esz;
            }
        }
        if (srh) {
            a.us += c0; a.vs += c0;
            a.srh_rm += c0; a.srh_lm += c0;
            if (a.rm) {
                const int64_t b = 2 * c0 * esz;
                a.rm = static_cast<char *>(a.rm) + b;
                a.lm = static_cast<char *>(a.lm) + b;
                a.mean6 = static_cast<char *>(a.mean6) + b;
            }
        }
        if (parts > 1) {
            // this part's helicity on s2, once its copies are done; the next
            // part's copies go on s meanwhile
            const hipEvent_t fork = thread_event(EV_FORK);
            if ((e = hipEventRecord(fork, s)) == hipSuccess &&
                (e = hipStreamWaitEvent(s2, fork, 0)) == hipSuccess) {
                forked = true;
                e = launch_cape_grid_accurate(a, math == 2, GRID_SRH, 256, s2);
            }
            // after the last part, or a failed one: s waits for s2
            if (e != hipSuccess || c0 + n >= a0.ncol) e = joined(e);
            if (e != hipSuccess) return e;
            continue;
        }
        if (wparts) e = launch_warm_parts(a, srh, math, s, s2, src, dst, nfirst, hs);
        else e = launch_grid(a, cape, srh, math, s, hs);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Everything a grid call needs; pointers as the caller passed them.
struct GridCall {
    int kind = GRID_CAPE;
    bool fp64 = false;
    const float *p = nullptr;
    int p_shared = 0;
    const int32_t *start = nullptr;
    const float *t = nullptr, *td = nullptr, *ps = nullptr, *ts = nullptr, *tds = nullptr;
    const float *u = nullptr, *v = nullptr, *us = nullptr, *vs = nullptr;
    int64_t nk = 0, ncol = 0;
    double pinc = 500.0, ml_depth = 500.0, depth = 3000.0;
    int source = 1, adiabat = 1;
    void *cape = nullptr, *cin = nullptr, *zout = nullptr;
    int32_t *mulev = nullptr;
    void *p_lfc = nullptr, *p_el = nullptr;    // LFC/EL outputs (float, or double when fp64)
    double *srh_rm = nullptr, *srh_lm = nullptr;
    void *rm = nullptr, *lm = nullptr, *mean6 = nullptr;  // float, or double when fp64
    uint32_t flags = 0;
    xcape_stats *stats = nullptr;
    void *stream = nullptr;
    // the no-allocation entry point's scratch (device address), or nullptr
    unsigned char *noalloc = nullptr;
};

// ---------------------------------------------------------------------------
// Host arrays in the reference's layout (each column contiguous, as Fortran's
// p3d(nk, n2)) stream through pinned staging.  Pageable host memory reaches
// the device at about 25 GB/s, pinned memory at about 57 GB/s (PCIe 5 x16 on
// MI355X), and the uploads, not the kernels, set the pace of a host-array
// call.  The call's columns go in chunks:
//   * host threads copy a chunk's inputs into one of STAGE_IN pinned input
//     buffers, free again as soon as its last upload has ended;
//   * one upload stream carries every chunk's inputs, one after the other, so
//     the copy engine runs at its single-stream rate without gaps;
//   * STAGE_SLOTS compute slots (device buffers, a stream and a pinned output
//     area each) run the chunk's kernels once its upload has ended, download
//     its outputs, and are drained into the caller's arrays when the calling
//     thread comes back to the slot.
// The first chunks are a quarter and a half of the others, so the first
// upload starts early.  A chunk of an array is one contiguous range of the
// caller's array and one contiguous run of the staging areas: a chunk moves
// in one upload and one download.

constexpr int STAGE_IN = 3;
constexpr int STAGE_SLOTS = 4;
constexpr int STAGE_SETS = 2;   // at most this many calls stream at once; others wait
constexpr size_t STAGE_ALIGN = 256;

size_t align_up(size_t x) { return (x + STAGE_ALIGN - 1) / STAGE_ALIGN * STAGE_ALIGN; }

// The per-column arrays of one direction (inputs or outputs) of a call, and
// their places in a slot's area for a chunk of n columns: array after array,
// each STAGE_ALIGN-aligned.
struct Slab {
    int count = 0;
    const char *src[12];   // inputs: the caller's arrays
    char *dst[12];         // outputs: the caller's arrays
    size_t cb[12];         // bytes per column
    int add(const void *in, void *out, size_t col_bytes)
    {
        src[count] = static_cast<const char *>(in);
        dst[count] = static_cast<char *>(out);
        cb[count] = col_bytes;
        return count++;
    }
    size_t place(int64_t n, size_t *off) const
    {
        size_t at = 0;
        for (int i = 0; i < count; ++i) {
            off[i] = at;
            at = align_up(at + (size_t)n * cb[i]);
        }
        return at;
    }
};

struct CopyJob {
    char *dst;
    const char *src;
    size_t n;
};

// Copies the jobs' bytes with up to `threads` threads, each taking an equal
// share of the jobs' concatenation.
void parallel_copy(const CopyJob *jobs, int njobs, int threads)
{
    size_t total = 0;
    for (int j = 0; j < njobs; ++j) total += jobs[j].n;
    if (total < ((size_t)4 << 20)) threads = 1;
    const size_t per = (total + threads - 1) / (size_t)threads;
    auto share = [jobs, njobs, total, per](int t) {
        const size_t lo = (size_t)t * per, hi = lo + per < total ? lo + per : total;
        size_t base = 0;
        for (int j = 0; j < njobs && base < hi; base += jobs[j].n, ++j) {
            const size_t b0 = lo > base ? lo : base;
            const size_t b1 = hi < base + jobs[j].n ? hi : base + jobs[j].n;
            if (b0 < b1) memcpy(jobs[j].dst + (b0 - base), jobs[j].src + (b0 - base), b1 - b0);
        }
    };
    std::vector<std::thread> pool;
    try {
        for (int t = 1; t < threads; ++t) pool.emplace_back(share, t);
    } catch (...) {
        // fewer threads than asked: the calling thread takes the other shares
    }
    for (int t = 1 + (int)pool.size(); t < threads; ++t) share(t);
    share(0);
    for (auto &th : pool) th.join();
}

int host_threads()
{
    static const int n = [] {
        const char *v = getenv("XCAPE_HOST_THREADS");
        const int k = v ? atoi(v) : 8;
        return k < 1 ? 1 : (k > 64 ? 64 : k);
    }();
    return n;
}

// Pinned input buffers, compute slots and streams of one streaming call; kept
// for the next call (xcape_release_host_staging frees the idle sets).
struct StagingSet {
    int device = -1;
    size_t in_cap = 0, out_cap = 0, dev_cap = 0;
    char *pin_in[STAGE_IN] = {}, *pin_out[STAGE_SLOTS] = {};
    char *dev = nullptr;
    hipStream_t up = nullptr;                 // every upload
    hipEvent_t uploaded[STAGE_IN] = {};       // a pinned input buffer's last upload
    hipStream_t stream[STAGE_SLOTS] = {};     // a slot's kernels and download
    hipEvent_t done[STAGE_SLOTS] = {};        // a slot's last download

    void free_pinned()
    {
        for (auto &x : pin_in) {
            if (x) (void)hipHostFree(x);
            x = nullptr;
        }
        for (auto &x : pin_out) {
            if (x) (void)hipHostFree(x);
            x = nullptr;
        }
        in_cap = out_cap = 0;
    }
    void destroy()
    {
        free_pinned();
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        dev_cap = 0;
        if (up) (void)hipStreamDestroy(up);
        up = nullptr;
        for (auto &x : uploaded) {
            if (x) (void)hipEventDestroy(x);
            x = nullptr;
        }
        for (int i = 0; i < STAGE_SLOTS; ++i) {
            if (stream[i]) (void)hipStreamDestroy(stream[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
            stream[i] = nullptr;
            done[i] = nullptr;
        }
        device = -1;
    }
    // sizes the set for the current device (the caller made it current)
    hipError_t reserve(int dv, size_t in, size_t out, size_t devb)
    {
        hipError_t e;
        if (device != dv) destroy();
        device = dv;
        const unsigned ev = hipEventDisableTiming;
        if (!up && (e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking))) return e;
        for (auto &x : uploaded)
            if (!x && (e = hipEventCreateWithFlags(&x, ev))) return e;
        for (int i = 0; i < STAGE_SLOTS; ++i) {
            if (!stream[i] && (e = hipStreamCreateWithFlags(&stream[i], hipStreamNonBlocking)))
                return e;
            if (!done[i] && (e = hipEventCreateWithFlags(&done[i], ev))) return e;
        }
        if (in > in_cap || out > out_cap) {
            free_pinned();
            for (auto &x : pin_in)
                if ((e = hipHostMalloc((void **)&x, in, hipHostMallocDefault))) return e;
            for (auto &x : pin_out)
                if ((e = hipHostMalloc((void **)&x, out, hipHostMallocDefault))) return e;
            in_cap = in;
            out_cap = out;
        }
        if (devb > dev_cap) {
            if (dev) (void)hipFree(dev);
            dev = nullptr;
            dev_cap = 0;
            if ((e = hipMalloc((void **)&dev, devb))) return e;
            dev_cap = devb;
        }
        return hipSuccess;
    }
};

std::mutex g_stage_mu;
std::condition_variable g_stage_cv;
std::vector<StagingSet *> g_stage_idle;
int g_stage_sets = 0;

StagingSet *acquire_staging(int dv)
{
    std::unique_lock<std::mutex> lk(g_stage_mu);
    for (;;) {
        for (size_t i = 0; i < g_stage_idle.size(); ++i)
            if (g_stage_idle[i]->device == dv) {
                StagingSet *st = g_stage_idle[i];
                g_stage_idle.erase(g_stage_idle.begin() + (long)i);
                return st;
            }
        if (g_stage_sets < STAGE_SETS) {
            ++g_stage_sets;
            return new StagingSet;
        }
        if (!g_stage_idle.empty()) {   // another device's set: reserve() rebuilds it
            StagingSet *st = g_stage_idle.back();
            g_stage_idle.pop_back();
            return st;
        }
        g_stage_cv.wait(lk);
    }
}

void release_staging(StagingSet *st)
{
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        g_stage_idle.push_back(st);
    }
    g_stage_cv.notify_one();
}

// Input megabytes per chunk; XCAPE_HOST_CHUNK_MB overrides the 128 MB default
// (read on every call).
int host_chunk_mb()
{
    const char *v = getenv("XCAPE_HOST_CHUNK_MB");
    const int k = v ? atoi(v) : 128;
    return k < 1 ? 1 : (k > 4096 ? 4096 : k);
}

// The streamed host-array call: tmpl carries the call's parameters, c its
// host arrays (column-contiguous).
int stream_host_grid(const GridCall &c, const CapeGridArgs &tmpl, bool want_cape, bool want_srh,
                     int math)
{
    const int64_t nk = c.nk, ncol = c.ncol;
    const size_t osz = c.fp64 ? sizeof(double) : sizeof(float);
    const size_t prof = (size_t)nk * sizeof(float);
    Slab in, out;
    const int i_t = in.add(c.t, nullptr, prof), i_td = in.add(c.td, nullptr, prof);
    const int i_p = c.p_shared ? -1 : in.add(c.p, nullptr, prof);
    const int i_ps = in.add(c.ps, nullptr, 4), i_ts = in.add(c.ts, nullptr, 4);
    const int i_tds = in.add(c.tds, nullptr, 4);
    const int i_st = c.start ? in.add(c.start, nullptr, 4) : -1;
    int i_u = -1, i_v = -1, i_us = -1, i_vs = -1;
    if (want_srh) {
        i_u = in.add(c.u, nullptr, prof);
        i_v = in.add(c.v, nullptr, prof);
        i_us = in.add(c.us, nullptr, 4);
        i_vs = in.add(c.vs, nullptr, 4);
    }
    int o_cape = -1, o_cin = -1, o_z = -1, o_mul = -1, o_lfc = -1, o_el = -1;
    int o_srm = -1, o_slm = -1, o_rm = -1, o_lm = -1, o_mw = -1;
    if (want_cape) {
        o_cape = out.add(nullptr, c.cape, osz);
        o_cin = out.add(nullptr, c.cin, osz);
        o_z = out.add(nullptr, c.zout, osz);
        o_mul = out.add(nullptr, c.mulev, 4);
        if (c.p_lfc) {
            o_lfc = out.add(nullptr, c.p_lfc, osz);
            o_el = out.add(nullptr, c.p_el, osz);
        }
    }
    if (want_srh) {
        o_srm = out.add(nullptr, c.srh_rm, 8);
        o_slm = out.add(nullptr, c.srh_lm, 8);
        if (c.rm) {
            o_rm = out.add(nullptr, c.rm, 2 * osz);
            o_lm = out.add(nullptr, c.lm, 2 * osz);
            o_mw = out.add(nullptr, c.mean6, 2 * osz);
        }
    }
    size_t in_col = 0;
    for (int i = 0; i < in.count; ++i) in_col += in.cb[i];
    // columns per chunk: about XCAPE_HOST_CHUNK_MB of input, a multiple of 64
    int64_t m = ((int64_t)host_chunk_mb() << 20) / (int64_t)in_col / 64 * 64;
    if (m < 64) m = 64;
    if (m > ncol) m = ncol;
    size_t off_in[12], off_out[12];
    const size_t in_bytes = in.place(m, off_in), out_bytes = out.place(m, off_out);
    CapeGridArgs shape = tmpl;   // the chunks' launches read column-contiguous profiles
    shape.ls = 1;
    shape.p = c.p_shared ? nullptr : c.p;
    const size_t scr_bytes = align_up((size_t)grid_scratch_floats(shape, want_cape, want_srh, math) *
                                      (size_t)m * sizeof(float));
    const size_t slot_bytes = in_bytes + out_bytes + scr_bytes + STAGE_ALIGN;   // + work counters
    const size_t shared = align_up((size_t)nk * sizeof(float)) +
                          align_up(ST_COUNT * sizeof(unsigned long long));
    int dv = 0;
    hipError_t e = hipGetDevice(&dv);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (c.stream && (e = hipStreamSynchronize((hipStream_t)c.stream)) != hipSuccess)
        return hip_fail(e, "hipStreamSynchronize(stream)");
    StagingSet *st = acquire_staging(dv);
    struct Finish {   // on every return: the set's streams drained, the set back in the pool
        StagingSet *st;
        ~Finish()
        {
            if (st->up) (void)hipStreamSynchronize(st->up);
            for (int i = 0; i < STAGE_SLOTS; ++i)
                if (st->stream[i]) (void)hipStreamSynchronize(st->stream[i]);
            release_staging(st);
        }
    } finish{st};
    if ((e = st->reserve(dv, in_bytes, out_bytes, shared + STAGE_SLOTS * slot_bytes)))
        return hip_fail(e, "pinned staging allocation");
    char *const p1d = st->dev;
    unsigned long long *const dstats =
        reinterpret_cast<unsigned long long *>(st->dev + align_up((size_t)nk * sizeof(float)));
#define XC_TRY(x) if ((e = (x)) != hipSuccess) return hip_fail(e, #x)
    if (c.p_shared)
        XC_TRY(hipMemcpyAsync(p1d, c.p, (size_t)nk * sizeof(float), hipMemcpyHostToDevice, st->up));
    if (c.stats && want_cape) XC_TRY(zero_counters(dstats, ST_COUNT, st->up));
    XC_TRY(hipStreamSynchronize(st->up));
    struct Pending {
        int64_t c0 = 0, n = 0;
        bool busy = false;
    } pend[STAGE_SLOTS];
    bool in_used[STAGE_IN] = {};
    CopyJob jobs[12];
    const int threads = host_threads();
    int64_t chunk_no = 0;
    // the outputs of the chunk that slot sl last ran, into the caller's arrays
    auto drain = [&](int sl) -> hipError_t {
        const hipError_t r = hipEventSynchronize(st->done[sl]);
        if (r != hipSuccess) return r;
        size_t off[12];
        out.place(pend[sl].n, off);
        for (int i = 0; i < out.count; ++i)
            jobs[i] = {out.dst[i] + (size_t)pend[sl].c0 * out.cb[i], st->pin_out[sl] + off[i],
                       (size_t)pend[sl].n * out.cb[i]};
        parallel_copy(jobs, out.count, threads);
        pend[sl].busy = false;
        return hipSuccess;
    };
    for (int64_t c0 = 0, n = 0; c0 < ncol; c0 += n, ++chunk_no) {
        const int r = (int)(chunk_no % STAGE_IN), sl = (int)(chunk_no % STAGE_SLOTS);
        // a quarter, a half, then whole chunks
        n = chunk_no >= 2 ? m : ((m >> (2 - chunk_no)) + 63) / 64 * 64;
        if (n < 64) n = 64;
        if (n > m) n = m;
        if (n > ncol - c0) n = ncol - c0;
        if (in_used[r]) XC_TRY(hipEventSynchronize(st->uploaded[r]));
        size_t oi[12], oo[12];
        const size_t ib = in.place(n, oi), ob = out.place(n, oo);
        for (int i = 0; i < in.count; ++i)
            jobs[i] = {st->pin_in[r] + oi[i], in.src[i] + (size_t)c0 * in.cb[i],
                       (size_t)n * in.cb[i]};
        parallel_copy(jobs, in.count, threads);
        // the slot's device buffers are free once its previous chunk is done
        if (pend[sl].busy) XC_TRY(drain(sl));
        char *din = st->dev + shared + (size_t)sl * slot_bytes;
        char *dout = din + in_bytes;
        float *dscr = reinterpret_cast<float *>(dout + out_bytes);
        unsigned int *dwork = reinterpret_cast<unsigned int *>(dout + out_bytes + scr_bytes);
        hipStream_t ss = st->stream[sl];
        XC_TRY(hipMemcpyAsync(din, st->pin_in[r], ib, hipMemcpyHostToDevice, st->up));
        XC_TRY(hipEventRecord(st->uploaded[r], st->up));
        in_used[r] = true;
        XC_TRY(hipStreamWaitEvent(ss, st->uploaded[r], 0));
        auto fin = [&](int i) { return i < 0 ? nullptr : reinterpret_cast<const float *>(din + oi[i]); };
        auto fout = [&](int i) { return i < 0 ? nullptr : static_cast<void *>(dout + oo[i]); };
        CapeGridArgs a = tmpl;
        a.ncol = n;
        a.ls = 1;
        a.t = fin(i_t); a.td = fin(i_td);
        a.p = c.p_shared ? nullptr : fin(i_p);
        a.p1d = c.p_shared ? reinterpret_cast<const float *>(p1d) : nullptr;
        a.start = reinterpret_cast<const int32_t *>(fin(i_st));
        a.ps = fin(i_ps); a.ts = fin(i_ts); a.tds = fin(i_tds);
        a.u = fin(i_u); a.v = fin(i_v); a.us = fin(i_us); a.vs = fin(i_vs);
        a.cape = fout(o_cape); a.cin = fout(o_cin); a.zout = fout(o_z);
        a.mulvl = static_cast<int32_t *>(fout(o_mul));
        a.p_lfc = fout(o_lfc); a.p_el = fout(o_el);
        a.srh_rm = static_cast<double *>(fout(o_srm)); a.srh_lm = static_cast<double *>(fout(o_slm));
        a.rm = fout(o_rm); a.lm = fout(o_lm); a.mean6 = fout(o_mw);
        a.stats = (c.stats && want_cape) ? dstats : nullptr;
        if ((e = run_grid(a, want_cape, want_srh, math, ss, dscr, dwork)) != hipSuccess)
            return hip_fail(e, "kernel launch");
        XC_TRY(hipMemcpyAsync(st->pin_out[sl], dout, ob, hipMemcpyDeviceToHost, ss));
        XC_TRY(hipEventRecord(st->done[sl], ss));
        pend[sl] = {c0, n, true};
    }
    // the last chunks' outputs, oldest first
    for (int64_t k = 0; k < STAGE_SLOTS; ++k) {
        const int sl = (int)((chunk_no + k) % STAGE_SLOTS);
        if (pend[sl].busy) XC_TRY(drain(sl));
    }
    if (c.stats && want_cape) {
        unsigned long long h[ST_COUNT];
        XC_TRY(hipMemcpy(h, dstats, sizeof h, hipMemcpyDeviceToHost));
        copy_stats(h, c.stats);
    }
#undef XC_TRY
    return XCAPE_OK;
}

int grid_entry(const GridCall &c)
{
    int rc = check_common(c.nk, c.ncol);
    if (rc) return rc;
    const bool want_cape = c.kind & GRID_CAPE, want_srh = c.kind & GRID_SRH;
    if (want_cape && (rc = check_cape_params(c.pinc, c.source, c.adiabat))) return rc;
    if (want_srh && !isfinite(c.depth)) return fail(XCAPE_EINVAL, "depth=%g not finite", c.depth);
    if (c.ncol > 0) {
        if (!c.p || !c.t || !c.td || !c.ps || !c.ts || !c.tds)
            return fail(XCAPE_EINVAL, "NULL input array");
        if (want_cape && (!c.cape || !c.cin || !c.mulev || !c.zout))
            return fail(XCAPE_EINVAL, "NULL CAPE output array");
        if ((c.p_lfc != nullptr) != (c.p_el != nullptr))
            return fail(XCAPE_EINVAL, "p_lfc and p_el must be both NULL or both set");
        if (want_srh && (!c.u || !c.v || !c.us || !c.vs || !c.srh_rm || !c.srh_lm))
            return fail(XCAPE_EINVAL, "NULL helicity array");
        if (want_srh && ((c.rm != nullptr) != (c.lm != nullptr) ||
                         (c.rm != nullptr) != (c.mean6 != nullptr)))
            return fail(XCAPE_EINVAL, "rm, lm and mean6 must be all NULL or all set");
    }
    const uint32_t known = XCAPE_DEVICE_PTRS | XCAPE_FAST_MATH | XCAPE_WARM_START |
                           XCAPE_NO_SURROGATE | XCAPE_LEVEL_MAJOR | XCAPE_NO_PREDICTION |
                           XCAPE_INPUT_F64;
    if (c.flags & ~known) return fail(XCAPE_EINVAL, "unknown flags 0x%x", c.flags & ~known);
    const bool f64in = c.flags & XCAPE_INPUT_F64;
    if (f64in && want_cape)
        return fail(XCAPE_EINVAL, "XCAPE_INPUT_F64 is for the helicity entry points only");
    if (c.fp64 && (c.flags & (XCAPE_FAST_MATH | XCAPE_WARM_START)))
        return fail(XCAPE_EINVAL, "the FP64 entry point supports neither FAST_MATH nor WARM_START");
    if (want_srh && !c.p_shared && c.start)
        return fail(XCAPE_EINVAL, "start_lev with per-column p is for CAPE calls only (loopcape_pl)");
    if (!(c.flags & XCAPE_DEVICE_PTRS) && c.start) {
        for (int64_t i = 0; i < c.ncol; ++i)
            if (c.start[i] < 1 || c.start[i] > c.nk)
                return fail(XCAPE_EINVAL, "start_lev[%lld]=%d outside 1..%lld", (long long)i,
                            c.start[i], (long long)c.nk);
    }
    if ((rc = ensure_device())) return rc;
    if (c.ncol == 0) return XCAPE_OK;
    const bool dev = c.flags & XCAPE_DEVICE_PTRS;
    if (c.noalloc && (!dev || c.stats || !want_cape || want_srh || c.fp64 || c.p_lfc))
        return fail(XCAPE_EINVAL, "the no-allocation entry point takes device pointers, "
                                  "computes CAPE only and collects no stats");
    // NULL stream: HIP's default stream for device-pointer calls (the
    // caller's work on it is ordered before ours), the calling thread's own
    // stream for host-pointer calls (synchronous anyway; keeps concurrent
    // callers apart)
    hipStream_t s = c.stream ? (hipStream_t)c.stream : (dev ? (hipStream_t)0 : thread_stream());
    const size_t osz = c.fp64 ? sizeof(double) : sizeof(float);
    const size_t n2 = (size_t)c.nk * (size_t)c.ncol;
    const size_t np = c.p_shared ? (size_t)c.nk : n2;
    const size_t ncol = (size_t)c.ncol;

    DevBuf<float> dp_, dt_, dtd_, dps_, dts_, dtds_, du_, dv_, dus_, dvs_;
    DevBuf<int32_t> dst_, dmul_;
    DevBuf<unsigned char> dcape_, dcin_, dz_, drm_, dlm_, dmw_, dlfc_, del_;
    DevBuf<double> dsrm_, dslm_;
    DevBuf<unsigned long long> dstats_;
    CapeGridArgs a{};
    a.nk_in = c.nk;
    a.ncol = c.ncol;
    a.ls = (c.flags & XCAPE_LEVEL_MAJOR) ? c.ncol : 1;
    a.in_f64 = f64in ? 1 : 0;
    a.pinc = (float)c.pinc;
    a.ml_depth = (float)c.ml_depth;
    a.source = c.source;
    a.adiabat = c.adiabat;
    a.depth = c.depth;
    a.warm = (c.flags & XCAPE_WARM_START) ? ((c.flags & XCAPE_NO_PREDICTION) ? 2 : 1) : 0;
    a.fallback_tol = 1.0e-3f;
    a.surrogate = (a.warm && !(c.flags & XCAPE_NO_SURROGATE)) ? surrogate_device() : nullptr;
    hipError_t e = hipSuccess;
    const int math = c.fp64 ? 2 : ((c.flags & XCAPE_FAST_MATH) ? 1 : 0);
    // host arrays in column order stream through pinned staging; level-major
    // host arrays go whole through stream-ordered buffers
    if (!dev && !(c.flags & XCAPE_LEVEL_MAJOR) && !f64in)
        return stream_host_grid(c, a, want_cape, want_srh, math);
    // input floats per element (double inputs: two)
    const size_t iw = f64in ? 2 : 1;
    if (dev) {
        a.p = c.p_shared ? nullptr : c.p;
        a.p1d = c.p_shared ? c.p : nullptr;
        a.start = c.start;
        a.t = c.t; a.td = c.td; a.ps = c.ps; a.ts = c.ts; a.tds = c.tds;
        a.cape = c.cape; a.cin = c.cin; a.zout = c.zout; a.mulvl = c.mulev;
        a.p_lfc = c.p_lfc; a.p_el = c.p_el;
        a.u = c.u; a.v = c.v; a.us = c.us; a.vs = c.vs;
        a.srh_rm = c.srh_rm; a.srh_lm = c.srh_lm; a.rm = c.rm; a.lm = c.lm; a.mean6 = c.mean6;
    } else {
#define XC_TRY(x) if ((e = (x)) != hipSuccess) return hip_fail(e, #x)
        XC_TRY(dp_.alloc(np * iw, s)); XC_TRY(dp_.upload(c.p, np * iw, s));
        XC_TRY(dt_.alloc(n2 * iw, s)); XC_TRY(dt_.upload(c.t, n2 * iw, s));
        XC_TRY(dtd_.alloc(n2 * iw, s)); XC_TRY(dtd_.upload(c.td, n2 * iw, s));
        XC_TRY(dps_.alloc(ncol * iw, s)); XC_TRY(dps_.upload(c.ps, ncol * iw, s));
        XC_TRY(dts_.alloc(ncol * iw, s)); XC_TRY(dts_.upload(c.ts, ncol * iw, s));
        XC_TRY(dtds_.alloc(ncol * iw, s)); XC_TRY(dtds_.upload(c.tds, ncol * iw, s));
        if (c.start) {
            XC_TRY(dst_.alloc(ncol, s)); XC_TRY(dst_.upload(c.start, ncol, s));
        }
        a.p = c.p_shared ? nullptr : dp_.ptr;
        a.p1d = c.p_shared ? dp_.ptr : nullptr;
        a.start = c.start ? dst_.ptr : nullptr;
        a.t = dt_.ptr; a.td = dtd_.ptr; a.ps = dps_.ptr; a.ts = dts_.ptr; a.tds = dtds_.ptr;
        if (want_cape) {
            XC_TRY(dcape_.alloc(ncol * osz, s)); XC_TRY(dcin_.alloc(ncol * osz, s));
            XC_TRY(dz_.alloc(ncol * osz, s)); XC_TRY(dmul_.alloc(ncol, s));
            a.cape = dcape_.ptr; a.cin = dcin_.ptr; a.zout = dz_.ptr; a.mulvl = dmul_.ptr;
            if (c.p_lfc) {
                XC_TRY(dlfc_.alloc(ncol * osz, s)); XC_TRY(del_.alloc(ncol * osz, s));
                a.p_lfc = dlfc_.ptr; a.p_el = del_.ptr;
            }
        }
        if (want_srh) {
            XC_TRY(du_.alloc(n2 * iw, s)); XC_TRY(du_.upload(c.u, n2 * iw, s));
            XC_TRY(dv_.alloc(n2 * iw, s)); XC_TRY(dv_.upload(c.v, n2 * iw, s));
            XC_TRY(dus_.alloc(ncol * iw, s)); XC_TRY(dus_.upload(c.us, ncol * iw, s));
            XC_TRY(dvs_.alloc(ncol * iw, s)); XC_TRY(dvs_.upload(c.vs, ncol * iw, s));
            XC_TRY(dsrm_.alloc(ncol, s)); XC_TRY(dslm_.alloc(ncol, s));
            a.u = du_.ptr; a.v = dv_.ptr; a.us = dus_.ptr; a.vs = dvs_.ptr;
            a.srh_rm = dsrm_.ptr; a.srh_lm = dslm_.ptr;
            if (c.rm) {
                XC_TRY(drm_.alloc(2 * ncol * osz, s)); XC_TRY(dlm_.alloc(2 * ncol * osz, s));
                XC_TRY(dmw_.alloc(2 * ncol * osz, s));
                a.rm = drm_.ptr; a.lm = dlm_.ptr; a.mean6 = dmw_.ptr;
            }
        }
    }
    if (c.stats && want_cape) {
        XC_TRY(dstats_.alloc(ST_COUNT, s));
        XC_TRY(zero_counters(dstats_.ptr, ST_COUNT, s));
        a.stats = dstats_.ptr;
    }
    if (c.noalloc) {
        // the whole call's scratch in the caller-independent array: the
        // warm ascent's counters first, then the staged profiles and records
        const size_t per = (size_t)grid_scratch_floats(a, true, false, math) * sizeof(float);
        const size_t need = NOALLOC_HEAD + per * (size_t)c.ncol;
        if (need > NOALLOC_SCRATCH_BYTES)
            return fail(XCAPE_EINVAL,
                        "xcape_cape_noalloc_f32: %lld columns of %lld levels need %zu bytes of "
                        "scratch; the limit is %zu (at most %zu such columns per call)",
                        (long long)c.ncol, (long long)c.nk, need, NOALLOC_SCRATCH_BYTES,
                        (NOALLOC_SCRATCH_BYTES - NOALLOC_HEAD) / per);
        e = run_grid(a, want_cape, want_srh, math, s,
                     reinterpret_cast<float *>(c.noalloc + NOALLOC_HEAD),
                     reinterpret_cast<unsigned int *>(c.noalloc));
        return e != hipSuccess ? hip_fail(e, "kernel launch") : XCAPE_OK;
    }
    if ((e = run_grid(a, want_cape, want_srh, math, s)) != hipSuccess)
        return hip_fail(e, "kernel launch");
    if (!dev) {
        if (want_cape) {
            XC_TRY(hipMemcpyAsync(c.cape, dcape_.ptr, ncol * osz, hipMemcpyDeviceToHost, s));
            XC_TRY(hipMemcpyAsync(c.cin, dcin_.ptr, ncol * osz, hipMemcpyDeviceToHost, s));
            XC_TRY(hipMemcpyAsync(c.zout, dz_.ptr, ncol * osz, hipMemcpyDeviceToHost, s));
            XC_TRY(dmul_.download(c.mulev, ncol, s));
            if (c.p_lfc) {
                XC_TRY(hipMemcpyAsync(c.p_lfc, dlfc_.ptr, ncol * osz, hipMemcpyDeviceToHost, s));
                XC_TRY(hipMemcpyAsync(c.p_el, del_.ptr, ncol * osz, hipMemcpyDeviceToHost, s));
            }
        }
        if (want_srh) {
            XC_TRY(dsrm_.download(c.srh_rm, ncol, s));
            XC_TRY(dslm_.download(c.srh_lm, ncol, s));
            if (c.rm) {
                const size_t nb = 2 * ncol * osz;
                XC_TRY(hipMemcpyAsync(c.rm, drm_.ptr, nb, hipMemcpyDeviceToHost, s));
                XC_TRY(hipMemcpyAsync(c.lm, dlm_.ptr, nb, hipMemcpyDeviceToHost, s));
                XC_TRY(hipMemcpyAsync(c.mean6, dmw_.ptr, nb, hipMemcpyDeviceToHost, s));
            }
        }
    }
    if (a.stats) {
        unsigned long long h[ST_COUNT];
        XC_TRY(hipMemcpyAsync(h, dstats_.ptr, sizeof h, hipMemcpyDeviceToHost, s));
        XC_TRY(hipStreamSynchronize(s));
        copy_stats(h, c.stats);
    } else if (!dev) {
        XC_TRY(hipStreamSynchronize(s));
    }
#undef XC_TRY
    return XCAPE_OK;
}

}  // namespace

extern "C" {

const char *xcape_version(void) { return "xcape-hip 0.1.0 (gfx950)"; }

const char *xcape_last_error(void) { return xcape::g_err.c_str(); }

int xcape_release_host_staging(void)
{
    std::vector<StagingSet *> idle;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        idle.swap(g_stage_idle);
        g_stage_sets -= (int)idle.size();
    }
    g_stage_cv.notify_all();
    int dv = -1;
    (void)hipGetDevice(&dv);
    for (StagingSet *st : idle) {
        if (st->device >= 0) (void)hipSetDevice(st->device);
        st->destroy();
        delete st;
    }
    if (dv >= 0) (void)hipSetDevice(dv);
    return (int)idle.size();
}

int xcape_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int xcape_cape_f32(const float *p, int p_shared, const int32_t *start_lev, const float *t,
                   const float *td, const float *ps, const float *ts, const float *tds,
                   int64_t nk, int64_t ncol, float pinc, int source, float ml_depth,
                   int adiabat, float *cape, float *cin, int32_t *mulev, float *zmulev,
                   uint32_t flags, xcape_stats *stats, void *stream)
{
    GridCall c;
    c.kind = GRID_CAPE;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.nk = nk; c.ncol = ncol; c.pinc = pinc; c.source = source; c.ml_depth = ml_depth;
    c.adiabat = adiabat;
    c.cape = cape; c.cin = cin; c.mulev = mulev; c.zout = zmulev;
    c.flags = flags; c.stats = stats; c.stream = stream;
    return grid_entry(c);
}

int xcape_cape_noalloc_f32(const float *p, int p_shared, const int32_t *start_lev,
                           const float *t, const float *td, const float *ps, const float *ts,
                           const float *tds, int64_t nk, int64_t ncol, float pinc, int source,
                           float ml_depth, int adiabat, float *cape, float *cin,
                           int32_t *mulev, float *zmulev, uint32_t flags, xcape_stats *stats,
                           void *stream)
{
    GridCall c;
    c.kind = GRID_CAPE;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.nk = nk; c.ncol = ncol; c.pinc = pinc; c.source = source; c.ml_depth = ml_depth;
    c.adiabat = adiabat;
    c.cape = cape; c.cin = cin; c.mulev = mulev; c.zout = zmulev;
    c.flags = flags; c.stats = stats; c.stream = stream;
    if (!(flags & XCAPE_DEVICE_PTRS))
        return fail(XCAPE_EINVAL, "xcape_cape_noalloc_f32 takes device pointers only "
                                  "(XCAPE_DEVICE_PTRS)");
    if (stats) return fail(XCAPE_EINVAL, "xcape_cape_noalloc_f32 collects no stats");
    int rc = ensure_device();
    if (rc || (rc = noalloc_scratch(&c.noalloc))) return rc;
    return grid_entry(c);
}

int xcape_cape_levels_f32(const float *p, int p_shared, const int32_t *start_lev,
                          const float *t, const float *td, const float *ps, const float *ts,
                          const float *tds, int64_t nk, int64_t ncol, float pinc, int source,
                          float ml_depth, int adiabat, float *cape, float *cin, int32_t *mulev,
                          float *zmulev, float *p_lfc, float *p_el, uint32_t flags,
                          void *stream)
{
    GridCall c;
    c.kind = GRID_CAPE;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.nk = nk; c.ncol = ncol; c.pinc = pinc; c.source = source; c.ml_depth = ml_depth;
    c.adiabat = adiabat;
    c.cape = cape; c.cin = cin; c.mulev = mulev; c.zout = zmulev;
    c.p_lfc = p_lfc; c.p_el = p_el;
    if (ncol > 0 && (!p_lfc || !p_el)) return fail(XCAPE_EINVAL, "NULL LFC/EL output array");
    c.flags = flags; c.stream = stream;
    return grid_entry(c);
}

int xcape_cape_levels_f64(const float *p, int p_shared, const int32_t *start_lev,
                          const float *t, const float *td, const float *ps, const float *ts,
                          const float *tds, int64_t nk, int64_t ncol, double pinc, int source,
                          double ml_depth, int adiabat, double *cape, double *cin,
                          int32_t *mulev, double *zmulev, double *p_lfc, double *p_el,
                          uint32_t flags, void *stream)
{
    GridCall c;
    c.kind = GRID_CAPE;
    c.fp64 = true;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.nk = nk; c.ncol = ncol; c.pinc = pinc; c.source = source; c.ml_depth = ml_depth;
    c.adiabat = adiabat;
    c.cape = cape; c.cin = cin; c.mulev = mulev; c.zout = zmulev;
    c.p_lfc = p_lfc; c.p_el = p_el;
    if (ncol > 0 && (!p_lfc || !p_el)) return fail(XCAPE_EINVAL, "NULL LFC/EL output array");
    c.flags = flags; c.stream = stream;
    return grid_entry(c);
}

int xcape_cape_f64(const float *p, int p_shared, const int32_t *start_lev, const float *t,
                   const float *td, const float *ps, const float *ts, const float *tds,
                   int64_t nk, int64_t ncol, double pinc, int source, double ml_depth,
                   int adiabat, double *cape, double *cin, int32_t *mulev, double *zmulev,
                   uint32_t flags, xcape_stats *stats, void *stream)
{
    GridCall c;
    c.kind = GRID_CAPE;
    c.fp64 = true;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.nk = nk; c.ncol = ncol; c.pinc = pinc; c.source = source; c.ml_depth = ml_depth;
    c.adiabat = adiabat;
    c.cape = cape; c.cin = cin; c.mulev = mulev; c.zout = zmulev;
    c.flags = flags; c.stats = stats; c.stream = stream;
    return grid_entry(c);
}

int xcape_srh_f32(const float *p, int p_shared, const int32_t *start_lev, const float *t,
                  const float *td, const float *u, const float *v, const float *ps,
                  const float *ts, const float *tds, const float *us, const float *vs,
                  int64_t nk, int64_t ncol, double depth, double *srh_rm, double *srh_lm,
                  float *rm, float *lm, float *mean6, uint32_t flags, void *stream)
{
    GridCall c;
    c.kind = GRID_SRH;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.u = u; c.v = v; c.us = us; c.vs = vs;
    c.nk = nk; c.ncol = ncol; c.depth = depth;
    c.srh_rm = srh_rm; c.srh_lm = srh_lm; c.rm = rm; c.lm = lm; c.mean6 = mean6;
    c.flags = flags & ~XCAPE_FAST_MATH; c.stream = stream;
    return grid_entry(c);
}

int xcape_srh_f64(const float *p, int p_shared, const int32_t *start_lev, const float *t,
                  const float *td, const float *u, const float *v, const float *ps,
                  const float *ts, const float *tds, const float *us, const float *vs,
                  int64_t nk, int64_t ncol, double depth, double *srh_rm, double *srh_lm,
                  double *rm, double *lm, double *mean6, uint32_t flags, void *stream)
{
    GridCall c;
    c.kind = GRID_SRH;
    c.fp64 = true;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.u = u; c.v = v; c.us = us; c.vs = vs;
    c.nk = nk; c.ncol = ncol; c.depth = depth;
    c.srh_rm = srh_rm; c.srh_lm = srh_lm; c.rm = rm; c.lm = lm; c.mean6 = mean6;
    c.flags = flags; c.stream = stream;
    return grid_entry(c);
}

int xcape_cape_srh_f32(const float *p, int p_shared, const int32_t *start_lev, const float *t,
                       const float *td, const float *u, const float *v, const float *ps,
                       const float *ts, const float *tds, const float *us, const float *vs,
                       int64_t nk, int64_t ncol, float pinc, int source, float ml_depth,
                       int adiabat, double depth, float *cape, float *cin, int32_t *mulev,
                       float *zmulev, double *srh_rm, double *srh_lm, float *rm, float *lm,
                       float *mean6, uint32_t flags, xcape_stats *stats, void *stream)
{
    GridCall c;
    c.kind = GRID_CAPE_SRH;
    c.p = p; c.p_shared = p_shared; c.start = start_lev;
    c.t = t; c.td = td; c.ps = ps; c.ts = ts; c.tds = tds;
    c.u = u; c.v = v; c.us = us; c.vs = vs;
    c.nk = nk; c.ncol = ncol; c.pinc = pinc; c.source = source; c.ml_depth = ml_depth;
    c.adiabat = adiabat; c.depth = depth;
    c.cape = cape; c.cin = cin; c.mulev = mulev; c.zout = zmulev;
    c.srh_rm = srh_rm; c.srh_lm = srh_lm; c.rm = rm; c.lm = lm; c.mean6 = mean6;
    c.flags = flags; c.stats = stats; c.stream = stream;
    return grid_entry(c);
}

}  // extern "C"
