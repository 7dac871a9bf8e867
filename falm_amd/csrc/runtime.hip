// runtime.hip — device selection, memory, error reporting, reduction
// scratch and per-family kernel timing (the falmMalloc*/falmMemcpy/
// falmWaitStream layer of the reference, src/dev/cuhandler.h:12-118).
#include <cstdarg>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace falm {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

falm_status hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return FALM_OK;
    set_error("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    if (e == hipErrorOutOfMemory) return FALM_ERR_MALLOC;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return FALM_ERR_NO_DEVICE;
    return FALM_ERR_HIP + (int)e;
}

falm_status check_domain(const falm_domain *d) {
    if (!d) {
        set_error("null domain");
        return FALM_ERR_INVALID;
    }
    for (int m = 0; m < 3; m++) {
        if (d->shape[m] < 2 * d->gc + 1 || d->gshape[m] < d->shape[m] || d->offset[m] < 0) {
            set_error("bad domain: shape %d %d %d gc %d", d->shape[0], d->shape[1], d->shape[2], d->gc);
            return FALM_ERR_INVALID;
        }
    }
    if (d->gc < 2) {
        set_error("bad domain: gc %d < 2 (the advection stencil has radius 2)", d->gc);
        return FALM_ERR_INVALID;
    }
    // one column must be indexable with 32-bit offsets
    if ((double)d->shape[0] * d->shape[1] * d->shape[2] > 2147483647.0) {
        set_error("bad domain: more than 2^31 cells per column");
        return FALM_ERR_INVALID;
    }
    return FALM_OK;
}

// ------------------------------------------------------- reduce scratch
static std::mutex g_mx;
static std::map<int, ReduceScratch> g_scratch;

falm_status reduce_scratch(size_t nblocks, ReduceScratch **out) {
    int dev = 0;
    FALM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mx);
    ReduceScratch &rs = g_scratch[dev];
    if (rs.max_blocks < nblocks || !rs.partials) {
        if (rs.partials) FALM_HIP(hipFree(rs.partials));
        size_t nb = nblocks < 1024 ? 1024 : nblocks;
        FALM_HIP(hipMalloc(&rs.partials, 8 * nb * sizeof(double)));
        rs.max_blocks = nb;
    }
    if (!rs.result) {
        FALM_HIP(hipMalloc(&rs.result, 8 * sizeof(double)));
        FALM_HIP(hipHostMalloc(&rs.host, 8 * sizeof(double), hipHostMallocDefault));
    }
    *out = &rs;
    return FALM_OK;
}

// --------------------------------------------------------------- timing
struct TimingState {
    int family = FALM_TIME_NONE;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    int launches = 0;
    double ms = 0.0, bytes = 0.0;
    hipEvent_t cur = nullptr;
};
static TimingState g_tm;

static hipEvent_t get_event() {
    if (!g_tm.pool.empty()) {
        hipEvent_t e = g_tm.pool.back();
        g_tm.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    // timing only: no system-scope fence (cache write-back) at the record
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

TimedLaunch::TimedLaunch(int fam, hipStream_t s) : family(fam), stream(s), on(false) {
    if (g_tm.family != FALM_TIME_NONE && g_tm.family == fam) {
        g_tm.cur = get_event();
        if (g_tm.cur && hipEventRecord(g_tm.cur, s) == hipSuccess) on = true;
    }
}

void TimedLaunch::done(double bytes) {
    if (!on) return;
    hipEvent_t e = get_event();
    if (!e || hipEventRecord(e, stream) != hipSuccess) return;
    g_tm.pending.push_back({g_tm.cur, e});
    g_tm.launches++;
    g_tm.bytes += bytes;
    // fold completed pairs now and then to bound the pending list
    if (g_tm.pending.size() > 4096) {
        for (auto &p : g_tm.pending) {
            float ms = 0.f;
            (void)hipEventSynchronize(p.second);
            (void)hipEventElapsedTime(&ms, p.first, p.second);
            g_tm.ms += ms;
            g_tm.pool.push_back(p.first);
            g_tm.pool.push_back(p.second);
        }
        g_tm.pending.clear();
    }
}

void TimedLaunch::add_bytes(double bytes) {
    if (on) g_tm.bytes += bytes;
}

}  // namespace falm

using namespace falm;

extern "C" {

const char *falm_last_error(void) { return g_err; }

const char *falm_status_string(falm_status s) {
    switch (s) {
    case FALM_OK: return "ok";
    case FALM_ERR_MALLOC: return "allocation failed";
    case FALM_ERR_SETUP: return "setup error";
    case FALM_ERR_INVALID: return "invalid argument";
    case FALM_ERR_ALIAS: return "output aliases an input";
    case FALM_ERR_UNSUPPORTED: return "unsupported option";
    case FALM_ERR_NONFINITE: return "non-finite value";
    case FALM_ERR_NOT_CONVERGED: return "solver did not converge";
    case FALM_ERR_BREAKDOWN: return "BiCGStab breakdown";
    case FALM_ERR_NO_DEVICE: return "no HIP device";
    default: return s >= FALM_ERR_HIP ? "HIP error" : "unknown status";
    }
}

falm_status falm_init(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device available");
        return FALM_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        set_error("device %d out of range (%d devices)", device, n);
        return FALM_ERR_INVALID;
    }
    FALM_HIP(hipSetDevice(device));
    return noalloc_setup(nullptr);  // falm_dot_noalloc's workspace address on this device
}

falm_status falm_malloc(void **ptr, size_t bytes) {
    if (!ptr) return FALM_ERR_INVALID;
    *ptr = nullptr;
    if (!bytes) return FALM_OK;
    FALM_HIP(hipMalloc(ptr, bytes));
    FALM_HIP(hipMemset(*ptr, 0, bytes));  // Matrix::alloc zero-fills (src/matrix.h:85-98)
    // complete the fill before work on any stream (non-blocking ones do not
    // wait for the null stream)
    FALM_HIP(hipDeviceSynchronize());
    return FALM_OK;
}

falm_status falm_free(void *ptr) {
    if (ptr) FALM_HIP(hipFree(ptr));
    return FALM_OK;
}

falm_status falm_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    if (!bytes) return FALM_OK;
    if (!dst || !src) return FALM_ERR_INVALID;
    hipMemcpyKind k;
    switch (kind) {
    case 1: k = hipMemcpyHostToHost; break;
    case 2: k = hipMemcpyHostToDevice; break;
    case 4: k = hipMemcpyDeviceToHost; break;
    case 8: k = hipMemcpyDeviceToDevice; break;
    default: set_error("falm_memcpy: bad kind %d", kind); return FALM_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    FALM_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
    FALM_HIP(hipStreamSynchronize(s));
    return FALM_OK;
}

falm_status falm_memset(void *ptr, int value, size_t bytes, void *stream) {
    if (!bytes) return FALM_OK;
    if (!ptr) return FALM_ERR_INVALID;
    FALM_HIP(hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream));
    return FALM_OK;
}

falm_status falm_stream_sync(void *stream) {
    FALM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return FALM_OK;
}

falm_status falm_timing_enable(int family) {
    for (auto &p : g_tm.pending) {
        g_tm.pool.push_back(p.first);
        g_tm.pool.push_back(p.second);
    }
    g_tm.pending.clear();
    g_tm.family = family;
    g_tm.launches = 0;
    g_tm.ms = 0.0;
    g_tm.bytes = 0.0;
    return FALM_OK;
}

falm_status falm_timing_read(int *launches, double *ms, double *bytes) {
    for (auto &p : g_tm.pending) {
        float m = 0.f;
        FALM_HIP(hipEventSynchronize(p.second));
        FALM_HIP(hipEventElapsedTime(&m, p.first, p.second));
        g_tm.ms += m;
        g_tm.pool.push_back(p.first);
        g_tm.pool.push_back(p.second);
    }
    g_tm.pending.clear();
    if (launches) *launches = g_tm.launches;
    if (ms) *ms = g_tm.ms;
    if (bytes) *bytes = g_tm.bytes;
    return FALM_OK;
}

}  // extern "C"
