// Error handler, device helpers, pinned host memory and kernel timing.
// Restates fmvd_cuda_utils.cpp:12-85 and fmvd_deconvolve.cpp:206-218 of the
// reference for HIP.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fmvd_cuda_utils.h"
#include "../../include/fmvd_deconvolve.h"
#include "hip_check.hpp"
#include "timing.hpp"

namespace {

// fmvd_cuda_utils.cpp:12-16: default handler prints and exits.
void exit_on_error(void*, const char*)
{
    printf("exiting... \n");
    fflush(stdout);
    exit(EXIT_FAILURE);
}

void (*g_handler)(void*, const char*) = exit_on_error;
void* g_param = nullptr;
thread_local std::string t_last_error;

}  // namespace

namespace fmvd {

void throw_hip(hipError_t e, const char* what, const char* file, int line)
{
    // The runtime keeps the last error of any call until hipGetLastError reads
    // it; this error is reported now, so clear it here or the next launch
    // check would report it again.
    (void)hipGetLastError();
    char msg[512];
    const char* base = strrchr(file, '/');
    snprintf(msg, sizeof msg, "HIP error %d: %s in %s (line %i): %s", (int)e, hipGetErrorString(e),
             base ? base + 1 : file, line, what);
    throw Error(msg);
}

void throw_msg(const std::string& msg) { throw Error(msg); }

void report_error(const char* msg)
{
    t_last_error = msg;
    fprintf(stderr, "%s\n", msg);
    g_handler(g_param, msg);
}

// ------------------------------------------------------------ timing
namespace {
struct TimedLaunch {
    int kind;
    hipEvent_t a, b;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<TimedLaunch> g_pending;
std::vector<hipEvent_t> g_pool;
double g_total_ms[kTimingKinds] = {0};
long long g_count[kTimingKinds] = {0};

hipEvent_t pool_get()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    FMVD_HIP(hipEventCreate(&e));
    return e;
}

// A drain error raised inside ~TimedScope, reported by the next timing read.
std::string g_drain_error;

void drain_locked()
{
    std::vector<TimedLaunch> pending;
    pending.swap(g_pending);  // a failing pair is dropped, not retried forever
    for (auto& t : pending) {
        FMVD_HIP(hipEventSynchronize(t.b));
        float ms = 0.f;
        FMVD_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        g_total_ms[t.kind] += ms;
        g_count[t.kind] += 1;
        g_pool.push_back(t.a);
        g_pool.push_back(t.b);
    }
}

// Reports (throws) a drain error kept from a destructor.
void raise_kept_error_locked()
{
    if (g_drain_error.empty()) return;
    std::string msg;
    msg.swap(g_drain_error);
    throw_msg(msg);
}
}  // namespace

bool timing_on() { return g_timing; }

thread_local hipEvent_t t_start = nullptr, t_stop = nullptr;

void take_launch_events(hipEvent_t& start, hipEvent_t& stop)
{
    start = t_start;
    stop = t_stop;
    t_start = t_stop = nullptr;
}

TimedScope::TimedScope(int kind, hipStream_t s, bool sample) : kind_(kind), s_(s), on_(g_timing && sample)
{
    if (!on_) return;
    std::lock_guard<std::mutex> lk(g_tmu);
    a_ = pool_get();
    b_ = pool_get();
    t_start = a_;
    t_stop = b_;
}

TimedScope::~TimedScope() noexcept
{
    if (!on_) return;
    std::lock_guard<std::mutex> lk(g_tmu);
    try {
        if (t_start == a_) {  // the scope launched nothing: no sample
            t_start = t_stop = nullptr;
            g_pool.push_back(a_);
            g_pool.push_back(b_);
            return;
        }
        g_pending.push_back({kind_, a_, b_});
        if (g_pending.size() > 4096) drain_locked();
    } catch (const std::exception& e) {
        if (g_drain_error.empty()) g_drain_error = e.what();
    }
}

}  // namespace fmvd

extern "C" {

void setErrorHandler(void (*handler)(void*, const char*), void* param)
{
    g_param = param;
    g_handler = handler ? handler : exit_on_error;
}

int iDivUp(int a, int b) { return (a % b != 0) ? (a / b + 1) : (a / b); }

int getNumCUDADevices(void)
{
    FMVD_API_GUARD(0, {
        int n = 0;
        FMVD_HIP(hipGetDeviceCount(&n));
        return n;
    })
}

void getCudaDeviceName(int dev, char* name)
{
    FMVD_API_GUARD(, {
        hipDeviceProp_t prop;
        FMVD_HIP(hipGetDeviceProperties(&prop, dev));
        memset(name, 0, 256);
        // some ROCm builds leave the marketing name empty; the ISA name is
        // then the most useful identifier
        strncpy(name, prop.name[0] ? prop.name : prop.gcnArchName, 255);
    })
}

void setCudaDevice(int dev)
{
    FMVD_API_GUARD(, { FMVD_HIP(hipSetDevice(dev)); })
}

int fmvd_last_error(char* buf, int len)
{
    if (buf && len > 0) {
        strncpy(buf, t_last_error.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int)t_last_error.size();
}

// fmvd_deconvolve.cpp:206-218
void* fmvd_malloc(size_t size)
{
    FMVD_API_GUARD(nullptr, {
        void* p = nullptr;
        FMVD_HIP(hipHostMalloc(&p, size, hipHostMallocDefault));
        return p;
    })
}

void fmvd_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

void fmvd_kernel_timing_enable(int on)
{
    std::lock_guard<std::mutex> lk(fmvd::g_tmu);
    fmvd::g_timing = on != 0;
}

void fmvd_kernel_timing_reset(void)
{
    std::lock_guard<std::mutex> lk(fmvd::g_tmu);
    try {
        fmvd::drain_locked();
        fmvd::raise_kept_error_locked();
    } catch (const std::exception& e) {
        fmvd::report_error(e.what());
    }
    for (int k = 0; k < fmvd::kTimingKinds; ++k) {
        fmvd::g_total_ms[k] = 0;
        fmvd::g_count[k] = 0;
    }
}

int fmvd_kernel_timing_read(int kind, double* total_ms, long long* launches)
{
    if (kind < 0 || kind >= fmvd::kTimingKinds) return -1;
    std::lock_guard<std::mutex> lk(fmvd::g_tmu);
    FMVD_API_GUARD(-1, {
        fmvd::drain_locked();
        fmvd::raise_kept_error_locked();
        if (total_ms) *total_ms = fmvd::g_total_ms[kind];
        if (launches) *launches = fmvd::g_count[kind];
        return 0;
    })
}

}  // extern "C"
