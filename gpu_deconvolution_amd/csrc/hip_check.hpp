// Error plumbing shared by every translation unit of the library.
//
// Internally every HIP failure throws fmvd::Error; each exported C entry point
// catches it, releases what it owns and reports it once through the
// process-global handler installed by setErrorHandler() (default: print and
// exit(EXIT_FAILURE), as fmvd_cuda_utils.cpp:12-16 does).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace fmvd {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);
[[noreturn]] void throw_msg(const std::string& msg);

// Reports an exception through the installed handler (called by the C ABI).
void report_error(const char* msg);

}  // namespace fmvd

#define FMVD_HIP(call)                                                        \
    do {                                                                      \
        hipError_t fmvd_e_ = (call);                                          \
        if (fmvd_e_ != hipSuccess) ::fmvd::throw_hip(fmvd_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define FMVD_LAUNCH_CHECK()                                                   \
    do {                                                                      \
        hipError_t fmvd_e_ = hipGetLastError();                               \
        if (fmvd_e_ != hipSuccess) ::fmvd::throw_hip(fmvd_e_, "kernel launch", __FILE__, __LINE__); \
    } while (0)

// Wraps the body of an exported C function: converts exceptions into one
// report through the error handler and returns `fail`.
#define FMVD_API_GUARD(fail, ...)                                             \
    try {                                                                     \
        __VA_ARGS__                                                           \
    } catch (const std::exception& fmvd_ex_) {                                \
        ::fmvd::report_error(fmvd_ex_.what());                                \
        return fail;                                                          \
    }
