// Error handling and device buffers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "simplic_amd.h"

namespace svp {
namespace api {

// message of the last failed call on this thread (svp_last_error)
inline thread_local std::string g_err;

struct SvpError : std::runtime_error {
    int code;
    SvpError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                            \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            throw ::svp::api::SvpError(SVP_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return SVP_OK;
    } catch (const SvpError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return SVP_E_ARG;
    }
}

// Owning device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* alloc(size_t n) {
        reset();
        bytes = std::max<size_t>(n, 1) * sizeof(T);
        HIPCHECK(hipMalloc(&p, bytes));
        // complete before return: the handles' work runs on non-blocking
        // streams, which do not wait for the null stream
        HIPCHECK(hipMemsetAsync(p, 0, bytes, nullptr));
        HIPCHECK(hipStreamSynchronize(nullptr));
        return (T*)p;
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* d = alloc<T>(v.size());
        if (!v.empty()) HIPCHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return d;
    }
};

}  // namespace api
}  // namespace svp
