// Plumbing shared by the translation units behind include/m3d_device.h: the
// context object, error propagation (status codes plus m3d_last_error) and an
// owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "m3d_device.h"

// A context lives until it has been destroyed by the caller and its last
// dependent object (operator, station set) is gone, whichever comes later.
struct m3d_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int operators = 0;  // live objects created on this context
    bool closed = false;
};

namespace m3d {
namespace abi {

struct Fail : std::runtime_error {
    int code;
    Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::m3d::abi::Fail(M3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Message slot of m3d_last_error (one per host thread).
std::string& last_error();
// Dependent-object bookkeeping of a context under one lock (objects may be
// created, moved to another context and destroyed from several host
// threads): attach counts an object in, detach counts it out and frees a
// destroyed context once nothing created on it is alive, close marks the
// context destroyed and frees it when nothing depends on it.
void context_attach(m3d_context* ctx);
void context_detach(m3d_context* ctx);
void context_close(m3d_context* ctx);

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return M3D_OK;
    } catch (const Fail& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return M3D_ERR_ARG;
    }
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Fail(M3D_ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        if (!p || n != std::max<size_t>(h.size(), 1)) alloc(h.size());
        if (!h.empty()) HIPCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

}  // namespace abi
}  // namespace m3d
