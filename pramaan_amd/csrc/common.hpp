// common.hpp -- error handling and device-buffer ownership for the C-ABI.
//
// Inside the library failures are C++ exceptions (pm::Error carries the
// C-ABI status code); every extern "C" entry point catches them, stores the
// message in a thread-local string and returns the code, so nothing ever
// unwinds across the ABI.  This plays the role of CUDA_CHECK
// (src/gpu/cuda_utils.cuh:8-15 in the reference), whose std::runtime_error
// text "<file>:<line> - <message>" the message format keeps.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pramaan_hip.h"

namespace pm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* file, int line, const char* what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory) ? PM_ERR_OUT_OF_MEMORY
                     : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? PM_ERR_NO_DEVICE
                                                                              : PM_ERR_HIP;
    fail(code, std::string("HIP error in ") + file + ":" + std::to_string(line) + " - " +
                   hipGetErrorString(e) + " (" + what + ")");
}

#define PM_HIP(call) ::pm::hip_check((call), __FILE__, __LINE__, #call)
#define PM_LAUNCHED() ::pm::hip_check(hipGetLastError(), __FILE__, __LINE__, "kernel launch")

// A named ROCTx range for the lifetime of the object: entry points and load
// phases show up on rocprofv3's marker trace (--marker-trace); without a
// tool attached a push or pop is a call that returns at once.
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

void set_last_error(const std::string& msg);
void clear_last_error();

// Runs `body` and converts any exception into a C-ABI status code.
template <class F>
int guarded(F&& body) {
    try {
        clear_last_error();
        body();
        return PM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return PM_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return PM_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown exception");
        return PM_ERR_INTERNAL;
    }
}

// Owning device array.  Freed on scope exit, also when an exception
// unwinds a solve (the reference's CudaPtr, pdhg_solver.cu:94-105).
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void resize(size_t n) {
        if (n == n_ && p_) return;
        reset();
        if (n == 0) return;
        PM_HIP(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
        n_ = n;
    }
    void reset() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T* h, size_t n, hipStream_t s) {
        if (n == 0) return;
        resize(n);
        PM_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
    void download(T* h, size_t n, hipStream_t s) const {
        if (n == 0) return;
        PM_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Page-locked host buffer for the small per-check readbacks.
template <class T>
class PinnedBuf {
public:
    PinnedBuf() = default;
    ~PinnedBuf() {
        if (p_) (void)hipHostFree(p_);
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    void resize(size_t n) {
        if (n <= n_) return;
        if (p_) (void)hipHostFree(p_);
        p_ = nullptr;
        PM_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
        n_ = n;
    }
    T* get() const { return p_; }
    T& operator[](size_t i) const { return p_[i]; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

}  // namespace pm
