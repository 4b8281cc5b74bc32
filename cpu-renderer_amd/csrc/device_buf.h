// device_buf.h — HIP error mapping and owned device buffers of the host side.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace cr {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// A transport failure inside a render (reference src/render.cpp:429-451).
struct PathError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw std::bad_alloc();
        throw HipError(std::string(what) + ": " + hipGetErrorString(e));
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        check(hipMalloc(&p, std::max<size_t>(n, 16)), "hipMalloc");
        bytes = std::max<size_t>(n, 16);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st), "upload");
}

}  // namespace cr
