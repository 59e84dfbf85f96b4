// common.hpp -- shared host/device helpers for the mapp4py_amd HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace mapp {

// Errors cross the C ABI as a status code plus a thread-local message; inside
// the library they travel as mapp::Error exceptions.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
    OK = 0,
    ERR_ARG = 1,      // invalid argument / shape
    ERR_FILE = 2,     // potential file could not be read or parsed
    ERR_HIP = 3,      // HIP runtime failure
    ERR_CAPACITY = 4, // a size limit of the implementation was exceeded
    ERR_STATE = 5,    // call order violated
};

void set_last_error(const std::string& m);

#define MAPP_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e__ = (expr);                                                             \
        if (e__ != hipSuccess)                                                               \
            throw ::mapp::Error(::mapp::ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e__) + \
                                                     " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define MAPP_REQUIRE(cond, code, msg)                  \
    do {                                               \
        if (!(cond)) throw ::mapp::Error((code), (msg)); \
    } while (0)

// Device buffer owned by the library (hipMalloc / hipFree).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // grows (never shrinks); contents are not preserved
    void reserve(size_t count) {
        if (count <= n && p) return;
        release();
        size_t c = count ? count : 1;
        MAPP_HIP(hipMalloc(&p, c * sizeof(T)));
        n = c;
    }
    T* get() const { return p; }
};

constexpr int kWave = 64;

}  // namespace mapp
