// common.hpp -- shared host/device helpers for the mapp4py_amd HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

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

// Per-kernel timing with HIP events recorded on the stream the kernels run
// on; enabled on request, read back after a synchronisation.
enum KernelId : int {
    K_EAM_RHO = 0,
    K_EAM_FORCE,
    K_BUILD_LIST,
    K_CELL_SORT,
    K_INTEGRATE,
    K_EAM_ENERGY,
    K_DMD_PASS1,
    K_DMD_PASS2,
    K_DMD_ENERGY,
    K_SURR_INFER,
    K_SURR_TRAIN,
    K_GCMC,
    K_DMD_CD,
    K_DMD_JSTATE,
    K_DMD_J,
    K_DMD_NEWTON,
    K_MD_MIN,
    K_DMD_PREPJ,
    K_DMD_JXA,
    K_REPLICA,
    K_COUNT
};

inline const char* kernel_name(int k)
{
    static const char* names[K_COUNT] = {"eam_rho",    "eam_force",  "build_list", "cell_sort",
                                         "integrate",  "eam_energy", "dmd_pass1",  "dmd_pass2",
                                         "dmd_energy", "surr_infer", "surr_train", "gcmc",
                                         "dmd_cd",     "dmd_j_state", "dmd_j",
                                         "dmd_newton", "md_min_vec", "dmd_prepj",
                                         "dmd_jxa",    "replica"};
    return (k >= 0 && k < K_COUNT) ? names[k] : "";
}

class Profiler {
public:
    bool on = false;
    double total_ms[K_COUNT] = {};
    long launches[K_COUNT] = {};
    ~Profiler()
    {
        for (auto& r : recs_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool_) (void)hipEventDestroy(e);
    }
    void reset()
    {
        flush();
        for (int k = 0; k < K_COUNT; ++k) {
            total_ms[k] = 0.0;
            launches[k] = 0;
        }
    }
    void begin(int kid, hipStream_t s)
    {
        if (!on) return;
        Rec r{kid, get(), get()};
        MAPP_HIP(hipEventRecord(r.a, s));
        recs_.push_back(r);
    }
    void end(hipStream_t s)
    {
        if (!on || recs_.empty()) return;
        MAPP_HIP(hipEventRecord(recs_.back().b, s));
    }
    // waits for the recorded events and folds them into the totals
    void flush()
    {
        for (auto& r : recs_) {
            MAPP_HIP(hipEventSynchronize(r.b));
            float ms = 0.f;
            MAPP_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            total_ms[r.kid] += ms;
            launches[r.kid] += 1;
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        recs_.clear();
    }

private:
    struct Rec {
        int kid;
        hipEvent_t a, b;
    };
    hipEvent_t get()
    {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        MAPP_HIP(hipEventCreate(&e));
        return e;
    }
    std::vector<Rec> recs_;
    std::vector<hipEvent_t> pool_;
};

}  // namespace mapp
