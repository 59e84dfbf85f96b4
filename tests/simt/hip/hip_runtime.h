// Test infrastructure: a CPU SIMT emulation of the HIP subset that gcmc.hip
// uses, so the kernel source itself can be run and checked on a machine
// without a GPU. Every GPU thread is an OS thread; the 64 lanes of a wave
// meet at every cross-lane operation (ballot, shuffles, readfirstlane, wave
// barrier) and a workgroup at __syncthreads, which is the lockstep the kernel
// relies on. Device memory is host memory; launches run one workgroup at a
// time; graph capture records launches and replays them.
#pragma once
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#include <algorithm>

#define __device__
#define __host__
#define __global__
#define __shared__ static
#define __launch_bounds__(x)
#define __forceinline__ inline

struct double4 { double x, y, z, w; };
struct float4 { float x, y, z, w; };
struct uint4 { unsigned x, y, z, w; };
struct int4 { int x, y, z, w; };
inline double4 make_double4(double a, double b, double c, double d) { return {a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }

struct EmuDim { unsigned x = 0, y = 0, z = 0; };
extern thread_local EmuDim threadIdx, blockIdx;
extern EmuDim blockDim, gridDim;

unsigned long long __ballot(int pred);
double __shfl_xor(double v, int mask, int width = 64);
double __shfl(double v, int src, int width = 64);
int __shfl(int v, int src, int width = 64);
void __syncthreads();
void emu_wave_barrier();
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline int __ffsll(long long m) { return __builtin_ffsll(m); }
#define __builtin_amdgcn_fence(a, b) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_wave_barrier() emu_wave_barrier()

typedef int hipError_t;
#define hipSuccess 0
typedef struct EmuStream* hipStream_t;
typedef struct EmuEvent* hipEvent_t;
typedef struct EmuGraph* hipGraph_t;
typedef struct EmuGraph* hipGraphExec_t;
enum { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum { hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1, hipDeviceAttributeMultiprocessorCount = 1 };
const char* hipGetErrorString(hipError_t);
hipError_t hipMalloc(void** p, size_t n);
template <class T> hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, int kind);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int kind, hipStream_t);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, int);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipGetDevice(int* d);
hipError_t hipDeviceGetAttribute(int* v, int, int);
template <class K> hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) { *n = 1; return 0; }
hipError_t hipGetLastError();
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, int);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t g);
hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t s);

namespace emu {
void run_grid(int grid, int block, const std::function<void()>& body);
bool capturing();
void record(std::function<void()> f);
template <class K, class... A>
void launch(K k, int grid, int block, int, hipStream_t, A... args)
{
    auto f = [=]() { run_grid(grid, block, [=]() { k(args...); }); };
    if (capturing()) record(f);
    else f();
}
}  // namespace emu
using std::max;
using std::min;
inline int __builtin_amdgcn_readfirstlane(int v) { return __shfl(v, 0, 64); }
inline long long __double_as_longlong(double d) { long long b; memcpy(&b, &d, 8); return b; }
inline double __longlong_as_double(long long b) { double d; memcpy(&d, &b, 8); return d; }
