// emu.cpp -- runtime of the CPU SIMT emulation (test infrastructure).
#include "hip/hip_runtime.h"
#include <barrier>
#include <memory>
#include <thread>

thread_local EmuDim threadIdx, blockIdx;
EmuDim blockDim, gridDim;

namespace {
struct Block {
    int nthreads;
    std::unique_ptr<std::barrier<>> all;
    std::vector<std::unique_ptr<std::barrier<>>> wave;
    std::vector<double> xd;            // per-thread exchange slots
    std::vector<long long> xi;
};
Block* g_block = nullptr;
bool g_capture = false;
std::vector<std::function<void()>>* g_rec = nullptr;
int nwave_lanes(int w) { int n = g_block->nthreads - 64 * w; return n < 64 ? n : 64; }
}  // namespace

struct EmuGraph { std::vector<std::function<void()>> ops; };

void emu_wave_barrier() { g_block->wave[threadIdx.x / 64]->arrive_and_wait(); }
void __syncthreads() { g_block->all->arrive_and_wait(); }

unsigned long long __ballot(int pred)
{
    const int t = threadIdx.x, w = t / 64;
    g_block->xi[t] = pred ? 1 : 0;
    emu_wave_barrier();
    unsigned long long m = 0;
    for (int l = 0; l < nwave_lanes(w); ++l)
        if (g_block->xi[64 * w + l]) m |= 1ull << l;
    emu_wave_barrier();
    return m;
}

double __shfl_xor(double v, int mask, int)
{
    const int t = threadIdx.x, w = t / 64, l = t % 64;
    g_block->xd[t] = v;
    emu_wave_barrier();
    double r = g_block->xd[64 * w + (l ^ mask)];
    emu_wave_barrier();
    return r;
}

double __shfl(double v, int src, int)
{
    const int t = threadIdx.x, w = t / 64;
    g_block->xd[t] = v;
    emu_wave_barrier();
    double r = g_block->xd[64 * w + src];
    emu_wave_barrier();
    return r;
}

int __shfl(int v, int src, int)
{
    const int t = threadIdx.x, w = t / 64;
    g_block->xi[t] = v;
    emu_wave_barrier();
    int r = (int)g_block->xi[64 * w + src];
    emu_wave_barrier();
    return r;
}

namespace emu {
void run_grid(int grid, int block, const std::function<void()>& body)
{
    gridDim.x = grid;
    blockDim.x = block;
    for (int b = 0; b < grid; ++b) {
        Block blk;
        blk.nthreads = block;
        blk.all = std::make_unique<std::barrier<>>(block);
        for (int w = 0; w * 64 < block; ++w)
            blk.wave.push_back(std::make_unique<std::barrier<>>(std::min(64, block - 64 * w)));
        blk.xd.assign(block, 0.0);
        blk.xi.assign(block, 0);
        g_block = &blk;
        std::vector<std::thread> th;
        for (int t = 0; t < block; ++t)
            th.emplace_back([&, t, b]() {
                threadIdx.x = t;
                blockIdx.x = b;
                body();
            });
        for (auto& x : th) x.join();
        g_block = nullptr;
    }
}
bool capturing() { return g_capture; }
void record(std::function<void()> f) { g_rec->push_back(std::move(f)); }
}  // namespace emu

const char* hipGetErrorString(hipError_t) { return "emulated error"; }
hipError_t hipMalloc(void** p, size_t n) { *p = calloc(1, n ? n : 1); return 0; }
hipError_t hipFree(void* p) { free(p); return 0; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, int) { memmove(d, s, n); return 0; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int, hipStream_t) { memmove(d, s, n); return 0; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, int) { *s = (hipStream_t)0x1; return 0; }
hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
hipError_t hipStreamDestroy(hipStream_t) { return 0; }
hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
hipError_t hipDeviceGetAttribute(int* v, int, int) { *v = 2; return 0; }
hipError_t hipGetLastError() { return 0; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = (hipEvent_t)0x1; return 0; }
hipError_t hipEventDestroy(hipEvent_t) { return 0; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return 0; }
hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 1.0f; return 0; }
hipError_t hipStreamBeginCapture(hipStream_t, int) { g_capture = true; g_rec = new std::vector<std::function<void()>>; return 0; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { g_capture = false; auto* G = new EmuGraph; G->ops = std::move(*g_rec); delete g_rec; g_rec = nullptr; *g = G; return 0; }
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) { *e = new EmuGraph(*g); return 0; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return 0; }
hipError_t hipGraphExecDestroy(hipGraphExec_t g) { delete g; return 0; }
hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t) { for (auto& f : g->ops) f(); return 0; }
