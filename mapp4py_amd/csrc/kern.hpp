// kern.hpp -- device helpers shared by the HIP translation units: launch
// geometry (wave64, 8 XCDs), deterministic block reductions, cell helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mapp {
namespace {

constexpr int kBlock = 256;
constexpr int kXcd = 8;
constexpr int kPartW = 8;  // doubles per partial-sum row

// number of workgroups, rounded up so the XCD remap below is a bijection
inline int grid_for(int n) { int b = (n + kBlock - 1) / kBlock; return ((b + kXcd - 1) / kXcd) * kXcd; }

// Workgroups are dealt to the 8 XCDs round-robin; give XCD x the contiguous
// block range [x*q, (x+1)*q) so neighbor gathers stay in that XCD's L2.
__device__ inline int xcd_block()
{
    const int nb = gridDim.x;
    const int q = nb / kXcd;
    return (blockIdx.x % kXcd) * q + blockIdx.x / kXcd;
}

// K FP64 sums over a workgroup of NT threads, written as one row of kPartW
template <int K, int NT = kBlock>
__device__ inline void block_sum_store(double (&v)[K], double* out)
{
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, 64);
    __shared__ double sh[NT / 64][K];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < K; ++k) sh[w][k] = v[k];
    __syncthreads();
    // the whole row is written (zeros past K) so any column count can be reduced
    if (threadIdx.x < kPartW) {
        double s = 0.0;
        if ((int)threadIdx.x < K)
            for (int q = 0; q < NT / 64; ++q) s += sh[q][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Fixed-order reduction of rows of partial sums: out[k] = sum_r part[r*W + k].
// Block b sums rows [b*chunk, (b+1)*chunk) and writes row b of out, so a
// launch over G blocks followed by one over a single block is a fixed-order
// (deterministic) two-level reduction.
__global__ __launch_bounds__(kBlock) void k_reduce_rows(int nrows, int ncols, const double* __restrict__ part,
                                                        double* __restrict__ out, int chunk)
{
    __shared__ double sh[kBlock][kPartW];
    double acc[kPartW];
    for (int k = 0; k < kPartW; ++k) acc[k] = 0.0;
    const int r0 = blockIdx.x * chunk;
    const int r1 = min(nrows, r0 + chunk);
    for (int r = r0 + threadIdx.x; r < r1; r += kBlock)
        for (int k = 0; k < ncols; ++k) acc[k] += part[(size_t)r * kPartW + k];
    out += (size_t)blockIdx.x * (gridDim.x > 1 ? kPartW : 0);
    for (int k = 0; k < kPartW; ++k) sh[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < kPartW; ++k) sh[threadIdx.x][k] += sh[threadIdx.x + s][k];
        __syncthreads();
    }
    if ((int)threadIdx.x < ncols) out[threadIdx.x] = sh[0][threadIdx.x];
}



__device__ inline int cell_coord(double s, int nc)
{
    int c = (int)(s * (double)nc);
    return c < nc - 1 ? c : nc - 1;
}

__device__ inline int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// r = sqrt(rsq) to ~1e-14 relative: FP32 root and one FP64 Newton step
__device__ inline double root64(double rsq)
{
    const float rf = __builtin_sqrtf((float)rsq);
    const double r0 = (double)rf;
    const double e = fma(-r0, r0, rsq);
    return fma(e, (double)(0.5f / rf), r0);
}

inline uint32_t morton3(uint32_t a, uint32_t b, uint32_t c)
{
    auto spread = [](uint32_t v) {
        v &= 0x3ff;
        v = (v | (v << 16)) & 0x030000FF;
        v = (v | (v << 8)) & 0x0300F00F;
        v = (v | (v << 4)) & 0x030C30C3;
        v = (v | (v << 2)) & 0x09249249;
        return v;
    };
    return spread(a) | (spread(b) << 1) | (spread(c) << 2);
}

}  // namespace
}  // namespace mapp
