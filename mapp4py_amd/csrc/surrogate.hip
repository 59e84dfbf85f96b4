// surrogate.hip -- training and inference of the DMD smeared-function MLPs
// on BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation).
//
// Tile layout of v_mfma_f32_16x16x16_bf16 on a wave of 64 lanes (l):
//   A (16 x 16, M x K): lane l holds A[l % 16][4 (l / 16) + 0..3]
//   B (16 x 16, K x N): lane l holds B[4 (l / 16) + 0..3][l % 16]
//   D (16 x 16, M x N): lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
// Layers are computed transposed (weights as A, activations as B with the
// samples along N), so a layer's D tile is directly the next layer's B
// fragment: activations never leave registers in inference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "dmd.hpp"
#include "kern.hpp"
#include "surrogate.hpp"

namespace mapp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float f)
{
    union {
        float f;
        uint32_t u;
    } v{f};
    const uint32_t u = v.u + 0x7fffu + ((v.u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
    union {
        uint32_t u;
        float f;
    } v{(uint32_t)h << 16};
    return v.f;
}

// tanh(x) = 1 - 2 / (exp(2x) + 1); saturates cleanly at +-1
__device__ inline float tanh_f(float x) { return 1.0f - __fdividef(2.0f, __expf(2.0f * x) + 1.0f); }

__device__ inline floatx4 mfma(bf16x4 a, bf16x4 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// C tile (m0, n0) += A x B over K = 16 * ks, operands read from LDS with
// element (m, k) of A at A[m * sam + k * sak] and (k, n) of B at B[k * sbk + n * sbn].
__device__ inline floatx4 mma_lds(const uint16_t* A, int sam, int sak, const uint16_t* B, int sbk, int sbn, int m0,
                                  int n0, int ks, floatx4 c, int lane)
{
    const int m = m0 + (lane & 15), n = n0 + (lane & 15), kq = 4 * (lane >> 4);
    for (int s = 0; s < ks; ++s) {
        const int k = 16 * s + kq;
        bf16x4 a, b;
        for (int i = 0; i < 4; ++i) {
            a[i] = (short)A[m * sam + (k + i) * sak];
            b[i] = (short)B[(k + i) * sbk + n * sbn];
        }
        c = mfma(a, b, c);
    }
    return c;
}

__device__ inline float lane_group_sum(float v)  // sum over the 16 lanes that share lane / 16
{
    for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct TrainArgs {
    const float* params;  // [nets][kSurrParams]
    const float* x;       // [nets][ns][4] normalized inputs (4th unused)
    const float* t;       // [nets][ns][6] normalized targets
    int ns, off, batch, net0;
    float* part;          // [nets][nblocks][kSurrParams]
    float* lpart;         // [nets][nblocks]
};

// One wave = one workgroup = 64 samples of one network: forward, MSE loss
// gradient and backward; writes this block's partial parameter gradient.
__global__ __launch_bounds__(64) void k_surr_train(TrainArgs A)
{
    __shared__ uint16_t sW2[64 * 64], sW3[16 * 64], sH1[64 * 64], sH2[64 * 64], sD[64 * 64];
    __shared__ float sX[64][4], sT[64][6];
    const int lane = threadIdx.x, blk = blockIdx.x, nb = gridDim.x;
    const int lnet = blockIdx.y, net = A.net0 + lnet;
    const float* P = A.params + (size_t)net * kSurrParams;
    for (int q = lane; q < 64 * 64; q += 64) sW2[q] = f2bf(P[kSurrW2 + q]);
    for (int q = lane; q < 16 * 64; q += 64) sW3[q] = q < kSurrOut * 64 ? f2bf(P[kSurrW3 + q]) : (uint16_t)0;
    const int sidx = (A.off + blk * 64 + lane) % A.ns;
    const float* xs = A.x + ((size_t)lnet * A.ns + sidx) * 4;
    const float* ts = A.t + ((size_t)lnet * A.ns + sidx) * 6;
    const float xv[4] = {xs[0], xs[1], xs[2], 1.0f};
    for (int c = 0; c < 4; ++c) sX[lane][c] = xv[c];
    for (int o = 0; o < kSurrOut; ++o) sT[lane][o] = ts[o];
    // layer 1 in FP32, one sample per lane: H1^T[f][s]
    for (int f = 0; f < 64; ++f) {
        const float* w = P + kSurrW1 + 4 * f;
        const float z = w[0] * xv[0] + w[1] * xv[1] + w[2] * xv[2] + w[3];
        sH1[f * 64 + lane] = f2bf(tanh_f(z));
    }
    __syncthreads();
    const int q4 = 4 * (lane >> 4), col = lane & 15;
    // layer 2: Z2^T = W2 H1^T + b2
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt) {
            floatx4 c;
            for (int i = 0; i < 4; ++i) c[i] = P[kSurrB2 + 16 * mt + q4 + i];
            c = mma_lds(sW2, 64, 1, sH1, 64, 1, 16 * mt, 16 * nt, 4, c, lane);
            for (int i = 0; i < 4; ++i) sH2[(16 * mt + q4 + i) * 64 + 16 * nt + col] = f2bf(tanh_f(c[i]));
        }
    __syncthreads();
    // layer 3 and the loss gradient dY = 2 (Y - T) / (batch * 6)
    const float gs = 2.0f / (float)(A.batch * kSurrOut);
    float lsum = 0.0f, dyr[4][4];
    for (int nt = 0; nt < 4; ++nt) {
        floatx4 c;
        for (int i = 0; i < 4; ++i) c[i] = q4 + i < kSurrOut ? P[kSurrB3 + q4 + i] : 0.0f;
        c = mma_lds(sW3, 64, 1, sH2, 64, 1, 0, 16 * nt, 4, c, lane);
        for (int i = 0; i < 4; ++i) {
            const int o = q4 + i, s = 16 * nt + col;
            float d = 0.0f;
            if (o < kSurrOut) {
                const float e = c[i] - sT[s][o];
                lsum += e * e;
                d = gs * e;
            }
            dyr[nt][i] = d;
            sD[o * 64 + s] = f2bf(d);
        }
    }
    __syncthreads();
    float* G = A.part + ((size_t)lnet * nb + blk) * kSurrParams;
    // dW3[o][f] = sum_s dY[o][s] H2[f][s]
    for (int nt = 0; nt < 4; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        c = mma_lds(sD, 64, 1, sH2, 1, 64, 0, 16 * nt, 4, c, lane);
        for (int i = 0; i < 4; ++i)
            if (q4 + i < kSurrOut) G[kSurrW3 + (q4 + i) * 64 + 16 * nt + col] = c[i];
    }
    for (int i = 0; i < 4; ++i) {  // db3[o] = sum_s dY[o][s]
        const float v = lane_group_sum(dyr[0][i] + dyr[1][i] + dyr[2][i] + dyr[3][i]);
        if (col == 0 && q4 + i < kSurrOut) G[kSurrB3 + q4 + i] = v;
    }
    // dZ2^T[f][s] = (sum_o W3[o][f] dY[o][s]) (1 - h2^2)
    float dz2[4][4][4];
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt) {
            floatx4 c = {0.f, 0.f, 0.f, 0.f};
            c = mma_lds(sW3, 1, 64, sD, 64, 1, 16 * mt, 16 * nt, 1, c, lane);
            for (int i = 0; i < 4; ++i) {
                const float h = bf2f(sH2[(16 * mt + q4 + i) * 64 + 16 * nt + col]);
                dz2[mt][nt][i] = c[i] * (1.0f - h * h);
            }
        }
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt)
            for (int i = 0; i < 4; ++i) sD[(16 * mt + q4 + i) * 64 + 16 * nt + col] = f2bf(dz2[mt][nt][i]);
    for (int mt = 0; mt < 4; ++mt)  // db2[o] = sum_s dZ2[o][s]
        for (int i = 0; i < 4; ++i) {
            const float v = lane_group_sum(dz2[mt][0][i] + dz2[mt][1][i] + dz2[mt][2][i] + dz2[mt][3][i]);
            if (col == 0) G[kSurrB2 + 16 * mt + q4 + i] = v;
        }
    __syncthreads();
    // dW2[o][f] = sum_s dZ2[o][s] H1[f][s]
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt) {
            floatx4 c = {0.f, 0.f, 0.f, 0.f};
            c = mma_lds(sD, 64, 1, sH1, 1, 64, 16 * mt, 16 * nt, 4, c, lane);
            for (int i = 0; i < 4; ++i) G[kSurrW2 + (16 * mt + q4 + i) * 64 + 16 * nt + col] = c[i];
        }
    // dZ1^T[f][s] = (sum_o W2[o][f] dZ2[o][s]) (1 - h1^2); dW1[f][c] = sum_s dZ1[f][s] X[s][c]
    float g1[4][4][4];
    for (int mt = 0; mt < 4; ++mt)
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 4; ++c) g1[mt][i][c] = 0.0f;
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt) {
            floatx4 c = {0.f, 0.f, 0.f, 0.f};
            c = mma_lds(sW2, 1, 64, sD, 64, 1, 16 * mt, 16 * nt, 4, c, lane);
            const int s = 16 * nt + col;
            for (int i = 0; i < 4; ++i) {
                const float h = bf2f(sH1[(16 * mt + q4 + i) * 64 + s]);
                const float dz = c[i] * (1.0f - h * h);
                for (int cc = 0; cc < 4; ++cc) g1[mt][i][cc] += dz * sX[s][cc];
            }
        }
    for (int mt = 0; mt < 4; ++mt)
        for (int i = 0; i < 4; ++i)
            for (int cc = 0; cc < 4; ++cc) {
                const float v = lane_group_sum(g1[mt][i][cc]);
                if (col == 0) G[kSurrW1 + (16 * mt + q4 + i) * 4 + cc] = v;
            }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) A.lpart[(size_t)lnet * nb + blk] = lsum;
}

// grad[net][p] = sum over blocks (fixed order); loss[net] likewise
__global__ __launch_bounds__(256) void k_surr_reduce(int nets, int nb, const float* __restrict__ part,
                                                     const float* __restrict__ lpart, float* __restrict__ grad,
                                                     float* __restrict__ loss)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nets * kSurrParams) {
        const int net = q / kSurrParams, p = q % kSurrParams;
        float s = 0.0f;
        for (int b = 0; b < nb; ++b) s += part[((size_t)net * nb + b) * kSurrParams + p];
        grad[q] = s;
    }
    if (q < nets) {
        float s = 0.0f;
        for (int b = 0; b < nb; ++b) s += lpart[(size_t)q * nb + b];
        loss[q] = s;
    }
}

__global__ __launch_bounds__(256) void k_surr_adam(int n, float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, float lr_t)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = 0.9f * m[q] + 0.1f * gq;
    const float vq = 0.999f * v[q] + 0.001f * gq * gq;
    m[q] = mq;
    v[q] = vq;
    w[q] -= lr_t * mq / (sqrtf(vq) + 1e-8f);
}

// hidden unit carried by element e (0..7) of k-step ks for lane group g (0..3)
// of the 16x16x32 B operand: two output rows of consecutive 16-row tiles
__host__ __device__ inline int hid_of(int ks, int g, int e) { return 16 * (2 * ks + (e >> 2)) + 4 * g + (e & 3); }

__device__ inline bf16x8 load8(const uint16_t* p)  // 16-byte aligned
{
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    bf16x8 r;
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
    for (int k = 0; k < 4; ++k) {
        r[2 * k] = __builtin_bit_cast(__bf16, (uint16_t)(w[k] & 0xffff));
        r[2 * k + 1] = __builtin_bit_cast(__bf16, (uint16_t)(w[k] >> 16));
    }
    return r;
}

__device__ inline floatx4 mfma32(bf16x8 a, bf16x8 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hardware round-to-nearest-even conversions (v_cvt_pk_bf16_f32)
__device__ inline void split8(const float (&h)[8], bf16x8& hi, bf16x8& lo)
{
    for (int e = 0; e < 8; ++e) {
        const __bf16 b = (__bf16)h[e];
        hi[e] = b;
        lo[e] = (__bf16)(h[e] - (float)b);
    }
}

// Inference on v_mfma_f32_16x16x32_bf16 (lane l of the B operand holds
// K rows 8 (l / 16) + 0..7 of column l % 16; D as for 16x16x16). Each wave
// takes groups of 64 requests (four tiles of 16 samples); weight fragments
// stay in registers for the whole launch. Every product is
// hi*hi + hi*lo + lo*hi of the split BF16 operands.
__global__ __launch_bounds__(256) void k_surr_infer(SurrNetDev N, int nreq, const SurrRequest* __restrict__ req,
                                                    float* __restrict__ y_out)
{
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int g = lane >> 4, col = lane & 15, q4 = 4 * g;
    __shared__ float4 sw1[kSurrHidden];  // first-layer weights and bias per hidden unit
    for (int f = threadIdx.x; f < kSurrHidden; f += blockDim.x)
        sw1[f] = make_float4(N.w1[4 * f], N.w1[4 * f + 1], N.w1[4 * f + 2], N.w1[4 * f + 3]);
    __syncthreads();
    bf16x8 w2h[4][2], w2l[4][2], w3h[2], w3l[2];
    for (int mt = 0; mt < 4; ++mt)
        for (int ks = 0; ks < 2; ++ks) {
            const uint16_t* p = N.w2p + (((size_t)(mt * 2 + ks) * 64 + lane) * 2) * 8;
            w2h[mt][ks] = load8(p);
            w2l[mt][ks] = load8(p + 8);
        }
    for (int ks = 0; ks < 2; ++ks) {
        const uint16_t* p = N.w3p + (((size_t)ks * 64 + lane) * 2) * 8;
        w3h[ks] = load8(p);
        w3l[ks] = load8(p + 8);
    }
    float b2[4][4], b3[4], sc[4], sh[4];
    for (int mt = 0; mt < 4; ++mt)
        for (int i = 0; i < 4; ++i) b2[mt][i] = N.b2[16 * mt + q4 + i];
    for (int i = 0; i < 4; ++i) {
        const int o = q4 + i;
        b3[i] = o < kSurrOut ? N.b3[o] : 0.0f;
        sc[i] = o < kSurrOut ? N.scale[o] : 0.0f;
        sh[i] = o < kSurrOut ? N.shift[o] : 0.0f;
    }
    for (int gr = wave; gr * 64 < nreq; gr += nwaves) {
        for (int nt = 0; nt < 4; ++nt) {
            const int r = gr * 64 + 16 * nt + col;
            const SurrRequest q = req[r < nreq ? r : nreq - 1];
            bf16x8 bh[2], bl[2];
            for (int ks = 0; ks < 2; ++ks) {
                float h[8];
                for (int e = 0; e < 8; ++e) {
                    const float4 w = sw1[hid_of(ks, g, e)];
                    h[e] = tanh_f(w.x * q.x0 + w.y * q.x1 + w.z * q.x2 + w.w);
                }
                split8(h, bh[ks], bl[ks]);
            }
            floatx4 z[4];
            for (int mt = 0; mt < 4; ++mt) {
                for (int i = 0; i < 4; ++i) z[mt][i] = b2[mt][i];
                for (int ks = 0; ks < 2; ++ks) {
                    z[mt] = mfma32(w2h[mt][ks], bh[ks], z[mt]);
                    z[mt] = mfma32(w2h[mt][ks], bl[ks], z[mt]);
                    z[mt] = mfma32(w2l[mt][ks], bh[ks], z[mt]);
                }
            }
            bf16x8 gh[2], gl[2];
            for (int ks = 0; ks < 2; ++ks) {
                float h[8];
                for (int e = 0; e < 8; ++e) h[e] = tanh_f(z[2 * ks + (e >> 2)][e & 3]);
                split8(h, gh[ks], gl[ks]);
            }
            floatx4 y;
            for (int i = 0; i < 4; ++i) y[i] = b3[i];
            for (int ks = 0; ks < 2; ++ks) {
                y = mfma32(w3h[ks], gh[ks], y);
                y = mfma32(w3h[ks], gl[ks], y);
                y = mfma32(w3l[ks], gh[ks], y);
            }
            if (r < nreq && q4 < kSurrOut)
                for (int i = 0; i < 4; ++i)
                    if (q4 + i < kSurrOut) y_out[(size_t)r * kSurrOut + q4 + i] = y[i] * sc[i] + sh[i];
        }
    }
}

}  // namespace

Surrogate::Surrogate(int ne, SurrBox box, int seed) : ne_(ne), nt_(ne * ne), box_(box)
{
    MAPP_REQUIRE(ne >= 1 && ne <= kMaxElems, ERR_ARG, "element count out of range");
    MAPP_REQUIRE(box.r_hi > box.r_lo && box.a_hi > box.a_lo && box.a_lo > 0.0f, ERR_ARG, "empty training box");
    params_.assign((size_t)nt_ * kSurrParams, 0.0f);
    scale_.assign((size_t)nt_ * kSurrOut, 1.0f);
    shift_.assign((size_t)nt_ * kSurrOut, 0.0f);
    // Glorot-uniform initial weights, zero biases
    std::mt19937 gen((unsigned)seed);
    auto fill = [&](float* w, int n, int fan_in, int fan_out) {
        std::uniform_real_distribution<float> u(-1.0f, 1.0f);
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        for (int k = 0; k < n; ++k) w[k] = a * u(gen);
    };
    for (int t = 0; t < nt_; ++t) {
        float* p = params_.data() + (size_t)t * kSurrParams;
        for (int f = 0; f < kSurrHidden; ++f) fill(p + kSurrW1 + 4 * f, 3, kSurrIn, kSurrHidden);
        fill(p + kSurrW2, kSurrHidden * kSurrHidden, kSurrHidden, kSurrHidden);
        fill(p + kSurrW3, kSurrOut * kSurrHidden, kSurrHidden, kSurrOut);
    }
    dparams_.reserve(params_.size());
    MAPP_HIP(hipMemcpy(dparams_.get(), params_.data(), params_.size() * sizeof(float), hipMemcpyHostToDevice));
}

void Surrogate::get_params(int net, float* p) const
{
    MAPP_REQUIRE(net >= 0 && net < nt_, ERR_ARG, "network index out of range");
    MAPP_HIP(hipMemcpy(p, dparams_.get() + (size_t)net * kSurrParams, kSurrParams * sizeof(float),
                       hipMemcpyDeviceToHost));
}

void Surrogate::set_params(int net, const float* p)
{
    MAPP_REQUIRE(net >= 0 && net < nt_, ERR_ARG, "network index out of range");
    MAPP_HIP(hipMemcpy(dparams_.get() + (size_t)net * kSurrParams, p, kSurrParams * sizeof(float),
                       hipMemcpyHostToDevice));
    dirty_ = true;
}

void Surrogate::get_norm(int net, float* scale, float* shift) const
{
    MAPP_REQUIRE(net >= 0 && net < nt_, ERR_ARG, "network index out of range");
    std::memcpy(scale, scale_.data() + (size_t)net * kSurrOut, kSurrOut * sizeof(float));
    std::memcpy(shift, shift_.data() + (size_t)net * kSurrOut, kSurrOut * sizeof(float));
}

void Surrogate::set_norm(int net, const float* scale, const float* shift)
{
    MAPP_REQUIRE(net >= 0 && net < nt_, ERR_ARG, "network index out of range");
    std::memcpy(scale_.data() + (size_t)net * kSurrOut, scale, kSurrOut * sizeof(float));
    std::memcpy(shift_.data() + (size_t)net * kSurrOut, shift, kSurrOut * sizeof(float));
    dirty_ = true;
}

// BF16 high/low copies of the weights in the layout the inference kernel loads
void Surrogate::refresh_inference()
{
    if (!dirty_) return;
    params_.resize((size_t)nt_ * kSurrParams);
    MAPP_HIP(hipMemcpy(params_.data(), dparams_.get(), params_.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> w1((size_t)nt_ * 64 * 4), b2((size_t)nt_ * 64), b3((size_t)nt_ * 8, 0.0f);
    const size_t f2 = (size_t)4 * 2 * 64 * 2 * 8, f3 = (size_t)2 * 64 * 2 * 8;  // per network
    std::vector<uint16_t> w2p((size_t)nt_ * f2), w3p((size_t)nt_ * f3, 0);
    auto put = [](uint16_t* dst, float w) {
        const uint16_t hi = f2bf(w);
        dst[0] = hi;
        dst[8] = f2bf(w - bf2f(hi));
    };
    for (int t = 0; t < nt_; ++t) {
        const float* p = params_.data() + (size_t)t * kSurrParams;
        std::memcpy(w1.data() + (size_t)t * 256, p + kSurrW1, 256 * sizeof(float));
        std::memcpy(b2.data() + (size_t)t * 64, p + kSurrB2, 64 * sizeof(float));
        std::memcpy(b3.data() + (size_t)t * 8, p + kSurrB3, kSurrOut * sizeof(float));
        for (int mt = 0; mt < 4; ++mt)
            for (int ks = 0; ks < 2; ++ks)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int o = 16 * mt + (l & 15), f = hid_of(ks, l >> 4, e);
                        put(w2p.data() + (size_t)t * f2 + (((size_t)(mt * 2 + ks) * 64 + l) * 2) * 8 + e,
                            p[kSurrW2 + o * 64 + f]);
                    }
        for (int ks = 0; ks < 2; ++ks)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int o = l & 15, f = hid_of(ks, l >> 4, e);
                    if (o < kSurrOut)
                        put(w3p.data() + (size_t)t * f3 + (((size_t)ks * 64 + l) * 2) * 8 + e, p[kSurrW3 + o * 64 + f]);
                }
    }
    dw1_.reserve(w1.size());
    db2_.reserve(b2.size());
    db3_.reserve(b3.size());
    dw2p_.reserve(w2p.size());
    dw3p_.reserve(w3p.size());
    MAPP_HIP(hipMemcpy(dw1_.get(), w1.data(), w1.size() * sizeof(float), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(db2_.get(), b2.data(), b2.size() * sizeof(float), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(db3_.get(), b3.data(), b3.size() * sizeof(float), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(dw2p_.get(), w2p.data(), w2p.size() * 2, hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(dw3p_.get(), w3p.data(), w3p.size() * 2, hipMemcpyHostToDevice));
    dirty_ = false;
}

SurrNetDev Surrogate::net_dev(int t) const
{
    SurrNetDev d{};
    d.w1 = dw1_.get() + (size_t)t * 256;
    d.b2 = db2_.get() + (size_t)t * 64;
    d.b3 = db3_.get() + (size_t)t * 8;
    d.w2p = dw2p_.get() + (size_t)t * 4 * 2 * 64 * 2 * 8;
    d.w3p = dw3p_.get() + (size_t)t * 2 * 64 * 2 * 8;
    for (int o = 0; o < kSurrOut; ++o) {
        d.scale[o] = scale_[(size_t)t * kSurrOut + o];
        d.shift[o] = shift_[(size_t)t * kSurrOut + o];
    }
    d.box = box_;
    return d;
}

void Surrogate::infer_dev(int net, int nreq, const SurrRequest* req, float* y, hipStream_t s)
{
    MAPP_REQUIRE(net >= 0 && net < nt_, ERR_ARG, "network index out of range");
    if (nreq <= 0) return;
    refresh_inference();
    int dev = 0, ncu = 0;
    MAPP_HIP(hipGetDevice(&dev));
    MAPP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int groups = (nreq + 63) / 64;
    const int blocks = std::max(1, std::min((groups + 3) / 4, ncu * 8));
    hipLaunchKernelGGL(k_surr_infer, dim3(blocks), dim3(256), 0, s, net_dev(net), nreq, req, y);
    MAPP_HIP(hipGetLastError());
}

void Surrogate::infer(int net, int n, const float* rin, float* out)
{
    MAPP_REQUIRE(n >= 0, ERR_ARG, "negative point count");
    if (n == 0) return;
    std::vector<SurrRequest> rq(n);
    const float sr = 2.0f / (box_.r_hi - box_.r_lo), sa = 2.0f / (box_.a_hi - box_.a_lo);
    for (int i = 0; i < n; ++i) {
        rq[i].x0 = (rin[3 * i] - box_.r_lo) * sr - 1.0f;
        rq[i].x1 = (rin[3 * i + 1] - box_.a_lo) * sa - 1.0f;
        rq[i].x2 = (rin[3 * i + 2] - box_.a_lo) * sa - 1.0f;
        rq[i].pad = 0.0f;
    }
    DevBuf<SurrRequest> dr;
    DevBuf<float> dout;
    dr.reserve(n);
    dout.reserve((size_t)n * kSurrOut);
    MAPP_HIP(hipMemcpy(dr.get(), rq.data(), n * sizeof(SurrRequest), hipMemcpyHostToDevice));
    infer_dev(net, n, dr.get(), dout.get(), nullptr);
    MAPP_HIP(hipDeviceSynchronize());
    MAPP_HIP(hipMemcpy(out, dout.get(), (size_t)n * kSurrOut * sizeof(float), hipMemcpyDeviceToHost));
}

double Surrogate::loss_grad(int net, int n, const float* x, const float* t, float* grad)
{
    MAPP_REQUIRE(net >= 0 && net < nt_, ERR_ARG, "network index out of range");
    MAPP_REQUIRE(n > 0 && n % 64 == 0, ERR_ARG, "the sample count must be a positive multiple of 64");
    std::vector<float> x4((size_t)n * 4);
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) x4[(size_t)i * 4 + c] = x[(size_t)i * 3 + c];
        x4[(size_t)i * 4 + 3] = 0.0f;
    }
    const int nb = n / 64;
    DevBuf<float> dx, dt, part, lpart, g, l;
    dx.reserve(x4.size());
    dt.reserve((size_t)n * kSurrOut);
    part.reserve((size_t)nb * kSurrParams);
    lpart.reserve(nb);
    g.reserve(kSurrParams);
    l.reserve(1);
    MAPP_HIP(hipMemcpy(dx.get(), x4.data(), x4.size() * sizeof(float), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(dt.get(), t, (size_t)n * kSurrOut * sizeof(float), hipMemcpyHostToDevice));
    TrainArgs A{dparams_.get(), dx.get(), dt.get(), n, 0, n, net, part.get(), lpart.get()};
    hipLaunchKernelGGL(k_surr_train, dim3(nb, 1), dim3(64), 0, nullptr, A);
    MAPP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_surr_reduce, dim3((kSurrParams + 255) / 256), dim3(256), 0, nullptr, 1, nb, part.get(),
                       lpart.get(), g.get(), l.get());
    MAPP_HIP(hipGetLastError());
    MAPP_HIP(hipDeviceSynchronize());
    MAPP_HIP(hipMemcpy(grad, g.get(), kSurrParams * sizeof(float), hipMemcpyDeviceToHost));
    float loss = 0.0f;
    MAPP_HIP(hipMemcpy(&loss, l.get(), sizeof(float), hipMemcpyDeviceToHost));
    return (double)loss / ((double)n * kSurrOut);
}

SurrTrainReport Surrogate::train(DMDSystem& d, int nsamples, int steps, int batch, double lr)
{
    MAPP_REQUIRE(d.args().ne == ne_, ERR_ARG, "the DMD system has a different element count");
    MAPP_REQUIRE(nsamples >= 64 && batch >= 64 && batch % 64 == 0 && steps >= 1, ERR_ARG,
                 "nsamples >= 64, steps >= 1 and batch a positive multiple of 64");
    const int ns = nsamples, nb = batch / 64;
    std::vector<float> hx((size_t)nt_ * ns * 4), ht((size_t)nt_ * ns * kSurrOut);
    {
        std::mt19937_64 gen(12345u + (unsigned)ns);
        std::uniform_real_distribution<double> ur(box_.r_lo, box_.r_hi), ua(box_.a_lo, box_.a_hi);
        std::vector<double> r(ns), a1(ns), a2(ns), aij(ns), out((size_t)ns * 9);
        DevBuf<double> dr, da, dout;
        dr.reserve(ns);
        da.reserve(ns);
        dout.reserve((size_t)ns * 9);
        const double sr = 2.0 / (box_.r_hi - box_.r_lo), sa = 2.0 / (box_.a_hi - box_.a_lo);
        for (int t = 0; t < nt_; ++t) {
            for (int i = 0; i < ns; ++i) {
                r[i] = ur(gen);
                a1[i] = ua(gen);
                a2[i] = ua(gen);
                aij[i] = std::sqrt(a1[i] * a1[i] + a2[i] * a2[i]);
            }
            MAPP_HIP(hipMemcpy(dr.get(), r.data(), ns * sizeof(double), hipMemcpyHostToDevice));
            MAPP_HIP(hipMemcpy(da.get(), aij.data(), ns * sizeof(double), hipMemcpyHostToDevice));
            d.smear_dev(t, ns, dr.get(), da.get(), dout.get());
            MAPP_HIP(hipMemcpy(out.data(), dout.get(), out.size() * sizeof(double), hipMemcpyDeviceToHost));
            // targets: g_phi, g_rho, dg_phi/dr, dg_rho/dr, dg_phi/dalpha, dg_rho/dalpha
            std::vector<double> y((size_t)ns * kSurrOut);
            for (int i = 0; i < ns; ++i) {
                const double* o = out.data() + (size_t)i * 9;
                double* yi = y.data() + (size_t)i * kSurrOut;
                yi[0] = o[0];
                yi[1] = o[1];
                yi[2] = o[3] * r[i];
                yi[3] = o[4] * r[i];
                yi[4] = o[6] * aij[i];
                yi[5] = o[7] * aij[i];
            }
            for (int c = 0; c < kSurrOut; ++c) {
                double m = 0.0, q = 0.0;
                for (int i = 0; i < ns; ++i) m += y[(size_t)i * kSurrOut + c];
                m /= ns;
                for (int i = 0; i < ns; ++i) {
                    const double e = y[(size_t)i * kSurrOut + c] - m;
                    q += e * e;
                }
                const double sd = std::sqrt(q / ns);
                shift_[(size_t)t * kSurrOut + c] = (float)m;
                scale_[(size_t)t * kSurrOut + c] = sd > 0.0 ? (float)sd : 1.0f;
            }
            for (int i = 0; i < ns; ++i) {
                float* xi = hx.data() + ((size_t)t * ns + i) * 4;
                xi[0] = (float)((r[i] - box_.r_lo) * sr - 1.0);
                xi[1] = (float)((a1[i] - box_.a_lo) * sa - 1.0);
                xi[2] = (float)((a2[i] - box_.a_lo) * sa - 1.0);
                xi[3] = 0.0f;
                for (int c = 0; c < kSurrOut; ++c)
                    ht[((size_t)t * ns + i) * kSurrOut + c] =
                        (float)((y[(size_t)i * kSurrOut + c] - shift_[(size_t)t * kSurrOut + c]) /
                                scale_[(size_t)t * kSurrOut + c]);
            }
        }
    }
    DevBuf<float> dx, dt;
    dx.reserve(hx.size());
    dt.reserve(ht.size());
    MAPP_HIP(hipMemcpy(dx.get(), hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(dt.get(), ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice));
    const size_t np = (size_t)nt_ * kSurrParams;
    dm_.reserve(np);
    dv_.reserve(np);
    MAPP_HIP(hipMemset(dm_.get(), 0, np * sizeof(float)));
    MAPP_HIP(hipMemset(dv_.get(), 0, np * sizeof(float)));
    dpart_.reserve((size_t)nt_ * nb * kSurrParams);
    dgrad_.reserve(np);
    dloss_.reserve((size_t)nt_ * (nb + 1));
    float* lpart = dloss_.get();
    float* lsum = dloss_.get() + (size_t)nt_ * nb;
    SurrTrainReport rep{0.0, 0.0, steps};
    hipStream_t s = nullptr;
    prof_.begin(K_SURR_TRAIN, s);
    double b1t = 1.0, b2t = 1.0;
    const double pi = 3.14159265358979323846;
    for (int st = 0; st < steps; ++st) {
        TrainArgs A{dparams_.get(), dx.get(), dt.get(), ns, (int)(((long long)st * batch) % ns), batch, 0,
                    dpart_.get(), lpart};
        hipLaunchKernelGGL(k_surr_train, dim3(nb, nt_), dim3(64), 0, s, A);
        hipLaunchKernelGGL(k_surr_reduce, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, nt_, nb,
                           dpart_.get(), lpart, dgrad_.get(), lsum);
        b1t *= 0.9;
        b2t *= 0.999;
        const double lr_s = lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(pi * st / std::max(1, steps - 1))));
        const float lr_t = (float)(lr_s * std::sqrt(1.0 - b2t) / (1.0 - b1t));
        hipLaunchKernelGGL(k_surr_adam, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, (int)np,
                           dparams_.get(), dgrad_.get(), dm_.get(), dv_.get(), lr_t);
        if (st == 0 || st == steps - 1) {
            std::vector<float> l(nt_);
            MAPP_HIP(hipMemcpy(l.data(), lsum, nt_ * sizeof(float), hipMemcpyDeviceToHost));
            double m = 0.0;
            for (float v : l) m += v;
            m /= (double)nt_ * batch * kSurrOut;
            (st == 0 ? rep.loss0 : rep.loss) = m;
        }
    }
    MAPP_HIP(hipGetLastError());
    prof_.end(s);
    MAPP_HIP(hipDeviceSynchronize());
    dirty_ = true;
    return rep;
}

}  // namespace mapp
