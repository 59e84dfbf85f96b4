// surrogate.hpp -- learned replacement of the DMD smeared functions.
//
// A multilayer perceptron per ordered element pair (a, b), 3 -> 64 -> 64 -> 6
// with tanh, maps (r, alpha_a, alpha_b) to the smeared quantities the
// quadrature gives for r*phi_ab and r*rho_ab: (g_phi, g_rho, dg_phi/dr,
// dg_rho/dr, dg_phi/dalpha, dg_rho/dalpha) with alpha = sqrt(alpha_a^2 +
// alpha_b^2). Training and inference run on BF16 MFMA tiles with FP32
// accumulation (inference splits every operand into a BF16 high and low part
// and sums three products, which carries about 16 mantissa bits). The DMD
// passes read per-slot network outputs through surrogate_read; pairs outside
// the training box keep the quadrature.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace mapp {

constexpr int kSurrIn = 3;
constexpr int kSurrHidden = 64;
constexpr int kSurrOut = 6;
constexpr int kMaxSurrNets = 16;  // ordered element pairs the surrogate covers (4 elements)
// flat parameter vector of one network
constexpr int kSurrW1 = 0;                             // [64][4]: weights of r, alpha_a, alpha_b, then bias
constexpr int kSurrW2 = kSurrW1 + kSurrHidden * 4;     // [64 out][64 in]
constexpr int kSurrB2 = kSurrW2 + kSurrHidden * kSurrHidden;
constexpr int kSurrW3 = kSurrB2 + kSurrHidden;         // [6 out][64 in]
constexpr int kSurrB3 = kSurrW3 + kSurrOut * kSurrHidden;
constexpr int kSurrParams = kSurrB3 + kSurrOut;        // 4806

// Training box: r in [r_lo, r_hi], each alpha in [a_lo, a_hi]; inputs are
// mapped linearly onto [-1, 1].
struct SurrBox {
    float r_lo, r_hi, a_lo, a_hi;
};

// Network outputs of the current configuration, one row of kSurrOut floats
// per request. Requests are grouped by network and, inside a network, by site
// in site order; offs[t * n + i] is the first request of network t for site
// i. A pass walks a site's pairs in the order the requests were made, so a
// cursor per network finds each pair's rows without any per-slot index.
struct SurrogateArgs {
    const float* y;
    const int* offs;
    int n;
    SurrBox box;
};

// pairs the networks cover (both alphas and r inside the training box)
__device__ inline bool surrogate_in_box(const SurrBox& b, double r, double ai, double aj)
{
    return r >= b.r_lo && r <= b.r_hi && ai >= b.a_lo && ai <= b.a_hi && aj >= b.a_lo && aj <= b.a_hi;
}

// Request rows of a pair: same-element pairs share one network call for both
// directions; a mixed pair (e_i, e_j) reads net(e_j, e_i) for phi and the
// density at i, net(e_i, e_j) for the density at j.
struct SurrCursor {
    int pos[kMaxSurrNets];
    __device__ void begin(const SurrogateArgs& S, int i, int nt)
    {
        for (int t = 0; t < nt; ++t) pos[t] = S.offs[(size_t)t * S.n + i];
    }
    __device__ bool next(const SurrogateArgs& S, int ne, int ei, int ej, double r, double ai, double aj, int& qin,
                         int& qout)
    {
        if (!surrogate_in_box(S.box, r, ai, aj)) return false;
        if (ei == ej) {
            qin = qout = pos[ei * ne + ei]++;
        } else {
            qin = pos[ej * ne + ei]++;
            qout = pos[ei * ne + ej]++;
        }
        return true;
    }
};

// rp / dr / da in the form of the quadrature: g, (1/r) dg/dr, (1/a) dg/da
template <bool DERIV>
__device__ inline void surrogate_values(const SurrogateArgs& S, int qin, int qout, double r, double a,
                                        double (&rp)[3], double (&dr)[3], double (&da)[3])
{
    const float* u = S.y + (size_t)qin * kSurrOut;
    rp[0] = u[0];
    rp[2] = u[1];
    if (DERIV) {
        const float* v = S.y + (size_t)qout * kSurrOut;
        rp[1] = v[1];
        const double ri = 1.0 / r, ai = 1.0 / a;
        dr[0] = u[2] * ri;
        dr[2] = u[3] * ri;
        dr[1] = v[3] * ri;
        da[0] = u[4] * ai;
        da[2] = u[5] * ai;
        da[1] = v[5] * ai;
    }
}

// One network as the inference kernel reads it.
struct SurrNetDev {
    const float* w1;         // [64][4] FP32
    const float* b2;         // [64]
    const float* b3;         // [6]
    // A fragments of W2 and W3 in lane order, [m-tile][k-step][lane][hi, lo][8]
    // (BF16 high part and the BF16 rounding of the remainder); the hidden
    // units of each k-step are permuted so that a layer's output tile is the
    // next layer's B fragment without any data movement
    const uint16_t* w2p;     // [4][2][64][2][8]
    const uint16_t* w3p;     // [1][2][64][2][8], output rows 6..15 zero
    float scale[kSurrOut], shift[kSurrOut];  // y = net * scale + shift
    SurrBox box;
};

// One inference request: normalized inputs (r, alpha_a, alpha_b).
struct SurrRequest {
    float x0, x1, x2, pad;
};

class DMDSystem;

struct SurrTrainReport {
    double loss0, loss;  // mean squared error of the normalized outputs, first and last step
    int steps;
};

// Host side of the per-element-pair networks of one DMD force field.
class Surrogate {
public:
    Surrogate(int ne, SurrBox box, int seed);
    int nnets() const { return nt_; }
    const SurrBox& box() const { return box_; }
    // nsamples quadrature samples per network (uniform in the box), then
    // `steps` Adam steps of `batch` samples per network (cosine learning rate
    // from lr to lr/20)
    SurrTrainReport train(DMDSystem& d, int nsamples, int steps, int batch, double lr);
    // loss and gradient of one network on given normalized inputs/targets
    // (the same BF16 MFMA kernel as training)
    double loss_grad(int net, int n, const float* x, const float* t, float* grad);
    void get_params(int net, float* p) const;
    void set_params(int net, const float* p);
    void get_norm(int net, float* scale, float* shift) const;
    void set_norm(int net, const float* scale, const float* shift);
    // denormalized outputs for raw (r, alpha_a, alpha_b) inputs
    void infer(int net, int n, const float* rin, float* out);
    // inference for device-resident requests of one network: y[r][6] for r < nreq
    void infer_dev(int net, int nreq, const SurrRequest* req, float* y, hipStream_t s);
    Profiler& profiler() { return prof_; }

private:
    void refresh_inference();
    SurrNetDev net_dev(int t) const;

    int ne_, nt_;
    SurrBox box_;
    std::vector<float> params_;  // host master copy [nt][kSurrParams]
    std::vector<float> scale_, shift_;
    DevBuf<float> dparams_, dm_, dv_, dpart_, dgrad_, dloss_;
    DevBuf<float> dw1_, db2_, db3_;
    DevBuf<uint16_t> dw2p_, dw3p_;
    bool dirty_ = true;
    Profiler prof_;
};

}  // namespace mapp
