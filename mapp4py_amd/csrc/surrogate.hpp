// surrogate.hpp -- learned replacement of the DMD smeared functions.
//
// A multilayer perceptron per ordered element pair (a, b) maps
// (r, alpha_i, alpha_j) to the smeared r*phi_ab / r*rho_ab quantities the
// quadrature gives: (phi, rho, dphi/dr, drho/dr, dphi/dalpha, drho/dalpha)
// with alpha = sqrt(alpha_i^2 + alpha_j^2). Inference runs as batched BF16
// MFMA tiles over the staged neighbor slots; the DMD passes then read the
// per-slot results through surrogate_eval. Pairs outside the training box
// keep the quadrature.
#pragma once
#include <hip/hip_runtime.h>

namespace mapp {

constexpr int kSurrHidden = 64;
constexpr int kSurrOut = 6;

// Per-slot network outputs of the current configuration. Slot = the
// neighbor-list position k * ldl + component. in_ holds net(e_j, e_i)
// (phi_ij and the density at i), out_ holds net(e_i, e_j) (density at j).
// ok[slot] == 0 marks pairs left to the quadrature.
struct SurrogateArgs {
    const float* in_;
    const float* out_;
    const unsigned char* ok;
};

// rp / dr / da in the form of the quadrature: g, (1/r) dg/dr, (1/a) dg/da
template <bool DERIV>
__device__ inline bool surrogate_read(const SurrogateArgs& S, size_t slot, double r, double a, double (&rp)[3],
                                      double (&dr)[3], double (&da)[3])
{
    if (!S.ok[slot]) return false;
    const float* u = S.in_ + slot * kSurrOut;
    rp[0] = u[0];
    rp[2] = u[1];
    if (DERIV) {
        const float* v = S.out_ + slot * kSurrOut;
        rp[1] = v[1];
        const double ri = 1.0 / r, ai = 1.0 / a;
        dr[0] = u[2] * ri;
        dr[2] = u[3] * ri;
        dr[1] = v[3] * ri;
        da[0] = u[4] * ai;
        da[2] = u[5] * ai;
        da[1] = v[5] * ai;
    }
    return true;
}

}  // namespace mapp
