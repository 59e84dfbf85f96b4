// box.hpp -- periodic box algebra shared by host and device code.
//
// The box matrix H is lower triangular with the box vectors as rows; B = H^-1.
// Conversions use the packed lower triangles h = (H00,H10,H20,H11,H21,H22)
// and b likewise, with the association and rounding of the reference
// (src/xmath.h:2636-2700, src/atoms.cpp:187-194). Every function that feeds the
// neighbor predicate is compiled with floating-point contraction disabled so
// that it rounds each product exactly as the reference's x86-64 build does.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mapp {

struct Box {
    double H[3][3];
    double B[3][3];
    double h[6], b[6];
    double depth_inv[3];
    double vol;
};

// lower-triangular inverse, depth and volume (src/xmath.h:1403-1485, 2399-2415)
inline Box make_box(const double* Hin)
{
#pragma clang fp contract(off)
    Box bx{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) bx.H[i][j] = j <= i ? Hin[i * 3 + j] : 0.0;
    const double(*H)[3] = bx.H;
    double(*B)[3] = bx.B;
    B[0][0] = 1.0 / H[0][0];
    B[1][1] = 1.0 / H[1][1];
    B[1][0] = -B[1][1] * (B[0][0] * H[1][0]);
    B[2][2] = 1.0 / H[2][2];
    B[2][1] = -B[2][2] * (B[1][1] * H[2][1]);
    B[2][0] = -B[2][2] * (B[0][0] * H[2][0] + B[1][0] * H[2][1]);
    bx.vol = H[0][0] * (H[1][1] * H[2][2]);
    bx.depth_inv[0] = std::sqrt(B[0][0] * B[0][0] + (B[1][0] * B[1][0] + B[2][0] * B[2][0]));
    bx.depth_inv[1] = std::sqrt(B[1][1] * B[1][1] + B[2][1] * B[2][1]);
    bx.depth_inv[2] = std::fabs(B[2][2]);
    const double hp[6] = {H[0][0], H[1][0], H[2][0], H[1][1], H[2][1], H[2][2]};
    const double bp[6] = {B[0][0], B[1][0], B[2][0], B[1][1], B[2][1], B[2][2]};
    for (int k = 0; k < 6; ++k) {
        bx.h[k] = hp[k];
        bx.b[k] = bp[k];
    }
    return bx;
}

// packed triangles as they travel to kernels
struct BoxDev {
    double h[6], b[6];
};

__host__ __device__ inline double wrap_unit(double s)
{
    while (s < 0.0) s += 1.0;
    while (s >= 1.0) s -= 1.0;
    return s;
}

// x -> s = x B, each component wrapped into [0,1)
__host__ __device__ inline void x_to_s(const double* b, double x0, double x1, double x2, double& s0, double& s1,
                                       double& s2)
{
#pragma clang fp contract(off)
    s0 = wrap_unit(b[0] * x0 + (b[1] * x1 + b[2] * x2));
    s1 = wrap_unit(b[3] * x1 + b[4] * x2);
    s2 = wrap_unit(b[5] * x2);
}

// s -> x = s H
__host__ __device__ inline void s_to_x(const double* h, double s0, double s1, double s2, double& x0, double& x1,
                                       double& x2)
{
#pragma clang fp contract(off)
    x0 = h[0] * s0 + (h[1] * s1 + h[2] * s2);
    x1 = h[3] * s1 + h[4] * s2;
    x2 = h[5] * s2;
}

// squared distance with the reference's association dx0^2 + (dx1^2 + dx2^2)
__host__ __device__ inline double rsq_exact(double a0, double a1, double a2, double b0, double b1, double b2)
{
#pragma clang fp contract(off)
    const double d0 = a0 - b0, d1 = a1 - b1, d2 = a2 - b2;
    return d0 * d0 + (d1 * d1 + d2 * d2);
}

// fractional coordinate of a periodic image n steps away, built by repeated
// +/-1 as the ghost layers of one rank are (src/exchange.h:833-866)
__host__ __device__ inline double shift_unit(double s, int n)
{
#pragma clang fp contract(off)
    for (; n > 0; --n) s += 1.0;
    for (; n < 0; ++n) s -= 1.0;
    return s;
}

}  // namespace mapp
