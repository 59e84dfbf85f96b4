// md_nst.hip -- MDNST on the GPU: velocity Verlet under a Nose-Hoover chain
// with a box that follows a target stress (src/md_nst.cpp). The per-atom
// work (the exponential velocity and position updates, the kinetic tensor)
// runs in three streaming kernels; the 3x3 box algebra, both thermostats and
// the box-velocity kicks are scalar work on the host between them. The
// neighbor list survives box changes: the image shifts follow the new box
// every step and the list is rebuilt only when an atom has moved half the
// skin (DynamicMD::update with a changed box, src/dynamic_md.h:38-56).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "kern.hpp"
#include "md.hpp"

namespace mapp {
namespace {

// lower-triangular 3x3 matrices packed by rows: 00, 10, 11, 20, 21, 22
struct Lt6 {
    double m[6];
};

Lt6 pack_lt(const double (&M)[3][3])
{
    return Lt6{{M[0][0], M[1][0], M[1][1], M[2][0], M[2][1], M[2][2]}};
}

// row vector times lower-triangular matrix: o_j = sum_{i >= j} a_i M_ij
__host__ __device__ inline void row_lt(const Lt6& M, double a0, double a1, double a2, double& o0, double& o1,
                                       double& o2)
{
    o0 = a0 * M.m[0] + a1 * M.m[1] + a2 * M.m[3];
    o1 = a1 * M.m[2] + a2 * M.m[4];
    o2 = a2 * M.m[5];
}

struct NstKern {
    Lt6 E, F0, F1, VH;
    double mass[kMaxElems];
};

// update_x_d__x__x_d, first half (src/md_nst.cpp:80-112): v <- xi v E +
// f F0 / m, the displacement u = (x V_H + v) F1, x <- x + u; partial sums of
// u for the centre-of-mass correction
__global__ __launch_bounds__(kBlock) void k_nst_drift(int n, double4* __restrict__ x, double4* __restrict__ v,
                                                      const float4* __restrict__ f, NstKern K, double xi,
                                                      double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double du[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        double4 xq = x[i];
        double4 vq = v[i];
        const float4 fq = f[i];
        const double m_inv = 1.0 / K.mass[(int)xq.w];
        double a0, a1, a2, b0, b1, b2;
        row_lt(K.E, vq.x, vq.y, vq.z, a0, a1, a2);
        row_lt(K.F0, (double)fq.x, (double)fq.y, (double)fq.z, b0, b1, b2);
        vq.x = xi * a0 + m_inv * b0;
        vq.y = xi * a1 + m_inv * b1;
        vq.z = xi * a2 + m_inv * b2;
        double w0, w1, w2;
        row_lt(K.VH, xq.x, xq.y, xq.z, w0, w1, w2);
        w0 += vq.x;
        w1 += vq.y;
        w2 += vq.z;
        row_lt(K.F1, w0, w1, w2, du[0], du[1], du[2]);
        xq.x += du[0];
        xq.y += du[1];
        xq.z += du[2];
        x[i] = xq;
        v[i] = vq;
    }
    block_sum_store<3>(du, part + (size_t)blk * kPartW);
}

// the centre-of-mass correction x <- x - c and DynamicMD::decide's skin
// check against the positions of the last build
__global__ __launch_bounds__(kBlock) void k_nst_shift(int n, double4* __restrict__ x, const double4* __restrict__ x0,
                                                      const double* __restrict__ com, double skin_sq4,
                                                      int* __restrict__ flags)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    // the centre-of-mass displacement: the drift's sums over n, read on the device
    const double c0 = com[0] / (double)n, c1 = com[1] / (double)n, c2 = com[2] / (double)n;
    double4 xq = x[i];
    xq.x -= c0;
    xq.y -= c1;
    xq.z -= c2;
    x[i] = xq;
    const double4 a = x0[i];
    if (rsq_exact(a.x, a.y, a.z, xq.x, xq.y, xq.z) > skin_sq4) atomicOr(&flags[1], 1);
}

// the 128-entry image-shift table of the record grid (refresh_frame)
__global__ __launch_bounds__(128) void k_shift_table(Rec16Frame R, uint4* __restrict__ out)
{
    const int c = threadIdx.x;
    uint4 sh = make_uint4(0u, 0u, 0u, 0u);
    if (c < 125) {
        const unsigned n0 = (unsigned)(c % 5 - 2), n1 = (unsigned)(c / 5 % 5 - 2), n2 = (unsigned)(c / 25 - 2);
        sh = make_uint4(n0 * R.cH[0] + n1 * R.cH[3] + n2 * R.cH[6], n1 * R.cH[4] + n2 * R.cH[7], n2 * R.cH[8], 0u);
    }
    out[c] = sh;
}

// second half (src/md_nst.cpp:133-152): v <- v E + f F0 / m when kick, and
// partial sums of the kinetic tensor m v (x) v
__global__ __launch_bounds__(kBlock) void k_nst_kick(int n, const double4* __restrict__ x, double4* __restrict__ v,
                                                     const float4* __restrict__ f, NstKern K, int kick,
                                                     double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double m = K.mass[(int)x[i].w];
        double4 vq = v[i];
        if (kick) {
            const float4 fq = f[i];
            double a0, a1, a2, b0, b1, b2;
            row_lt(K.E, vq.x, vq.y, vq.z, a0, a1, a2);
            row_lt(K.F0, (double)fq.x, (double)fq.y, (double)fq.z, b0, b1, b2);
            const double m_inv = 1.0 / m;
            vq.x = a0 + m_inv * b0;
            vq.y = a1 + m_inv * b1;
            vq.z = a2 + m_inv * b2;
            v[i] = vq;
        }
        acc[0] = vq.x * vq.x * m;
        acc[1] = vq.y * vq.x * m;
        acc[2] = vq.z * vq.x * m;
        acc[3] = vq.y * vq.y * m;
        acc[4] = vq.z * vq.y * m;
        acc[5] = vq.z * vq.z * m;
    }
    block_sum_store<6>(acc, part + (size_t)blk * kPartW);
}

__global__ __launch_bounds__(kBlock) void k_nst_scale_v(int n, double4* __restrict__ v, double s)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    double4 q = v[i];
    q.x *= s;
    q.y *= s;
    q.z *= s;
    v[i] = q;
}

// ---------------- host algebra ----------------

// phi(x) = (e^x - 1) / x and its first two derivatives, by their Taylor
// series below |x| = 1/64 (MDMath::f, df, ddf, src/md_nst.cpp:843-871)
double phi0(double x)
{
    if (std::fabs(x) < 0.015625) return 1.0 + x / 2.0 + x * x / 6.0 + x * x * x / 24.0;
    return (std::exp(x) - 1.0) / x;
}
double phi1(double x)
{
    if (std::fabs(x) < 0.015625) return 0.5 + x / 3.0 + x * x / 8.0 + x * x * x / 30.0;
    const double e = std::exp(x);
    return (e - (e - 1.0) / x) / x;
}
double phi2(double x)
{
    if (std::fabs(x) < 0.015625) return 1.0 / 3.0 + x / 4.0 + x * x / 10.0 + x * x * x / 36.0;
    const double e = std::exp(x);
    return (e - 2.0 * ((e - (e - 1.0) / x) / x)) / x;
}

constexpr double kSqrtEps = 1.4901161193847656e-08;
constexpr double kEps = 2.220446049250313e-16;

// A function g of a lower-triangular 3x3 matrix Z from its values on the
// diagonal and their divided differences: g(Z)_ii = g(z_i), g(Z)_ij =
// z_ij g[z_i, z_j], g(Z)_20 = z_20 g[z_2, z_0] + z_21 z_10 g[z_0, z_1, z_2].
// Near-coincident nodes take the derivative at their mean (thresholds of
// src/md_nst.cpp:890-901, 960-986).
struct TriFn {
    double (*g)(double);
    bool is_exp;
    double d1(double x, double y, double gx, double gy) const
    {
        const double m = 0.5 * (x + y);
        const double lim = is_exp ? kSqrtEps : std::sqrt(kEps * phi0(m) / phi2(m));
        if (std::fabs(x - y) < lim) return is_exp ? std::exp(m) : phi1(m);
        return (gx - gy) / (x - y);
    }
    double d2(double x, double y, double z, double g_yz, double g_zx, double g_xy) const
    {
        const double m = (x + y + z) / 3.0;
        const double lim = is_exp ? kSqrtEps : std::sqrt(kEps * phi0(m) / phi2(m));
        if (std::fabs(x - z) > lim) return (g_xy - g_yz) / (x - z);
        if (std::fabs(y - x) > lim) return (g_yz - g_zx) / (y - x);
        if (std::fabs(z - y) > lim) return (g_zx - g_xy) / (z - y);
        return is_exp ? std::exp(m) : phi2(m);
    }
    // g(s (A + a I)) scaled by out_scale, for the lower triangle of A
    void apply(const double (&A)[3][3], double a, double s, double out_scale, double (&G)[3][3]) const
    {
        double z[3], gz[3];
        for (int i = 0; i < 3; ++i) {
            z[i] = s * (A[i][i] + a);
            gz[i] = g(z[i]);
        }
        std::memset(G, 0, sizeof G);
        for (int i = 0; i < 3; ++i) G[i][i] = out_scale * gz[i];
        double dd[3][3] = {};
        for (int i = 1; i < 3; ++i)
            for (int j = 0; j < i; ++j)
                if (A[i][j] != 0.0) {
                    dd[i][j] = d1(z[i], z[j], gz[i], gz[j]);
                    G[i][j] = out_scale * s * A[i][j] * dd[i][j];
                }
        if (A[1][0] != 0.0 && A[2][1] != 0.0)
            G[2][0] += out_scale * s * s * A[2][1] * A[1][0] * d2(z[0], z[1], z[2], dd[2][1], dd[2][0], dd[1][0]);
    }
};

double exp_fn(double x) { return std::exp(x); }

// MDMath::calc (src/md_nst.cpp:887-1002): E = exp(-dt2 (A + a I)),
// F0 = dt2 phi(-dt2 (A + a I)), F1 = dt phi(dt A)
void nst_calc(double a, double dt, double dt2, const double (&A)[3][3], double (&E)[3][3], double (&F0)[3][3],
              double (&F1)[3][3])
{
    const TriFn fe{exp_fn, true}, fp{phi0, false};
    fe.apply(A, a, -dt2, 1.0, E);
    fp.apply(A, a, -dt2, dt2, F0);
    fp.apply(A, 0.0, dt, dt, F1);
}

// ThermostatNHC(dt, t_relax, L = 3, niters = 1)::operator() (src/thermostat.cpp:47-143)
struct HostNhc {
    double ddt, ddt2, freq_sq, eta[3] = {0.0, 0.0, 0.0};
    HostNhc(double dt, double t_relax) : ddt(dt), ddt2(0.5 * dt), freq_sq(1.0 / (t_relax * t_relax)) {}
    double operator()(double T_r, int ndof)
    {
        double* e = eta;
        double rescale = 1.0, prev, tmp;
        e[2] += ddt2 * (e[1] * e[1] - freq_sq);
        prev = -e[2];
        e[1] += ddt2 * phi0(prev * ddt2) * (prev * e[1] + e[0] * e[0] - freq_sq);
        prev = -e[1];
        e[0] += ddt2 * phi0(prev * ddt2) * (prev * e[0] + freq_sq * (T_r - 1.0));
        prev = -e[0];
        tmp = std::exp(ddt * prev);
        rescale *= tmp;
        T_r *= tmp * tmp;
        prev = (T_r - 1.0) * freq_sq;
        e[0] += ddt2 * phi0(-e[1] * ddt2) * (-e[1] * e[0] + prev);
        prev = ndof * e[0] * e[0] - freq_sq;
        e[1] += ddt2 * phi0(-e[2] * ddt2) * (-e[2] * e[1] + prev);
        prev = e[1] * e[1] - freq_sq;
        e[2] += ddt2 * prev;
        return rescale;
    }
};

inline int lt_voigt(int i, int j) { return i + j * 3 - j * (j + 1) / 2; }

}  // namespace

// Image shifts of the 16-byte records for the current box (the record grid q
// depends on the cut-off only), made on the device in stream order so that a
// box change costs no host synchronisation: code c = (n0+2) + 5(n1+2) +
// 25(n2+2) shifts by n0 h0 + n1 h1 + n2 h2 on the grid (modulo 2^32); codes
// 125..127 are unused
void MDSystem::refresh_frame()
{
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d)
            r16_.cH[3 * k + d] = (unsigned)(unsigned long long)std::llrint(box_.H[k][d] * r16_.inv_q);
    shift_.reserve(128);
    hipLaunchKernelGGL(k_shift_table, dim3(1), dim3(128), 0, stream_, r16_, shift_.get());
    r16_.shift = shift_.get();
}

int MDSystem::nst_run(const double* masses, double kB, double T, double dt, const double* S9, const double* tau9,
                      int nreset, int nsteps, MDReport* rep, double* T_part_out, double* H9, double* hist)
{
    MAPP_REQUIRE(nsteps >= 0 && dt > 0.0 && T > 0.0 && kB > 0.0 && nreset >= 0, ERR_ARG, "invalid NST parameters");
    const EamTables& t = ff_->tables();
    NstKern K{};
    for (int a = 0; a < t.ne; ++a) {
        MAPP_REQUIRE(masses[a] > 0.0, ERR_ARG, "masses must be positive");
        K.mass[a] = masses[a];
    }
    const double dt2 = dt / 2.0;
    const double ndof = 3.0 * n_;
    // pre_init (src/md_nst.cpp:291-319): free components (NaN) carry no box
    // motion; S_dev keeps the target's lower triangle, less the hydrostatic
    // part on the diagonal
    bool sdof[3][3] = {};
    double S[3][3] = {}, S_dev[3][3] = {}, tau[3][3] = {}, prefac[3][3] = {}, VH[3][3] = {};
    int ndof_baro = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) {
            tau[i][j] = tau9 ? tau9[3 * i + j] : 1000.0 * dt;
            MAPP_REQUIRE(tau[i][j] > 0.0, ERR_ARG, "tau must be positive");
            if (std::isnan(S9[3 * i + j])) continue;
            sdof[i][j] = true;
            S[i][j] = S9[3 * i + j];
            prefac[i][j] = 1.0 / (tau[i][j] * tau[i][j] * ndof * kB * T);
            ++ndof_baro;
        }
    MAPP_REQUIRE(ndof_baro > 0, ERR_ARG, "at least one stress component must be set (not NaN)");
    const double s_hyd = (S[0][0] + S[1][1] + S[2][2]) / 3.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) S_dev[i][j] = S[i][j] - (i == j ? s_hyd : 0.0);
    double T_baro = 0.0;
    HostNhc tp(dt2, 100.0 * dt), tb(dt2, 1000.0 * dt);

    const int nb = grid_for(n_);
    double* rows = part_.get() + (size_t)2 * nb * kPartW;
    double* red = red_.get() + 40;
    auto reduce = [&](int ncols, double* out) {
        reduce_rows(rows, nb, ncols, red);
        MAPP_HIP(hipMemcpyAsync(out, red, sizeof(double) * ncols, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    double mvv[6];
    hipLaunchKernelGGL(k_nst_kick, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), K, 0, rows);
    reduce(6, mvv);
    double T_part = (mvv[0] + (mvv[3] + mvv[5])) / (ndof * kB);
    MDReport r = force_calc();
    double S_pe[3][3] = {};
    auto stress = [&]() {
        for (int j = 0; j < 3; ++j)
            for (int i = j; i < 3; ++i) S_pe[i][j] = -r.e[1 + lt_voigt(i, j)] / box_.vol;
    };
    stress();
    double B_ref[3][3], vol_ref = box_.vol;
    std::memcpy(B_ref, box_.B, sizeof B_ref);

    // MDNST::update_V_H (src/md_nst.cpp:320-351)
    auto update_VH = [&]() {
        double L[3][3] = {}, SL[3][3] = {}, M2[3][3] = {};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = j; k <= i; ++k) L[i][j] += B_ref[i][k] * box_.H[k][j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) SL[i][j] += S_dev[i][k] * L[k][j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = 0; k < 3; ++k) M2[i][j] += L[k][i] * SL[k][j];
        for (int i = 0; i < 3; ++i)
            if (sdof[i][i]) VH[i][i] += (kB * T_part + box_.vol * s_hyd) * prefac[i][i] * dt2;
        T_baro = 0.0;
        for (int j = 0; j < 3; ++j)
            for (int i = j; i < 3; ++i) {
                if (!sdof[i][j]) continue;
                VH[i][j] += (vol_ref * M2[i][j] - box_.vol * S_pe[i][j] + mvv[lt_voigt(i, j)]) * prefac[i][j] * dt2;
                T_baro += VH[i][j] * VH[i][j] * tau[i][j] * tau[i][j];
            }
        T_baro *= ndof * T / ndof_baro;
    };
    auto scale_VH = [&](double s) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j) VH[i][j] *= s;
        T_baro *= s * s;
    };
    auto scale_part = [&](double s) {
        const double s2 = s * s;
        for (double& m : mvv) m *= s2;
        T_part *= s2;
    };

    const double skin_sq4 = 0.25 * skin_ * skin_;
    const int rebuilds0 = nrebuild_;
    const double vol0 = box_.vol;
    double fac, fac_x_d = 1.0;
    for (int step = 0; step < nsteps; ++step) {
        MAPP_REQUIRE(box_.vol > 0.125 * vol0 && box_.vol < 8.0 * vol0 && std::isfinite(T_part), ERR_STATE,
                     "the box left [vol0 / 8, 8 vol0]: NST run stopped");
        scale_VH(tb(T_baro / T, ndof_baro));
        fac_x_d *= fac = tp(T_part / T, (int)ndof);
        scale_part(fac);
        update_VH();
        // update_x_d__x__x_d (src/md_nst.cpp:80-153)
        double E[3][3], F0[3][3], F1[3][3];
        nst_calc((VH[0][0] + VH[1][1] + VH[2][2]) / ndof, dt, dt2, VH, E, F0, F1);
        K.E = pack_lt(E);
        K.F0 = pack_lt(F0);
        K.F1 = pack_lt(F1);
        K.VH = pack_lt(VH);
        prof_.begin(K_INTEGRATE, stream_);
        hipLaunchKernelGGL(k_nst_drift, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), K,
                           fac_x_d, rows);
        prof_.end(stream_);
        reduce_rows(rows, nb, 3, red);
        MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
        prof_.begin(K_INTEGRATE, stream_);
        hipLaunchKernelGGL(k_nst_shift, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), x0_.get(), red, skin_sq4,
                           flags_.get());
        prof_.end(stream_);
        // H <- H (I + F1 V_H)
        double M[3][3] = {}, Hn[9] = {};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j) {
                for (int k = j; k <= i; ++k) M[i][j] += F1[i][k] * VH[k][j];
                if (i == j) M[i][j] += 1.0;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = j; k <= i; ++k) Hn[3 * i + j] += box_.H[i][k] * M[k][j];
        int flag = 0;
        MAPP_HIP(hipMemcpyAsync(&flag, flags_.get() + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        box_ = make_box(Hn);
        if (flag) {
            setup_grid();
            rebuild();
        } else {
            refresh_frame();
        }
        // the force call's energy and virial and the kick's kinetic tensor
        // come back in one synchronisation (the kick uses the box velocity of
        // the first half-step, not the new stress)
        force_passes(nullptr);
        prof_.begin(K_INTEGRATE, stream_);
        hipLaunchKernelGGL(k_nst_kick, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), K, 1,
                           rows);
        prof_.end(stream_);
        reduce_rows(rows, nb, 6, red);
        MAPP_HIP(hipMemcpyAsync(r.e, red_.get(), sizeof(double) * 7, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipMemcpyAsync(mvv, red, sizeof(double) * 6, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        prof_.flush();
        stress();
        T_part = (mvv[0] + (mvv[3] + mvv[5])) / (ndof * kB);
        update_VH();
        fac_x_d = fac = tp(T_part / T, (int)ndof);
        scale_part(fac);
        scale_VH(tb(T_baro / T, ndof_baro));
        if (nreset && (step + 1) % nreset == 0) {
            std::memcpy(B_ref, box_.B, sizeof B_ref);
            vol_ref = box_.vol;
        }
        if (hist) {
            double* h = hist + 14 * (size_t)step;
            h[0] = T_part;
            h[1] = r.e[0];
            h[2] = box_.H[0][0];
            h[3] = box_.H[1][0];
            h[4] = box_.H[1][1];
            h[5] = box_.H[2][0];
            h[6] = box_.H[2][1];
            h[7] = box_.H[2][2];
            for (int j = 0; j < 3; ++j)
                for (int i = j; i < 3; ++i) h[8 + lt_voigt(i, j)] = S_pe[i][j];
        }
    }
    hipLaunchKernelGGL(k_nst_scale_v, dim3(nb), dim3(kBlock), 0, stream_, n_, v_.get(), fac_x_d);
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    if (rep) *rep = r;
    if (T_part_out) *T_part_out = T_part;
    if (H9)
        for (int k = 0; k < 9; ++k) H9[k] = box_.H[k / 3][k % 3];
    return nrebuild_ - rebuilds0;
}

}  // namespace mapp
