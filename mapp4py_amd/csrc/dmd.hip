// dmd.hip -- diffusive MD on gfx950: per-component neighbor list with the
// alpha-dependent cutoff, Gauss-Hermite smeared EAM in two passes (density /
// free energy, then forces on x and alpha and the chemical-potential sums),
// the skin check, and the conjugate-gradient free-energy minimiser.
//
// One thread per site loops over the site's components (c_dim <= 4), so the
// site force is accumulated in registers with no atomics. The smeared
// functions use FP64 tables and FP64 moments: the CG driver converges the
// gradient to 1e-6, below the noise floor FP32 spline values would leave.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>

#include "dmd.hpp"
#include "linesearch.hpp"
#include "cfg.hpp"
#include "kern.hpp"
#include "surrogate.hpp"

namespace mapp {

namespace {

constexpr double kPiInSq = 0.564189583547756286948079451561;  // 1/sqrt(pi)

struct DGrid {
    int nc[3], R[3];
    int ne, C;
    double scl;
    double cut_sk[kMaxElems * kMaxElems];  // r_c + skin (not squared: the DMD test is on r)
};

__device__ inline int dcell(double s, int nc)
{
    int c = (int)(s * (double)nc);
    return c < nc - 1 ? c : nc - 1;
}

__global__ __launch_bounds__(kBlock) void k_dcell_keys(int n, const double4* __restrict__ x, BoxDev bd, DGrid g,
                                                       const int* __restrict__ cell_rank, double4* __restrict__ s_tmp,
                                                       uint32_t* __restrict__ keys, int* __restrict__ vals)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 xi = x[i];
    double s0, s1, s2;
    x_to_s(bd.b, xi.x, xi.y, xi.z, s0, s1, s2);
    s_tmp[i] = make_double4(s0, s1, s2, 0.0);
    const int c = dcell(s0, g.nc[0]) + g.nc[0] * (dcell(s1, g.nc[1]) + g.nc[1] * dcell(s2, g.nc[2]));
    keys[i] = (uint32_t)cell_rank[c];
    vals[i] = i;
}

// sites into cell order; x re-expressed from the wrapped s (src/cell.h:208)
__global__ __launch_bounds__(kBlock) void k_dpermute(int n, int C, const int* __restrict__ perm,
                                                     const double4* __restrict__ s_tmp, BoxDev bd,
                                                     const double* __restrict__ al_old, const double* __restrict__ c_old,
                                                     const int* __restrict__ el_old, const int* __restrict__ id_old,
                                                     double4* __restrict__ x_new, double4* __restrict__ s4,
                                                     double4* __restrict__ x0, double* __restrict__ al_new,
                                                     double* __restrict__ c_new, int* __restrict__ el_new,
                                                     double* __restrict__ al0, int* __restrict__ id_new,
                                                     int* __restrict__ slot_of)
{
    const int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= n) return;
    const int o = perm[p];
    const double4 s = s_tmp[o];
    double a0, a1, a2;
    s_to_x(bd.h, s.x, s.y, s.z, a0, a1, a2);
    const double4 xn = make_double4(a0, a1, a2, 0.0);
    x_new[p] = xn;
    x0[p] = xn;
    s4[p] = s;
    for (int k = 0; k < C; ++k) {
        al_new[p * C + k] = al_old[o * C + k];
        al0[p * C + k] = al_old[o * C + k];
        c_new[p * C + k] = c_old[o * C + k];
        el_new[p * C + k] = el_old[o * C + k];
    }
    const int id = id_old[o];
    id_new[p] = id;
    slot_of[id] = p;
}

__global__ __launch_bounds__(kBlock) void k_dcell_start(int n, int ncells, const uint32_t* __restrict__ keys,
                                                        int* __restrict__ cell_start)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r > ncells) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < (uint32_t)r) lo = mid + 1; else hi = mid;
    }
    cell_start[r] = lo;
}

__device__ inline double dsqrt_exact(double v)
{
#pragma clang fp contract(off)
    return sqrt(v);
}

// Per-component full list (src/neighbor_dmd.cpp:102-194 predicate): for
// components ic of site i and jc of the image of site j, both present and
// preceded only by present components, keep j*C+jc when
// r - x_max sqrt(alpha_i^2 + alpha_j^2) < r_c + skin.
__global__ __launch_bounds__(kBlock) void k_dbuild_list(int n, const double4* __restrict__ s4, BoxDev bd, DGrid g,
                                                        const double* __restrict__ al, const double* __restrict__ cc,
                                                        const int* __restrict__ el, const int* __restrict__ cell_rank,
                                                        const int* __restrict__ cell_start, int maxn, int ldl,
                                                        uint32_t* __restrict__ list, int* __restrict__ nn,
                                                        int* __restrict__ flags)
{
#pragma clang fp contract(off)
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const int C = g.C;
    const double4 si = s4[i];
    double xi0, xi1, xi2;
    s_to_x(bd.h, si.x, si.y, si.z, xi0, xi1, xi2);
    int npres_i = 0;
    while (npres_i < C && cc[i * C + npres_i] >= 0.0) ++npres_i;
    int cnt[kMaxComp] = {0, 0, 0, 0};
    const int c0 = dcell(si.x, g.nc[0]), c1 = dcell(si.y, g.nc[1]), c2 = dcell(si.z, g.nc[2]);
    if (npres_i > 0) {
        for (int o2 = -g.R[2]; o2 <= g.R[2]; ++o2) {
            const int u2 = c2 + o2, n2 = floordiv(u2, g.nc[2]), w2 = u2 - n2 * g.nc[2];
            for (int o1 = -g.R[1]; o1 <= g.R[1]; ++o1) {
                const int u1 = c1 + o1, n1 = floordiv(u1, g.nc[1]), w1 = u1 - n1 * g.nc[1];
                for (int o0 = -g.R[0]; o0 <= g.R[0]; ++o0) {
                    const int u0 = c0 + o0, n0 = floordiv(u0, g.nc[0]), w0 = u0 - n0 * g.nc[0];
                    const int rank = cell_rank[w0 + g.nc[0] * (w1 + g.nc[1] * w2)];
                    const int jb = cell_start[rank], je = cell_start[rank + 1];
                    const bool self_img = (n0 | n1 | n2) == 0;
                    const uint32_t code = (uint32_t)((n0 + 2) + 5 * (n1 + 2) + 25 * (n2 + 2)) << kIndexBits;
                    for (int j = jb; j < je; ++j) {
                        if (self_img && j == i) continue;
                        const double4 sj = s4[j];
                        double xj0, xj1, xj2;
                        s_to_x(bd.h, shift_unit(sj.x, n0), shift_unit(sj.y, n1), shift_unit(sj.z, n2), xj0, xj1, xj2);
                        const double r = dsqrt_exact(rsq_exact(xi0, xi1, xi2, xj0, xj1, xj2));
                        for (int ic = 0; ic < npres_i; ++ic) {
                            const double ai = al[i * C + ic];
                            const int ei = el[i * C + ic];
                            for (int jc = 0; jc < C && cc[j * C + jc] >= 0.0; ++jc) {
                                const double aj = al[j * C + jc];
                                const double aij = dsqrt_exact(ai * ai + aj * aj);
                                if (r - g.scl * aij < g.cut_sk[ei * g.ne + el[j * C + jc]]) {
                                    if (cnt[ic] < maxn)
                                        list[(size_t)cnt[ic] * ldl + (i * C + ic)] = (uint32_t)(j * C + jc) | code;
                                    ++cnt[ic];
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    int mx = 0;
    for (int ic = 0; ic < C; ++ic) {
        nn[i * C + ic] = cnt[ic];
        mx = max(mx, cnt[ic]);
    }
    if (mx > maxn) atomicMax(&flags[0], mx);
}

__device__ inline void dpair_dx(const double4& xi, const double4& xj, uint32_t code, const double* __restrict__ h,
                                double& dx, double& dy, double& dz)
{
#pragma clang fp contract(off)
    dx = xi.x - xj.x;
    dy = xi.y - xj.y;
    dz = xi.z - xj.z;
    if (code != (uint32_t)kZeroImage) {
        const int n2 = (int)(code / 25u) - 2;
        const int rem = (int)(code % 25u);
        const int n1 = rem / 5 - 2, n0 = rem % 5 - 2;
        dx -= h[0] * n0 + h[1] * n1 + h[2] * n2;
        dy -= h[3] * n1 + h[4] * n2;
        dz -= h[5] * n2;
    }
}

// Separation, distance, pair width and the cut test of one listed pair,
// rounded the same way in every kernel (contraction off): the passes and the
// surrogate staging must take identical decisions pair by pair.
struct PairGeo {
    double dx, dy, dz, r, a;
    bool live;
};

__device__ inline PairGeo pair_geo(const double4& xi, const double4& xj, uint32_t code, const double* __restrict__ h,
                                   double ai, double aj, double xmax, double cut)
{
#pragma clang fp contract(off)
    PairGeo g;
    dpair_dx(xi, xj, code, h, g.dx, g.dy, g.dz);
    g.r = sqrt(g.dx * g.dx + g.dy * g.dy + g.dz * g.dz);
    g.a = sqrt(ai * ai + aj * aj);
    g.live = g.r - g.a * xmax < cut;
    return g;
}

__device__ inline double hv(const double4& c, double p) { return ((c.w * p + c.z) * p + c.y) * p + c.x; }
// an FP32 row: the cubic in FP32 at the FP32-rounded offset, returned as FP64
// for the FP64 moment sums
__device__ inline double hv(const float4& c, double p)
{
    const float q = (float)p;
    return (double)(((c.w * q + c.z) * q + c.y) * q + c.x);
}

// the quadrature passes' table rows: FP64 (parity mode) or FP32
template <bool F32>
using DRow = typename std::conditional<F32, float4, double4>::type;
template <bool F32>
__device__ inline const DRow<F32>* drows(const DmdArgs& A)
{
    if constexpr (F32)
        return A.tabf;
    else
        return A.tab;
}

// Smeared r*f over the odd extension, N-point Gauss-Hermite
// (src/ff_eam_dmd.cpp:195-253): rp = g, dr = (1/r) dg/dr, da = (1/a) dg/da.
template <bool DERIV>
__device__ inline void smear_quad(const DmdArgs& A, int tt, double r, double a, double (&rp)[3], double (&dr)[3],
                                  double (&da)[3])
{
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int l = 0; l < A.N; ++l) {
        const double xl = A.xi[l];
        const double u = r - xl * a;
        double p = fabs(u) * A.dr_inv;
        int m = (int)p;
        m = m < A.nr - 2 ? m : A.nr - 2;
        p -= (double)m;
        p = p < 1.0 ? p : 1.0;
        const double4* row = A.tab + ((size_t)tt * A.nr + m) * 3;
        const double sgn = u < 0.0 ? -A.wi[l] : A.wi[l];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double t = hv(row[k], p) * sgn;
            H[k][0] += t;
            if (DERIV) {
                H[k][1] += t * xl;
                H[k][2] += t * (xl * xl);
            }
        }
    }
    const double rinv = 1.0 / r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rp[k] = kPiInSq * rinv * H[k][0];
        if (DERIV) {
            const double ainv = 1.0 / a;
            dr[k] = -(rinv * rinv) * (rp[k] + 2.0 * kPiInSq * ainv * H[k][1]);
            da[k] = -(ainv * ainv) * (rp[k] - 2.0 * kPiInSq * rinv * H[k][2]);
        }
    }
}

__device__ inline void femb5(const DmdArgs& A, int e, double rho, double& F, double& dF)
{
    double p = rho * A.drho_inv;
    int m = (int)p;
    m = m < A.nrho - 2 ? m : A.nrho - 2;
    m = m > 0 ? m : 0;
    p -= (double)m;
    p = p < 1.0 ? p : 1.0;
    const double* c = A.F5 + ((size_t)e * A.nrho + m) * 5;
    F = (((c[4] * p + c[3]) * p + c[2]) * p + c[1]) * p + c[0];
    dF = (((4.0 * c[4] * p + 3.0 * c[3]) * p + 2.0 * c[2]) * p + c[1]) * A.drho_inv;
    if (rho > A.rho_max) F += dF * (rho - A.rho_max);
}

// F''(rho) on the interval femb5 uses (no second derivative past rho_max)
__device__ inline double ddemb5(const DmdArgs& A, int e, double rho)
{
    double p = rho * A.drho_inv;
    int m = (int)p;
    m = m < A.nrho - 2 ? m : A.nrho - 2;
    m = m > 0 ? m : 0;
    p -= (double)m;
    p = p < 1.0 ? p : 1.0;
    const double* cf = A.F5 + ((size_t)e * A.nrho + m) * 5;
    return ((12.0 * cf[4] * p + 6.0 * cf[3]) * p + 2.0 * cf[2]) * A.drho_inv * A.drho_inv;
}

__device__ inline double xlogx(double v) { return v == 0.0 ? 0.0 : v * log(v); }

// pass 1: rho per component, F(rho), F'(rho); free energy and entropy sums.
// energy_only mirrors __energy_calc (the c != 0 test on the site term).
template <bool SURR>
__global__ __launch_bounds__(kBlock) void k_dmd_pass1(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                      const double* __restrict__ cc, const int* __restrict__ el,
                                                      const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                      int ldl, DmdArgs A, BoxDev bd, SurrogateArgs S, int energy_only,
                                                      double* __restrict__ E, double* __restrict__ dE,
                                                      double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};  // free energy without the kT S term, S
    if (i < n) {
        const int C = A.C;
        const double4 xi = x[i];
        double cv = 1.0;
        SurrCursor cur;
        if (SURR) cur.begin(S, i, A.ne * A.ne);
        for (int ic = 0; ic < C; ++ic) {
            const int ci_idx = i * C + ic;
            const double ci = cc[ci_idx];
            if (ci > 0.0) cv -= ci;
            if (ci < 0.0) continue;
            const double ai = al[ci_idx];
            const int ei = el[ci_idx];
            double rho = 0.0, ep = 0.0;
            const int cnt = nn[ci_idx];
            for (int k = 0; k < cnt; ++k) {
                const uint32_t e = list[(size_t)k * ldl + ci_idx];
                const int jc = (int)(e & kIndexMask);
                const double4 xj = x[jc / C];
                const double aj = al[jc];
                const int ej = el[jc];
                const PairGeo pg = pair_geo(xi, xj, e >> kIndexBits, bd.h, ai, aj, A.xmax, A.cut[ei * A.ne + ej]);
                if (!pg.live) continue;
                const double r = pg.r, a = pg.a;
                double rp[3], drr[3], dal[3];
                int qin, qout;
                if (SURR && cur.next(S, A.ne, ei, ej, r, ai, aj, qin, qout))
                    surrogate_values<false>(S, qin, qout, r, a, rp, drr, dal);
                else
                    smear_quad<false>(A, ei * A.ne + ej, r, a, rp, drr, dal);
                rho += cc[jc] * rp[2];
                ep += 0.5 * ci * cc[jc] * rp[0];
            }
            double F, dF;
            femb5(A, ei, rho, F, dF);
            E[ci_idx] = F;
            dE[ci_idx] = dF;
            if (!energy_only || ci != 0.0) acc[0] += ci * (F + A.mu0[ei] - 3.0 * A.kBT * log(ai));
            acc[0] += ep;
            acc[1] += xlogx(ci);
        }
        acc[1] += xlogx(cv);
    }
    block_sum_store<2>(acc, part + (size_t)blk * kPartW);
}

// pass 2: forces on x (site), on alpha and the chemical-potential sums per
// component, virial; owner computes every pair from its own side
template <bool SURR>
__global__ __launch_bounds__(kBlock) void k_dmd_pass2(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                      const double* __restrict__ cc, const int* __restrict__ el,
                                                      const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                      int ldl, DmdArgs A, BoxDev bd, SurrogateArgs S,
                                                      const double* __restrict__ E, const double* __restrict__ dE,
                                                      double4* __restrict__ f, double* __restrict__ fa,
                                                      double* __restrict__ mu, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const int C = A.C;
        const double4 xi = x[i];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        SurrCursor cur;
        if (SURR) cur.begin(S, i, A.ne * A.ne);
        for (int ic = 0; ic < C; ++ic) {
            const int ci_idx = i * C + ic;
            const double ci = cc[ci_idx];
            if (ci < 0.0) {
                fa[ci_idx] = 0.0;
                mu[ci_idx] = 0.0;
                continue;
            }
            const double ai = al[ci_idx];
            const int ei = el[ci_idx];
            const double dEi = dE[ci_idx];
            double fai = 0.0, mui = 0.0;
            const int cnt = nn[ci_idx];
            for (int k = 0; k < cnt; ++k) {
                const uint32_t e = list[(size_t)k * ldl + ci_idx];
                const int jc = (int)(e & kIndexMask);
                const double4 xj = x[jc / C];
                const double aj = al[jc];
                const int ej = el[jc];
                const PairGeo pg = pair_geo(xi, xj, e >> kIndexBits, bd.h, ai, aj, A.xmax, A.cut[ei * A.ne + ej]);
                if (!pg.live) continue;
                const double r = pg.r, a = pg.a, dx = pg.dx, dy = pg.dy, dz = pg.dz;
                double rp[3], drr[3], dal[3];
                int qin, qout;
                if (SURR && cur.next(S, A.ne, ei, ej, r, ai, aj, qin, qout))
                    surrogate_values<true>(S, qin, qout, r, a, rp, drr, dal);
                else
                    smear_quad<true>(A, ei * A.ne + ej, r, a, rp, drr, dal);
                const double cj = cc[jc], dEj = dE[jc];
                const double fpair = -(drr[2] * dEi + drr[1] * dEj + drr[0]) * ci * cj;
                const double apair = -(dal[2] * dEi + dal[1] * dEj + dal[0]) * ci * cj;
                mui += cj * (rp[0] + rp[1] * dEj);
                fx += fpair * dx;
                fy += fpair * dy;
                fz += fpair * dz;
                fai += ai * apair;
                const double hf = 0.5 * fpair;
                acc[0] += hf * dx * dx;
                acc[1] += hf * dx * dy;
                acc[2] += hf * dx * dz;
                acc[3] += hf * dy * dy;
                acc[4] += hf * dy * dz;
                acc[5] += hf * dz * dz;
            }
            fa[ci_idx] = fai + 3.0 * A.kBT * ci / ai;
            mu[ci_idx] = E[ci_idx] + mui;
        }
        f[i] = make_double4(fx, fy, fz, 0.0);
    }
    block_sum_store<6>(acc, part + (size_t)blk * kPartW);
}

// ---------------- quadrature-only passes, batched ----------------
//
// Without the surrogate every listed pair takes the quadrature, so the
// passes resolve B list entries per batch with their loads staged across the
// batch: all list entries, then the neighbor data, then for each Hermite
// point the table rows of every pair of the batch. A wave keeps many
// independent loads in flight instead of one dependent chain per pair (the
// passes are bound by load latency, not FP64 arithmetic). Entries past the
// end of a row, and pairs beyond the cut, read a valid row and are selected
// away, so a batch has no branches. The arithmetic per pair is that of
// smear_quad, point by point in the same order.

// list entry k of component ci (entries past the row end: ci itself, zero image)
__device__ inline uint32_t dlist_entry(const uint32_t* __restrict__ list, int ldl, int maxn, int k, int ci, int cnt)
{
    const uint32_t raw = list[(size_t)(k < maxn ? k : maxn - 1) * ldl + ci];
    return k < cnt ? raw : ((uint32_t)ci | ((uint32_t)kZeroImage << kIndexBits));
}

// spline interval and offset of |u| (src/ff_eam_dmd.cpp:209-217 clamps)
__device__ inline int dknot(const DmdArgs& A, double u, double& p)
{
    double pm = fabs(u) * A.dr_inv;
    int m = (int)pm;
    m = m < A.nr - 2 ? m : A.nr - 2;
    pm -= (double)m;
    p = pm < 1.0 ? pm : 1.0;
    return m;
}

// Table sources of the batched passes. fetch<B, K>(ks, tt, m, live, c) gives,
// for each pair b of a batch, interval m[b] of function ks[j] (0 r*phi,
// 1 r*rho_{i->j}, 2 r*rho_{j->i}) of element pair tt[b]; a pair that is not
// live reads a valid row that the caller selects away. basis(p) holds what
// the offset p in [0, 1] contributes to every function of the interval, and
// eval(c, basis) is the interval's cubic there. Acc1 and Acc2 are the types
// in which passes 1 and 2 sum the Hermite points of one pair into its
// moments, and recip is the reciprocal of a pair's distance and width.

// coefficient rows [ne*ne][nr][3] in global memory, FP64 or FP32
template <bool F32>
struct RowsGlobal {
    using Row = DRow<F32>;
    const Row* T;
    int nr;
    template <int B, int K>
    __device__ void fetch(const int (&ks)[K], const int (&tt)[B], const int (&m)[B], const bool (&live)[B],
                          Row (&c)[B][K]) const
    {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const size_t row = live[b] ? ((size_t)tt[b] * nr + m[b]) * 3 : 0;
#pragma unroll
            for (int j = 0; j < K; ++j) c[b][j] = T[row + ks[j]];
        }
    }
    using Acc1 = double;
    using Acc2 = double;
    using Basis = double;
    __device__ static Basis basis(double p) { return p; }
    __device__ static double eval(const Row& c, double p) { return hv(c, p); }
    __device__ static double recip(double v) { return 1.0 / v; }
};

// FP32 knots. Every table function is the Hermite cubic of its interval's
// two knot values and slopes per unit offset (c2 = 3 dy - 2 s0 - s1,
// c3 = s0 + s1 - 2 dy, as the spline builder makes them), so one (value,
// slope) pair per knot holds a function in 8 bytes per knot, and an interval
// is two adjacent pairs. Only the distinct functions are kept: fn maps
// (element pair, function) to one of them. The top nwin knots of every
// distinct function sit in LDS; intervals below m_lo read the global copy.
// A random 16-byte row from L2 costs the vector cache one tag lookup per
// lane; from LDS a few bank cycles. The cubic is evaluated in the Hermite
// basis, y0 + (y1 - y0) h1 + s0 h2 + s1 h3, with h1..h3 of the offset shared
// by the three functions. Pass 2 sums one pair's Hermite points in FP32;
// pass 1, whose pair terms make the free energy, and both passes' sums over
// pairs stay FP64.
struct KnotsLds {
    using Row = float4;  // y0, s0, y1, s1
    const float2* glob;  // [nf][nr]
    const float2* win;   // [nf][nwin], LDS
    const int* fn;       // [ne*ne*3], LDS
    int nr, nwin, m_lo;
    template <int B, int K>
    __device__ void fetch(const int (&ks)[K], const int (&tt)[B], const int (&m)[B], const bool (&live)[B],
                          float4 (&c)[B][K]) const
    {
        bool low[B], any = false;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            low[b] = live[b] && m[b] < m_lo;
            any |= low[b];
            // m <= nr - 2, so r + 1 <= nwin - 1; pairs not in the window read row 0 (nwin >= 2)
            const int r = live[b] && !low[b] ? m[b] - m_lo : 0;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const float2* q = win + fn[tt[b] * 3 + ks[j]] * nwin + r;
                const float2 a = q[0], d = q[1];
                c[b][j] = make_float4(a.x, a.y, d.x, d.y);
            }
        }
        if (__ballot(any) != 0ull) {
#pragma unroll
            for (int b = 0; b < B; ++b)
                if (low[b])
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const float2* q = glob + (size_t)fn[tt[b] * 3 + ks[j]] * nr + m[b];
                        const float2 a = q[0], d = q[1];
                        c[b][j] = make_float4(a.x, a.y, d.x, d.y);
                    }
        }
    }
    using Acc1 = double;
    using Acc2 = float;
    struct Basis {
        float h1, h2, h3;
    };
    __device__ static Basis basis(double p)
    {
        const float q = (float)p, q1 = 1.0f - q, q2 = q * q;
        return Basis{q2 * (3.0f - 2.0f * q), q * q1 * q1, -q2 * q1};
    }
    __device__ static float eval(const float4& c, const Basis& h)
    {
        return c.x + (c.z - c.x) * h.h1 + c.y * h.h2 + c.w * h.h3;
    }
    // the hardware reciprocal and one Newton step (squares its relative
    // error), in place of the correctly rounded division
    __device__ static double recip(double v)
    {
        const double y = __builtin_amdgcn_rcp(v);
        return fma(fma(-v, y, 1.0), y, y);
    }
};

// Neighbor data of the batched passes: width, concentration and element of
// a listed component (pass 1), and its F'(rho) (pass 2). Every item is a
// gather from a random address, so the LDS passes read them packed: (alpha,
// c) in one 16-byte record, and (F', element) in another that pass 1 writes
// for pass 2.
struct NbPlain {
    const double* al;
    const double* cc;
    const int* el;
    const double* dE;
    __device__ void p1(int jc, double& a, double& c, int& e) const
    {
        a = al[jc];
        c = cc[jc];
        e = el[jc];
    }
    __device__ void p2(int jc, double& a, double& c, int& e, double& d) const
    {
        p1(jc, a, c, e);
        d = dE[jc];
    }
    // pass 1's output for pass 2 beside dE: none
    __device__ void put(int, double, int) const {}
};

struct NbPacked {
    const double2* ac;  // alpha, c
    const int* el;
    double2* ed;        // F'(rho), element (pass 1 writes it)
    __device__ void p1(int jc, double& a, double& c, int& e) const
    {
        const double2 v = ac[jc];
        a = v.x;
        c = v.y;
        e = el[jc];
    }
    __device__ void p2(int jc, double& a, double& c, int& e, double& d) const
    {
        const double2 v = ac[jc], w = ed[jc];
        a = v.x;
        c = v.y;
        d = w.x;
        e = (int)w.y;
    }
    __device__ void put(int ci, double d, int e) const { ed[ci] = make_double2(d, (double)e); }
};

// pass 1 of site i: rho of every component (smeared r*rho_ba) and the pair
// energy (smeared r*phi); only table functions 0 and 2 are read
template <int B, class Src, class Nb>
__device__ inline void pass1_site(int i, const double4* __restrict__ x, const double* __restrict__ al,
                                  const double* __restrict__ cc, const int* __restrict__ el,
                                  const int* __restrict__ nn, const uint32_t* __restrict__ list, int ldl, int maxn,
                                  const DmdArgs& A, const BoxDev& bd, int energy_only, const Src& T, const Nb& nb,
                                  double* __restrict__ E, double* __restrict__ dE, double (&acc)[2])
{
    const int ks[2] = {0, 2};
    const int C = A.C;
    const double4 xi = x[i];
    double cv = 1.0;
    for (int ic = 0; ic < C; ++ic) {
        const int ci_idx = i * C + ic;
        const double ci = cc[ci_idx];
        if (ci > 0.0) cv -= ci;
        if (ci < 0.0) continue;
        const double ai = al[ci_idx];
        const int ei = el[ci_idx];
        double rho = 0.0, ep = 0.0;
        const int cnt = nn[ci_idx];
        for (int k0 = 0; k0 < cnt; k0 += B) {
            uint32_t e[B];
            double4 xj[B];
            double aj[B], cj[B];
            int tt[B];
#pragma unroll
            for (int b = 0; b < B; ++b) e[b] = dlist_entry(list, ldl, maxn, k0 + b, ci_idx, cnt);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int jc = (int)(e[b] & kIndexMask);
                int ej;
                xj[b] = x[jc / C];
                nb.p1(jc, aj[b], cj[b], ej);
                tt[b] = ei * A.ne + ej;
            }
            PairGeo pg[B];
            bool live[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                pg[b] = pair_geo(xi, xj[b], e[b] >> kIndexBits, bd.h, ai, aj[b], A.xmax, A.cut[tt[b]]);
                live[b] = pg[b].live && k0 + b < cnt;
            }
            using Acc = typename Src::Acc1;
            Acc H0[B], H2[B];
#pragma unroll
            for (int b = 0; b < B; ++b) H0[b] = H2[b] = 0;
            for (int l = 0; l < A.N; ++l) {
                const double xl = A.xi[l], wl = A.wi[l];
                int m[B];
                double p[B], s[B];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const double u = pg[b].r - xl * pg[b].a;
                    m[b] = dknot(A, u, p[b]);
                    s[b] = u < 0.0 ? -wl : wl;
                }
                typename Src::Row c[B][2];
                T.template fetch<B, 2>(ks, tt, m, live, c);
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const typename Src::Basis h = Src::basis(p[b]);
                    const Acc sb = (Acc)s[b];
                    H0[b] += Src::eval(c[b][0], h) * sb;
                    H2[b] += Src::eval(c[b][1], h) * sb;
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const double rinv = Src::recip(pg[b].r);
                const double rp0 = kPiInSq * rinv * (double)H0[b], rp2 = kPiInSq * rinv * (double)H2[b];
                rho += live[b] ? cj[b] * rp2 : 0.0;
                ep += live[b] ? 0.5 * ci * cj[b] * rp0 : 0.0;
            }
        }
        double F, dF;
        femb5(A, ei, rho, F, dF);
        E[ci_idx] = F;
        dE[ci_idx] = dF;
        nb.put(ci_idx, dF, ei);
        if (!energy_only || ci != 0.0) acc[0] += ci * (F + A.mu0[ei] - 3.0 * A.kBT * log(ai));
        acc[0] += ep;
        acc[1] += xlogx(ci);
    }
    acc[1] += xlogx(cv);
}

// pass 2 of site i: forces on x and alpha, chemical-potential sums, virial.
// With UNIT it also writes the gradient per unit c that
// ForceFieldEAMDMD::calc_mu leaves in f and f_alpha
// (src/ff_eam_dmd.cpp:1189-1278): every pair weighted by c_j and the
// component's fcoef (src/ff_eam_dmd.cpp:995-1046) instead of c_i c_j, and
// d/dalpha with -3 kT / alpha, defined also where c_i = 0.
template <int B, bool UNIT, class Src, class Nb>
__device__ inline void pass2_site(int i, const double4* __restrict__ x, const double* __restrict__ al,
                                  const double* __restrict__ cc, const int* __restrict__ el,
                                  const int* __restrict__ nn, const uint32_t* __restrict__ list, int ldl, int maxn,
                                  const DmdArgs& A, const BoxDev& bd, const Src& T, const Nb& nb,
                                  const double* __restrict__ E,
                                  const double* __restrict__ dE, double4* __restrict__ f, double* __restrict__ fa,
                                  double* __restrict__ mu, double4* __restrict__ gu, double* __restrict__ gau,
                                  const JState& J, double (&acc)[6])
{
    const int ks[3] = {0, 1, 2};
    const int C = A.C;
    const double4 xi = x[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double c_tot = 0.0;
    if (UNIT)
        for (int q = 0; q < C; ++q) {
            const double v = cc[i * C + q];
            if (v > 0.0) c_tot += v;
        }
    for (int ic = 0; ic < C; ++ic) {
        const int ci_idx = i * C + ic;
        const double ci = cc[ci_idx];
        if (ci < 0.0) {
            fa[ci_idx] = 0.0;
            mu[ci_idx] = 0.0;
            if (UNIT) {
                gau[ci_idx] = 0.0;
                J.cnt[ci_idx] = 0;
                J.ddE[ci_idx] = 0.0;
            }
            continue;
        }
        const double ai = al[ci_idx];
        const int ei = el[ci_idx];
        const double dEi = dE[ci_idx];
        // fcoef: 1 with one component; else c_i / sum of positive c, or
        // 1 for the zero components of an empty site
        const double fcoef = C == 1 ? 1.0 : (c_tot == 0.0 ? (ci == 0.0 ? 1.0 : 0.0) : ci / c_tot);
        double fai = 0.0, mui = 0.0, gai = 0.0, rho = 0.0;
        const int cnt = nn[ci_idx];
        for (int k0 = 0; k0 < cnt; k0 += B) {
            uint32_t e[B];
            double4 xj[B];
            double aj[B], cj[B], dEj[B];
            int tt[B];
#pragma unroll
            for (int b = 0; b < B; ++b) e[b] = dlist_entry(list, ldl, maxn, k0 + b, ci_idx, cnt);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int jc = (int)(e[b] & kIndexMask);
                int ej;
                xj[b] = x[jc / C];
                nb.p2(jc, aj[b], cj[b], ej, dEj[b]);
                tt[b] = ei * A.ne + ej;
            }
            PairGeo pg[B];
            bool live[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                pg[b] = pair_geo(xi, xj[b], e[b] >> kIndexBits, bd.h, ai, aj[b], A.xmax, A.cut[tt[b]]);
                live[b] = pg[b].live && k0 + b < cnt;
            }
            using Acc = typename Src::Acc2;
            Acc H[B][3][3];
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int k = 0; k < 3; ++k) H[b][k][0] = H[b][k][1] = H[b][k][2] = 0;
            for (int l = 0; l < A.N; ++l) {
                const double xl = A.xi[l], wl = A.wi[l];
                const Acc xa = (Acc)xl, xa2 = (Acc)(xl * xl);
                int m[B];
                double p[B], s[B];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const double u = pg[b].r - xl * pg[b].a;
                    m[b] = dknot(A, u, p[b]);
                    s[b] = u < 0.0 ? -wl : wl;
                }
                typename Src::Row c[B][3];
                T.template fetch<B, 3>(ks, tt, m, live, c);
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const typename Src::Basis h = Src::basis(p[b]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const Acc t = Src::eval(c[b][k], h) * (Acc)s[b];
                        H[b][k][0] += t;
                        H[b][k][1] += t * xa;
                        H[b][k][2] += t * xa2;
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const double rinv = Src::recip(pg[b].r), ainv = Src::recip(pg[b].a);
                double rp[3], drr[3], dal[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    rp[k] = kPiInSq * rinv * (double)H[b][k][0];
                    drr[k] = -(rinv * rinv) * (rp[k] + 2.0 * kPiInSq * ainv * (double)H[b][k][1]);
                    dal[k] = -(ainv * ainv) * (rp[k] - 2.0 * kPiInSq * rinv * (double)H[b][k][2]);
                }
                const double fpair = -(drr[2] * dEi + drr[1] * dEj[b] + drr[0]) * ci * cj[b];
                const double apair = -(dal[2] * dEi + dal[1] * dEj[b] + dal[0]) * ci * cj[b];
                const double du = cj[b] * (rp[0] + rp[1] * dEj[b]);
                const double fp = live[b] ? fpair : 0.0;
                mui += live[b] ? du : 0.0;
                fai += live[b] ? ai * apair : 0.0;
                const double dx = pg[b].dx, dy = pg[b].dy, dz = pg[b].dz;
                fx += fp * dx;
                fy += fp * dy;
                fz += fp * dz;
                if (UNIT && k0 + b < cnt) {  // the Jacobian's slot values (as k_dmd_jrp)
                    const size_t slot = (size_t)(k0 + b) * ldl + ci_idx;
                    J.idx[slot] = e[b] & kIndexMask;
                    J.r2[slot] = live[b] ? rp[2] : 0.0;
                    J.r1[slot] = live[b] ? rp[1] : 0.0;
                    J.t0[slot] = live[b] ? rp[0] + dEi * rp[2] + dEj[b] * rp[1] : 0.0;
                    rho += live[b] ? cj[b] * rp[2] : 0.0;
                }
                if (UNIT) {
                    const double gp = live[b] ? (drr[2] * dEi + drr[1] * dEj[b] + drr[0]) * cj[b] * fcoef : 0.0;
                    gx += gp * dx;
                    gy += gp * dy;
                    gz += gp * dz;
                    gai += live[b] ? ai * (dal[2] * dEi + dal[1] * dEj[b] + dal[0]) * cj[b] : 0.0;
                }
                const double hf = 0.5 * fp;
                acc[0] += hf * dx * dx;
                acc[1] += hf * dx * dy;
                acc[2] += hf * dx * dz;
                acc[3] += hf * dy * dy;
                acc[4] += hf * dy * dz;
                acc[5] += hf * dz * dz;
            }
        }
        fa[ci_idx] = fai + 3.0 * A.kBT * ci / ai;
        mu[ci_idx] = E[ci_idx] + mui;
        if (UNIT) {
            gau[ci_idx] = gai - 3.0 * A.kBT / ai;
            J.cnt[ci_idx] = cnt;
            J.ddE[ci_idx] = ddemb5(A, ei, rho);
        }
    }
    f[i] = make_double4(fx, fy, fz, 0.0);
    if (UNIT) gu[i] = make_double4(gx, gy, gz, 0.0);
}

// the batched passes over the global rows, one site per thread
template <int B, bool F32>
__global__ __launch_bounds__(kBlock) void k_dmd_pass1q(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                       const double* __restrict__ cc, const int* __restrict__ el,
                                                       const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                       int ldl, int maxn, DmdArgs A, BoxDev bd, int energy_only,
                                                       double* __restrict__ E, double* __restrict__ dE,
                                                       double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};  // free energy without the kT S term, S
    if (i < n)
        pass1_site<B>(i, x, al, cc, el, nn, list, ldl, maxn, A, bd, energy_only, RowsGlobal<F32>{drows<F32>(A), A.nr},
                      NbPlain{al, cc, el, dE}, E, dE, acc);
    block_sum_store<2>(acc, part + (size_t)blk * kPartW);
}

template <int B, bool UNIT, bool F32>
__global__ __launch_bounds__(kBlock) void k_dmd_pass2q(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                       const double* __restrict__ cc, const int* __restrict__ el,
                                                       const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                       int ldl, int maxn, DmdArgs A, BoxDev bd,
                                                       const double* __restrict__ E, const double* __restrict__ dE,
                                                       double4* __restrict__ f, double* __restrict__ fa,
                                                       double* __restrict__ mu, double* __restrict__ part,
                                                       double4* __restrict__ gu, double* __restrict__ gau, JState J)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < n)
        pass2_site<B, UNIT>(i, x, al, cc, el, nn, list, ldl, maxn, A, bd, RowsGlobal<F32>{drows<F32>(A), A.nr},
                            NbPlain{al, cc, el, dE}, E, dE, f, fa, mu, gu, gau, J, acc);
    block_sum_store<6>(acc, part + (size_t)blk * kPartW);
}

// Persistent passes with the FP32 knots in LDS: one workgroup of NT threads
// per CU loads the knot window once and walks chunks of NT sites. The chunks
// form 8 contiguous ranges, one per XCD, and the workgroups of an XCD stride
// through its range side by side (neighbor halos shared in that XCD's L2).
template <int NT>
struct SiteWalk {
    int c, end, step;
    __device__ explicit SiteWalk(int n)
    {
        const int nchunks = (n + NT - 1) / NT;
        const int per = (nchunks + kXcd - 1) / kXcd;
        const int xcd = (int)blockIdx.x % kXcd;
        c = xcd * per + (int)blockIdx.x / kXcd;
        end = min((xcd + 1) * per, nchunks);
        step = (int)gridDim.x / kXcd;
    }
};

template <int NT>
__device__ inline KnotsLds load_knots(const KnotArgs& K, float2* win, int* fn)
{
    for (int r = threadIdx.x; r < K.nf * K.nwin; r += NT) {
        const int q = r / K.nwin;
        win[r] = K.glob[(size_t)q * K.nr + K.m_lo + (r - q * K.nwin)];
    }
    for (int t = threadIdx.x; t < K.nfn; t += NT) fn[t] = K.fn[t];
    __syncthreads();
    return KnotsLds{K.glob, win, fn, K.nr, K.nwin, K.m_lo};
}

// (alpha, c) of every component in one record for the LDS passes
__global__ __launch_bounds__(kBlock) void k_pack_ac(int nc, const double* __restrict__ al,
                                                    const double* __restrict__ cc, double2* __restrict__ ac)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < nc) ac[k] = make_double2(al[k], cc[k]);
}

template <int B, int NT>
__global__ __launch_bounds__(NT) void k_dmd_pass1k(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                   const double* __restrict__ cc, const int* __restrict__ el,
                                                   const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                   int ldl, int maxn, DmdArgs A, BoxDev bd, KnotArgs K,
                                                   const double2* __restrict__ ac, double2* __restrict__ ed,
                                                   int energy_only, double* __restrict__ E, double* __restrict__ dE,
                                                   double* __restrict__ part)
{
    extern __shared__ float2 knot_win[];
    __shared__ int knot_fn[kMaxElems * kMaxElems * 3];
    const KnotsLds T = load_knots<NT>(K, knot_win, knot_fn);
    const NbPacked nb{ac, el, ed};
    double acc[2] = {0.0, 0.0};
    for (SiteWalk<NT> w(n); w.c < w.end; w.c += w.step) {
        const int i = w.c * NT + (int)threadIdx.x;
        if (i < n) pass1_site<B>(i, x, al, cc, el, nn, list, ldl, maxn, A, bd, energy_only, T, nb, E, dE, acc);
    }
    block_sum_store<2, NT>(acc, part + (size_t)blockIdx.x * kPartW);
}

template <int B, int NT>
__global__ __launch_bounds__(NT) void k_dmd_pass2k(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                   const double* __restrict__ cc, const int* __restrict__ el,
                                                   const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                   int ldl, int maxn, DmdArgs A, BoxDev bd, KnotArgs K,
                                                   const double2* __restrict__ ac, double2* __restrict__ ed,
                                                   const double* __restrict__ E, const double* __restrict__ dE,
                                                   double4* __restrict__ f, double* __restrict__ fa,
                                                   double* __restrict__ mu, double* __restrict__ part)
{
    extern __shared__ float2 knot_win[];
    __shared__ int knot_fn[kMaxElems * kMaxElems * 3];
    const KnotsLds T = load_knots<NT>(K, knot_win, knot_fn);
    const NbPacked nb{ac, el, ed};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (SiteWalk<NT> w(n); w.c < w.end; w.c += w.step) {
        const int i = w.c * NT + (int)threadIdx.x;
        if (i < n)
            pass2_site<B, false>(i, x, al, cc, el, nn, list, ldl, maxn, A, bd, T, nb, E, dE, f, fa, mu, nullptr,
                                 nullptr, JState{}, acc);
    }
    block_sum_store<6, NT>(acc, part + (size_t)blockIdx.x * kPartW);
}

// ---------------- concentration rates ----------------

// ForceFieldEAMDMD::calc_Q with theta (src/ff_eam_dmd.cpp:1611-1633): the
// exchange barrier between two sites in units of kT, the inverse squared
// width alpha_Q_sq and theta (the Jacobian's derivative term)
__device__ inline void exchange_q(double gamma_i, double gamma_j, double mu_ji, double& Q, double& alpha_Q_sq,
                                  double& theta)
{
    const double a = 5.0 * (gamma_i - gamma_j - 6.0 * mu_ji);
    const double xi = -2.0 * (gamma_i + gamma_j);
    const double dsq = sqrt((xi - a) * (xi - a) - 8.0 * a * gamma_i);
    const double z0 = -4.0 * gamma_i / ((xi - a) - dsq);
    const double dz0 = 30.0 * z0 * (1.0 - z0) / dsq;
    const double dQ = z0 * z0 * z0 * (6.0 * z0 * z0 - 15.0 * z0 + 10.0);
    Q = z0 * z0 * (1.0 - z0) * (1.0 - z0) * ((1.0 - z0) * gamma_i + z0 * gamma_j) + dQ * mu_ji;
    alpha_Q_sq = 1.0 / ((1.0 - z0) * gamma_i + z0 * gamma_j);
    theta = alpha_Q_sq * (gamma_i - gamma_j) * dz0 - dQ;
}

// ForceFieldEAMDMD::__c_d_calc (src/ff_eam_dmd.cpp:1359-1473) over the second
// list of NeighborDMD::create_2nd_list (src/neighbor_dmd.cpp:224-273): the
// entries of the component's list with the same element and r^2 < r_crd^2.
// One thread per component; the owner sums every exchange of its component
// from its own side (the rate of a pair seen from the other side is minus
// this one: the barrier there is Q - mu_ji). Partial max |c_d| per workgroup.
__global__ __launch_bounds__(kBlock) void k_dmd_cd(int nc, const double4* __restrict__ x, const double* __restrict__ al,
                                                   const double* __restrict__ cc, const int* __restrict__ el,
                                                   const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                   int ldl, int C, BoxDev bd, CdArgs P,
                                                   const double* __restrict__ mu, double* __restrict__ cd,
                                                   double* __restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double sh[kBlock];
    const int blk = xcd_block();
    const int ic = blk * kBlock + threadIdx.x;
    double dci = 0.0;
    if (ic < nc) {
        const double ci = cc[ic];
        if (ci >= 0.0) {
            const int site = ic / C, ei = el[ic];
            const double4 xi = x[site];
            const double alpha_i = al[ic], alpha_sq_i = alpha_i * alpha_i, mu_i = mu[ic];
            double cv_i = 1.0;
            for (int k = 0; k < C; ++k) {
                const double v = cc[site * C + k];
                if (v > 0.0) cv_i -= v;
            }
            const double zeta = P.zeta[ei], c1 = P.c1[ei], rsq_crd = P.rsq_crd[ei];
            const int cnt = nn[ic];
            for (int k = 0; k < cnt; ++k) {
                const uint32_t e = list[(size_t)k * ldl + ic];
                const int jc = (int)(e & kIndexMask);
                if (el[jc] != ei) continue;
                double dx, dy, dz;
                dpair_dx(xi, x[jc / C], e >> kIndexBits, bd.h, dx, dy, dz);
                const double rsq = dx * dx + (dy * dy + dz * dz);
                if (rsq >= rsq_crd) continue;
                const double alpha_j = al[jc], alpha_sq_j = alpha_j * alpha_j;
                const double gamma_i = rsq / (zeta * alpha_sq_i), gamma_j = rsq / (zeta * alpha_sq_j);
                const double mu_ji = P.beta * (mu[jc] - mu_i);
                double Q, gamma_ij_inv, theta;
                exchange_q(gamma_i, gamma_j, mu_ji, Q, gamma_ij_inv, theta);
                const double coef0 = -c1 * rsq * gamma_ij_inv * exp(-Q + 0.5 * mu_ji);
                double cv_j = 1.0;
                const int sj = jc / C;
                for (int q = 0; q < C; ++q) {
                    const double v = cc[sj * C + q];
                    if (v > 0.0) cv_j -= v;
                }
                dci += coef0 * (ci * cv_j * exp(-0.5 * mu_ji) / (alpha_sq_i * alpha_i) -
                                cc[jc] * cv_i * exp(0.5 * mu_ji) / (alpha_sq_j * alpha_j));
            }
        }
        cd[ic] = dci;
    }
    sh[threadIdx.x] = fabs(dci);
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blk * kPartW] = sh[0];
}

// ---------------- Jacobian of the concentration rates ----------------

// Products with the Jacobian of the rates (ForceFieldEAMDMD::__J,
// src/ff_eam_dmd.cpp:1475-1575) at a fixed state. The state kernels below
// run once per state and keep what every product reads, so a product is
// three gathers:
//  * every entry of every component's list, slot-major at its list position
//    (so a wave stores one row at a time): the partner, rho_{j->i},
//    rho_{i->j} and t0 = r phi + dE_i rho_{j->i} + dE_j rho_{i->j}, zero past
//    the smeared cut. c_d's second pass writes them; k_dmd_jrp otherwise;
//  * the second list (same element, r < r_crd) with the exchange
//    coefficients of src/ff_eam_dmd.cpp:1430-1462 folded in: per entry
//    f0 = coef0 M0, f1 = coef0 d_i c_i, f2 = coef0 d_j cv_i, and per component
//    S0 = sum coef0 M0, S1 = sum coef0 d_i cv_j, S2 = sum coef0 d_j c_j, so that
//    Ax_i = S1 x_i + S2 xs_i - S0 b0_i + sum_j (f0 b0_j - f1 xs_j - f2 x_j)
//    with xs the sum of x over a site's present components.

// quadrature values of every list entry, rho and F''(rho)
template <int B>
__global__ __launch_bounds__(kBlock) void k_dmd_jrp(int n, const double4* __restrict__ x, const double* __restrict__ al,
                                                    const double* __restrict__ cc, const int* __restrict__ el,
                                                    const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                    int ldl, int maxn, DmdArgs A, BoxDev bd,
                                                    const double* __restrict__ dE, JState J)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const int C = A.C;
    const double4 xi = x[i];
    for (int ic = 0; ic < C; ++ic) {
        const int ci_idx = i * C + ic;
        const double ci = cc[ci_idx];
        if (ci < 0.0) {
            J.ddE[ci_idx] = 0.0;
            J.cnt[ci_idx] = 0;
            continue;
        }
        const double ai = al[ci_idx], dEi = dE[ci_idx];
        const int ei = el[ci_idx];
        double rho = 0.0;
        const int cnt = nn[ci_idx];
        for (int k0 = 0; k0 < cnt; k0 += B) {
            uint32_t e[B];
            double4 xj[B];
            double aj[B], cj[B];
            int tt[B];
#pragma unroll
            for (int b = 0; b < B; ++b) e[b] = dlist_entry(list, ldl, maxn, k0 + b, ci_idx, cnt);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int jc = (int)(e[b] & kIndexMask);
                xj[b] = x[jc / C];
                aj[b] = al[jc];
                cj[b] = cc[jc];
                tt[b] = ei * A.ne + el[jc];
            }
            PairGeo pg[B];
            bool live[B];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                pg[b] = pair_geo(xi, xj[b], e[b] >> kIndexBits, bd.h, ai, aj[b], A.xmax, A.cut[tt[b]]);
                live[b] = pg[b].live && k0 + b < cnt;
            }
            double H[B][3];
#pragma unroll
            for (int b = 0; b < B; ++b) H[b][0] = H[b][1] = H[b][2] = 0.0;
            for (int l = 0; l < A.N; ++l) {
                const double xl = A.xi[l], wl = A.wi[l];
                double4 c[B][3];
                double p[B], sg[B];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const double u = pg[b].r - xl * pg[b].a;
                    const int m = dknot(A, u, p[b]);
                    sg[b] = u < 0.0 ? -wl : wl;
                    const size_t row = live[b] ? ((size_t)tt[b] * A.nr + m) * 3 : 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) c[b][k] = A.tab[row + k];
                }
#pragma unroll
                for (int b = 0; b < B; ++b)
#pragma unroll
                    for (int k = 0; k < 3; ++k) H[b][k] += hv(c[b][k], p[b]) * sg[b];
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (k0 + b >= cnt) continue;
                const int jc = (int)(e[b] & kIndexMask);
                const double rinv = 1.0 / pg[b].r;
                const double v0 = live[b] ? kPiInSq * rinv * H[b][0] : 0.0;
                const double v1 = live[b] ? kPiInSq * rinv * H[b][1] : 0.0;
                const double v2 = live[b] ? kPiInSq * rinv * H[b][2] : 0.0;
                const size_t slot = (size_t)(k0 + b) * ldl + ci_idx;
                J.idx[slot] = (uint32_t)jc;
                J.r2[slot] = v2;
                J.r1[slot] = v1;
                J.t0[slot] = v0 + dEi * v2 + dE[jc] * v1;
                rho += cj[b] * v2;
            }
        }
        J.cnt[ci_idx] = cnt;
        // F''(rho) on the interval femb5 uses (no second derivative past rho_max)
        double p = rho * A.drho_inv;
        int m = (int)p;
        m = m < A.nrho - 2 ? m : A.nrho - 2;
        m = m > 0 ? m : 0;
        p -= (double)m;
        p = p < 1.0 ? p : 1.0;
        const double* cf = A.F5 + ((size_t)ei * A.nrho + m) * 5;
        J.ddE[ci_idx] = ((12.0 * cf[4] * p + 6.0 * cf[3]) * p + 2.0 * cf[2]) * A.drho_inv * A.drho_inv;
    }
}

// the second list with the exchange coefficients of the state. Each
// component keeps its own exchanges; seen from j a term is minus this one
// (coef0 and M0 are symmetric in i and j, d_i and d_j trade places).
__global__ __launch_bounds__(kBlock) void k_dmd_jex(int nc, const double4* __restrict__ x, const double* __restrict__ al,
                                                    const double* __restrict__ cc, const int* __restrict__ el,
                                                    const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                    int ldl, int C, BoxDev bd, CdArgs P,
                                                    const double* __restrict__ mu, JState J)
{
#pragma clang fp contract(off)
    const int ic = xcd_block() * kBlock + threadIdx.x;
    if (ic >= nc) return;
    const double ci = cc[ic];
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    int m = 0;
    if (ci >= 0.0) {
        const int site = ic / C, ei = el[ic];
        const double4 xi = x[site];
        const double alpha_i = al[ic], alpha_sq_i = alpha_i * alpha_i, mu_i = mu[ic];
        double cv_i = 1.0;
        for (int k = 0; k < C; ++k) {
            const double v = cc[site * C + k];
            if (v > 0.0) cv_i -= v;
        }
        const double zeta = P.zeta[ei], c1 = P.c1[ei], rsq_crd = P.rsq_crd[ei];
        const int cnt = nn[ic];
        for (int k = 0; k < cnt; ++k) {
            const uint32_t e = list[(size_t)k * ldl + ic];
            const int jc = (int)(e & kIndexMask);
            if (el[jc] != ei) continue;
            double dx, dy, dz;
            dpair_dx(xi, x[jc / C], e >> kIndexBits, bd.h, dx, dy, dz);
            const double rsq = dx * dx + (dy * dy + dz * dz);
            if (rsq >= rsq_crd) continue;
            const double alpha_j = al[jc], alpha_sq_j = alpha_j * alpha_j;
            const double gamma_i = rsq / (zeta * alpha_sq_i), gamma_j = rsq / (zeta * alpha_sq_j);
            const double mu_ji = P.beta * (mu[jc] - mu_i);
            double Q, gamma_ij_inv, theta_i;
            exchange_q(gamma_i, gamma_j, mu_ji, Q, gamma_ij_inv, theta_i);
            const double coef0 = -c1 * rsq * gamma_ij_inv * exp(-Q + 0.5 * mu_ji);
            const double theta_j = theta_i + 1.0;
            const int sj = jc / C;
            double cv_j = 1.0;
            for (int q = 0; q < C; ++q) {
                const double v = cc[sj * C + q];
                if (v > 0.0) cv_j -= v;
            }
            const double cj = cc[jc];
            const double d_i = exp(-0.5 * mu_ji) / (alpha_sq_i * alpha_i);
            const double d_j = exp(0.5 * mu_ji) / (alpha_sq_j * alpha_j);
            const double M0 = P.beta * (theta_i * ci * cv_j * d_i - theta_j * cj * cv_i * d_j);
            const size_t slot = (size_t)m * ldl + ic;
            J.idx2[slot] = (uint32_t)jc;
            J.f0[slot] = coef0 * M0;
            J.f1[slot] = coef0 * d_i * ci;
            J.f2[slot] = coef0 * d_j * cv_i;
            S0 += coef0 * M0;
            S1 += coef0 * d_i * cv_j;
            S2 += coef0 * d_j * cj;
            ++m;
        }
    }
    J.cnt2[ic] = m;
    J.S[(size_t)ic * 3] = S0;
    J.S[(size_t)ic * 3 + 1] = S1;
    J.S[(size_t)ic * 3 + 2] = S2;
}

// product, first sum: A1_i = c_i F''(rho_i) sum_j rho_{j->i} x_j (the change
// of F'(rho_i), times c_i); every component sums its own list (the
// reference's half list adds the j side of each pair in the same loop). The
// thread of a site's first component also writes xs of the site.
__global__ __launch_bounds__(kBlock) void k_dmd_ja1(int nc, int C, const double* __restrict__ cc, int ldl, JState J,
                                                    const double* __restrict__ xv, double* __restrict__ a1,
                                                    double* __restrict__ xs)
{
    const int ic = xcd_block() * kBlock + threadIdx.x;
    if (ic >= nc) return;
    const double ci = cc[ic];
    double s = 0.0;
    const int cnt = J.cnt[ic];
    for (int k = 0; k < cnt; ++k) {
        const size_t slot = (size_t)k * ldl + ic;
        s += J.r2[slot] * xv[J.idx[slot]];
    }
    a1[ic] = ci >= 0.0 ? s * (ci * J.ddE[ic]) : 0.0;
    if (ic % C == 0) {
        double t = 0.0;
        for (int q = 0; q < C; ++q)
            if (cc[ic + q] >= 0.0) t += xv[ic + q];
        xs[ic / C] = t;
    }
}

// product, second sum: b0_i = sum_j rho_{i->j} A1_j + t0 x_j, the change of
// the chemical potential mu_i
__global__ __launch_bounds__(kBlock) void k_dmd_jb0(int nc, int ldl, JState J, const double* __restrict__ a1,
                                                    const double* __restrict__ xv, double* __restrict__ b0)
{
    const int ic = xcd_block() * kBlock + threadIdx.x;
    if (ic >= nc) return;
    double s = 0.0;
    const int cnt = J.cnt[ic];
    for (int k = 0; k < cnt; ++k) {
        const size_t slot = (size_t)k * ldl + ic;
        const uint32_t jc = J.idx[slot];
        s += J.r1[slot] * a1[jc] + J.t0[slot] * xv[jc];
    }
    b0[ic] = s;
}

// product over the second list (src/ff_eam_dmd.cpp:1536-1572)
__global__ __launch_bounds__(kBlock) void k_dmd_jax(int nc, int C, int ldl, JState J, const double* __restrict__ b0,
                                                    const double* __restrict__ xv, const double* __restrict__ xs,
                                                    double* __restrict__ ax)
{
    const int ic = xcd_block() * kBlock + threadIdx.x;
    if (ic >= nc) return;
    double acc = 0.0;
    const int cnt = J.cnt2[ic];
    for (int k = 0; k < cnt; ++k) {
        const size_t slot = (size_t)k * ldl + ic;
        const uint32_t jc = J.idx2[slot];
        acc += J.f0[slot] * b0[jc] - J.f1[slot] * xs[jc / C] - J.f2[slot] * xv[jc];
    }
    const double* S = J.S + (size_t)ic * 3;
    ax[ic] = (S[1] * xv[ic] + S[2] * xs[ic / C] - S[0] * b0[ic]) + acc;
}

// ---------------- implicit concentration step ----------------

// F = c / beta + a - c_d, the residual of DAEImplicit::newton
// (src/dae_imp.cpp:148-161)
__global__ __launch_bounds__(kBlock) void k_newton_f(int nc, double beta_inv, const double* __restrict__ c,
                                                     const double* __restrict__ a, const double* __restrict__ cd,
                                                     double* __restrict__ F)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < nc) F[i] = beta_inv * c[i] + a[i] - cd[i];
}

// y = -y + x / beta: the Jacobian of F applied to x, from y = J x
// (src/dae_imp.cpp:163-172)
__global__ __launch_bounds__(kBlock) void k_newton_jac(int nc, double beta_inv, const double* __restrict__ x,
                                                       double* __restrict__ y)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < nc) y[i] = -y[i] + beta_inv * x[i];
}

// GMRES: partial sums of v . Q_j for j < nv (Q_j at Q + j * ldq)
__global__ __launch_bounds__(kBlock) void k_gm_dots(int nc, int nv, const double* __restrict__ v,
                                                    const double* __restrict__ Q, size_t ldq,
                                                    double* __restrict__ part)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[kPartW];
#pragma unroll
    for (int k = 0; k < kPartW; ++k) acc[k] = 0.0;
    if (i < nc) {
        const double vi = v[i];
#pragma unroll
        for (int k = 0; k < kPartW; ++k)
            if (k < nv) acc[k] = vi * Q[(size_t)k * ldq + i];
    }
    block_sum_store<kPartW>(acc, part + (size_t)blockIdx.x * kPartW);
}

// GMRES, one classical Gram-Schmidt pass (src/gmres.cpp:66-76):
// qn = v - sum_j h_j Q_j with h on the device; partial sums of qn^2
__global__ __launch_bounds__(kBlock) void k_gm_orth(int nc, int nv, const double* __restrict__ v,
                                                    const double* __restrict__ Q, size_t ldq,
                                                    const double* __restrict__ h, double* __restrict__ qn,
                                                    double* __restrict__ part)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < nc) {
        double q = v[i];
        for (int k = 0; k < nv; ++k) q -= h[k] * Q[(size_t)k * ldq + i];
        qn[i] = q;
        acc[0] = q * q;
    }
    block_sum_store<1>(acc, part + (size_t)blockIdx.x * kPartW);
}

__global__ __launch_bounds__(kBlock) void k_gm_scale(int nc, double s, const double* __restrict__ in,
                                                     double* __restrict__ out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < nc) out[i] = s * in[i];
}

struct GmY {
    double y[kPartW];
};

// x = sum_k y_k Q_k (src/gmres.cpp:183-189)
__global__ __launch_bounds__(kBlock) void k_gm_combine(int nc, int nv, GmY y, const double* __restrict__ Q,
                                                       size_t ldq, double* __restrict__ x)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nc) return;
    double s = 0.0;
    for (int k = 0; k < nv; ++k) s += y.y[k] * Q[(size_t)k * ldq + i];
    x[i] = s;
}

// update_c (src/dae_imp.cpp:89-117): the largest step fraction r <= 1 that
// keeps every present component in [0, 1], per component (the nextafter
// walk makes the bound hold in FP64), minimum per workgroup; absent
// components get no step
__global__ __launch_bounds__(kBlock) void k_newton_r(int nc, const double* __restrict__ c, double* __restrict__ del,
                                                     double* __restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double sh[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double r = 1.0;
    if (i < nc) {
        const double ci = c[i];
        if (ci >= 0.0) {
            const double di = del[i];
            const double tmp = ci - di;
            if (tmp > 1.0) {
                r = (ci - 1.0) / di;
                while (ci - r * di > 1.0) r = nextafter(r, 0.0);
            }
            if (tmp < 0.0) {
                r = ci / di;
                while (ci - r * di < 0.0) r = nextafter(r, 0.0);
            }
        } else {
            del[i] = 0.0;
        }
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * kPartW] = sh[0];
}

// c = (c - r del) rounded through 1 + (.) - 1, as src/dae_imp.cpp:119-124
__global__ __launch_bounds__(kBlock) void k_newton_apply(int nc, double r, const double* __restrict__ del,
                                                         double* __restrict__ c)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nc) return;
    const double v = c[i] - r * del[i];  // no contraction: one rounding for r del, one for the difference
    c[i] = (v + 1.0) - 1.0;              // strict IEEE: not folded to v
}

// chemical potential of new_derivative<true> (src/ff_dmd.cpp:305-366) and
// the sum of c * mu_ext
__global__ __launch_bounds__(kBlock) void k_dmd_chem(int n, int C, const double* __restrict__ al,
                                                     const double* __restrict__ cc, const int* __restrict__ el,
                                                     double kBT, DmdArgs A, const double* __restrict__ ext_mu,
                                                     double* __restrict__ mu, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        double cv = 1.0;
        for (int k = 0; k < C; ++k)
            if (cc[i * C + k] >= 0.0) cv -= cc[i * C + k];
        const double logcv = cv != 0.0 ? log(cv) : log(2.220446049250313e-16);
        for (int k = 0; k < C; ++k) {
            const int q = i * C + k;
            const double c = cc[q];
            if (c < 0.0) continue;
            const double logc = c != 0.0 ? log(c) : log(2.220446049250313e-16);
            const int e = el[q];
            mu[q] = -mu[q] - kBT * (logc - logcv - 3.0 * log(al[q])) - A.mu0[e] + ext_mu[e];
            acc[0] += c * ext_mu[e];
        }
    }
    block_sum_store<1>(acc, part + (size_t)blk * kPartW);
}

// DynamicDMD::decide (src/dynamic_dmd.cpp:103-126)
__global__ __launch_bounds__(kBlock) void k_dmd_decide(int n, int C, const double4* __restrict__ x,
                                                       const double4* __restrict__ x0, const double* __restrict__ al,
                                                       const double* __restrict__ al0, double scale, double half_skin,
                                                       int* __restrict__ flags)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 a = x0[i], b = x[i];
    const double dr = sqrt(rsq_exact(a.x, a.y, a.z, b.x, b.y, b.z));
    double da = al[i * C] - al0[i * C];
    for (int k = 0; k < C; ++k) da = fmax(da, al[i * C + k] - al0[i * C + k]);
    if (dr + da * scale > half_skin) atomicOr(&flags[1], 1);
}

// CG vectors after a rebuild inside a minimisation: slot p now holds the site
// that was at slot perm[p]. Directions and forces move with the site; the
// line-search origin x0 also takes the site's wrap into the box (the
// difference of its new and old position), so x0 + t h stays consistent with
// the new list's image codes.
__global__ __launch_bounds__(kBlock) void k_dcg_follow(int n, int C, const int* __restrict__ perm,
                                                       const double4* __restrict__ x_new,
                                                       const double4* __restrict__ x_old,
                                                       const double4* __restrict__ x0, const double4* __restrict__ h,
                                                       const double4* __restrict__ f0, const double* __restrict__ a0,
                                                       const double* __restrict__ ha, const double* __restrict__ fa0,
                                                       double4* __restrict__ x0_o, double4* __restrict__ h_o,
                                                       double4* __restrict__ f0_o, double* __restrict__ a0_o,
                                                       double* __restrict__ ha_o, double* __restrict__ fa0_o)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const int o = perm[p];
    const double4 xn = x_new[p], xo = x_old[o], b = x0[o];
    x0_o[p] = make_double4(b.x + (xn.x - xo.x), b.y + (xn.y - xo.y), b.z + (xn.z - xo.z), b.w);
    h_o[p] = h[o];
    f0_o[p] = f0[o];
    for (int k = 0; k < C; ++k) {
        a0_o[p * C + k] = a0[o * C + k];
        ha_o[p * C + k] = ha[o * C + k];
        fa0_o[p * C + k] = fa0[o * C + k];
    }
}

// ---------------- CG vector kernels ----------------

// partial dot products over the x part (slots) and the alpha part (components)
__global__ __launch_bounds__(kBlock) void k_dot2(int n, int nc, const double4* __restrict__ ax,
                                                 const double* __restrict__ aa, const double4* __restrict__ bx,
                                                 const double* __restrict__ ba, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    const int stride = gridDim.x * kBlock;
    for (int k = i; k < n; k += stride) {
        const double4 u = ax[k], v = bx[k];
        acc[0] += u.x * v.x + u.y * v.y + u.z * v.z;
    }
    for (int k = i; k < nc; k += stride) acc[0] += aa[k] * ba[k];
    block_sum_store<1>(acc, part + (size_t)blk * kPartW);
}

// max |h_x|, max |h_alpha| and min over h_alpha<0 of -alpha/h_alpha
__global__ __launch_bounds__(kBlock) void k_cg_limits(int n, int nc, const double4* __restrict__ hx,
                                                      const double* __restrict__ ha, const double* __restrict__ al,
                                                      const double* __restrict__ cc, double* __restrict__ part)
{
    // grid-stride partial max |h_x|, max |h_alpha|, min -alpha/h_alpha; one
    // row per workgroup (max and min are exact, so any order gives the same)
    __shared__ double sh[3][kBlock];
    double mx = 0.0, ma = 0.0, mn = 1e300;
    const int stride = gridDim.x * kBlock;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) {
        const double4 h = hx[k];
        mx = fmax(mx, fmax(fabs(h.x), fmax(fabs(h.y), fabs(h.z))));
    }
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nc; k += stride) {
        if (cc[k] < 0.0) continue;
        ma = fmax(ma, fabs(ha[k]));
        if (ha[k] < 0.0) mn = fmin(mn, -al[k] / ha[k]);
    }
    sh[0][threadIdx.x] = mx;
    sh[1][threadIdx.x] = ma;
    sh[2][threadIdx.x] = mn;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
            sh[2][threadIdx.x] = fmin(sh[2][threadIdx.x], sh[2][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) part[(size_t)blockIdx.x * kPartW + threadIdx.x] = sh[threadIdx.x][0];
}

// rows of partial extrema -> out: columns [0, nmax) by max, [nmax, nmax + nmin) by min
__global__ __launch_bounds__(kBlock) void k_extrema_rows(int nrows, int nmax, int nmin, const double* __restrict__ part,
                                                         double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    for (int c = 0; c < nmax + nmin; ++c) {
        const bool is_max = c < nmax;
        double v = is_max ? -1e300 : 1e300;
        for (int r = threadIdx.x; r < nrows; r += kBlock) {
            const double u = part[(size_t)r * kPartW + c];
            v = is_max ? fmax(v, u) : fmin(v, u);
        }
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + s])
                                         : fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = sh[0];
        __syncthreads();
    }
}

// x = x0 + t h (positions and alpha; absent components untouched)
__global__ __launch_bounds__(kBlock) void k_cg_move(int n, int nc, double t, const double4* __restrict__ x0,
                                                    const double4* __restrict__ hx, const double* __restrict__ a0,
                                                    const double* __restrict__ ha, const double* __restrict__ cc,
                                                    double4* __restrict__ x, double* __restrict__ al)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double4 a = x0[i], h = hx[i];
        x[i] = make_double4(a.x + t * h.x, a.y + t * h.y, a.z + t * h.z, 0.0);
    }
    if (i < nc && cc[i] >= 0.0) al[i] = a0[i] + t * ha[i];
}

// h = ratio h + f
__global__ __launch_bounds__(kBlock) void k_cg_dir(int n, int nc, double ratio, const double4* __restrict__ fx,
                                                   const double* __restrict__ fa, double4* __restrict__ hx,
                                                   double* __restrict__ ha)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double4 f = fx[i], h = hx[i];
        hx[i] = make_double4(ratio * h.x + f.x, ratio * h.y + f.y, ratio * h.z + f.z, 0.0);
    }
    if (i < nc) ha[i] = ratio * ha[i] + fa[i];
}

// h = (h + a u) g over the position and alpha parts (u optional)
__global__ __launch_bounds__(kBlock) void k_dlb_axpy(int n, int nc, double a, const double4* __restrict__ ux,
                                                     const double* __restrict__ ua, double g,
                                                     double4* __restrict__ hx, double* __restrict__ ha)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        double4 h = hx[i];
        if (ux) {
            const double4 u = ux[i];
            h = make_double4(h.x + a * u.x, h.y + a * u.y, h.z + a * u.z, 0.0);
        }
        hx[i] = make_double4(h.x * g, h.y * g, h.z * g, 0.0);
    }
    if (i < nc) ha[i] = (ua ? ha[i] + a * ua[i] : ha[i]) * g;
}

// the newest L-BFGS pair: s = t h, y = f0 - f over both parts
__global__ __launch_bounds__(kBlock) void k_dlb_pair(int n, int nc, double t, const double4* __restrict__ hx,
                                                     const double* __restrict__ ha, const double4* __restrict__ f0x,
                                                     const double* __restrict__ fa0, const double4* __restrict__ fx,
                                                     const double* __restrict__ fa, double4* __restrict__ sx,
                                                     double* __restrict__ sa, double4* __restrict__ yx,
                                                     double* __restrict__ ya)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double4 h = hx[i], g = f0x[i], f = fx[i];
        sx[i] = make_double4(t * h.x, t * h.y, t * h.z, 0.0);
        yx[i] = make_double4(g.x - f.x, g.y - f.y, g.z - f.z, 0.0);
    }
    if (i < nc) {
        sa[i] = t * ha[i];
        ya[i] = fa0[i] - fa[i];
    }
}

// an L-BFGS pair follows its sites into a rebuilt slot order
__global__ __launch_bounds__(kBlock) void k_dlb_follow(int n, int C, const int* __restrict__ perm,
                                                       const double4* __restrict__ sx, const double* __restrict__ sa,
                                                       double4* __restrict__ dx, double* __restrict__ da)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const int o = perm[p];
    dx[p] = sx[o];
    for (int k = 0; k < C; ++k) da[p * C + k] = sa[o * C + k];
}

// grid-stride partial max |f_x|, |f_alpha|: one row per workgroup
__global__ __launch_bounds__(kBlock) void k_gmax(int n, int nc, const double4* __restrict__ fx,
                                                 const double* __restrict__ fa, double* __restrict__ part)
{
    __shared__ double sh[kBlock];
    double m = 0.0;
    const int stride = gridDim.x * kBlock;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) {
        const double4 f = fx[k];
        m = fmax(m, fmax(fabs(f.x), fmax(fabs(f.y), fabs(f.z))));
    }
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nc; k += stride) m = fmax(m, fabs(fa[k]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * kPartW] = sh[0];
}

// ---------------- second derivatives over x and alpha: the Newton stage ----------------
//
// ForceFieldEAMDMD::__prepJ_n_res / __J (src/ff_eam_dmd.cpp:364-831) on the
// owner-computes component lists. The prep pass evaluates every listed pair
// once at the current state and keeps nine FP64 values per list slot
// (slot-major, [q][k * ldl + component]); each Jacobian product then reads
// them twice: the first pass with the densities' F'' fixed collects the
// change of F'(rho) per component, the second applies it. Vectors over
// (x, alpha) are flat in slot order: x of site s at [3 s, 3 s + 3), alpha of
// component k at [3 n + k].

constexpr int kXaQ = 9;  // fr, fa, frr, fra, faa, r2, a2, r1, a1 (see k_dmd_prepj)

// prep pass: per list slot of a present component i and its entry j,
//   fr  = D_r g0 + dE_i D_r g2 + dE_j D_r g1,   fa = the same with D_a,
//   frr, fra, faa: the same with D_rr, D_ra, D_aa,
//   r2 = D_r g2, a2 = D_a g2, r1 = D_r g1, a1 = D_a g1,
// with D_r = (1/r) d/dr, D_a = (1/a) d/da of the smeared g0 = r phi,
// g1 = r rho_ij, g2 = r rho_ji from the moments xi^0..xi^4 (the second
// derivatives as src/ff_eam_dmd.cpp:484-492 writes them); zero beyond the cut.
// It also writes F''(rho) per component (zero above rho_max), fcoef, and the
// residual f = the gradient per unit c (as calc_mu's f and f_alpha), with
// partial sums of f^2.
__global__ __launch_bounds__(kBlock) void k_dmd_prepj(int n, const double4* __restrict__ x,
                                                      const double* __restrict__ al, const double* __restrict__ cc,
                                                      const int* __restrict__ el, const int* __restrict__ nn,
                                                      const uint32_t* __restrict__ list, int ldl, DmdArgs A,
                                                      BoxDev bd, const double* __restrict__ dE,
                                                      double* __restrict__ slots, size_t lds,
                                                      double* __restrict__ ddE, double* __restrict__ fc,
                                                      double* __restrict__ f, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        const int C = A.C;
        const double4 xi = x[i];
        const double rho_max = A.rho_max;
        double c_tot = 0.0;
        for (int q = 0; q < C; ++q) {
            const double v = cc[i * C + q];
            if (v > 0.0) c_tot += v;
        }
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const size_t fa_off = (size_t)3 * n;
        for (int ic = 0; ic < C; ++ic) {
            const int ci_idx = i * C + ic;
            const double ci = cc[ci_idx];
            if (ci < 0.0) {
                ddE[ci_idx] = 0.0;
                fc[ci_idx] = 0.0;
                f[fa_off + ci_idx] = 0.0;
                continue;
            }
            const double fcoef = C == 1 ? 1.0 : (c_tot == 0.0 ? (ci == 0.0 ? 1.0 : 0.0) : ci / c_tot);
            const double ai = al[ci_idx];
            const int ei = el[ci_idx];
            const double dEi = dE[ci_idx];
            double rho = 0.0, gai = 0.0;
            const int cnt = nn[ci_idx];
            for (int k = 0; k < cnt; ++k) {
                const size_t slot = (size_t)k * ldl + ci_idx;
                const uint32_t e = list[slot];
                const int jc = (int)(e & kIndexMask);
                const int tt = ei * A.ne + el[jc];
                const double aj = al[jc], cj = cc[jc], dEj = dE[jc];
                const PairGeo pg = pair_geo(xi, x[jc / C], e >> kIndexBits, bd.h, ai, aj, A.xmax, A.cut[tt]);
                double q[kXaQ] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (pg.live) {
                    const double r = pg.r, a = pg.a;
                    double H[3][5];
#pragma unroll
                    for (int t = 0; t < 3; ++t)
#pragma unroll
                        for (int m = 0; m < 5; ++m) H[t][m] = 0.0;
                    for (int l = 0; l < A.N; ++l) {
                        const double xl = A.xi[l];
                        double p;
                        const double u = r - xl * a;
                        const int m = dknot(A, u, p);
                        const double s = u < 0.0 ? -A.wi[l] : A.wi[l];
                        const double4* row = A.tab + ((size_t)tt * A.nr + m) * 3;
#pragma unroll
                        for (int t = 0; t < 3; ++t) {
                            double w = hv(row[t], p) * s;
#pragma unroll
                            for (int mm = 0; mm < 5; ++mm) {
                                H[t][mm] += w;
                                w *= xl;
                            }
                        }
                    }
                    const double r_inv = 1.0 / r, r2_inv = r_inv * r_inv, rsq = r * r;
                    const double a_inv = 1.0 / a, a2_inv = a_inv * a_inv, a4_inv = a2_inv * a2_inv, asq = a * a;
                    double d1[3], d2[3], d3[3], d4[3], d5[3], g0[3];
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        const double rp = kPiInSq * r_inv * H[t][0];
                        const double dr = -r2_inv * (rp + 2.0 * kPiInSq * a_inv * H[t][1]);
                        const double da = -a2_inv * (rp - 2.0 * kPiInSq * r_inv * H[t][2]);
                        g0[t] = rp;
                        d1[t] = dr;
                        d2[t] = da;
                        d3[t] = -r2_inv * (3.0 * dr - 2.0 * da);
                        d4[t] = -r2_inv * a2_inv * (3.0 * rp + 3.0 * rsq * dr + asq * da + 4.0 * kPiInSq * a_inv * H[t][3]);
                        d5[t] = -a4_inv * (3.0 * rp + 6.0 * asq * da - 4.0 * kPiInSq * r_inv * H[t][4]);
                    }
                    q[0] = d1[0] + d1[2] * dEi + d1[1] * dEj;
                    q[1] = d2[0] + d2[2] * dEi + d2[1] * dEj;
                    q[2] = d3[0] + d3[2] * dEi + d3[1] * dEj;
                    q[3] = d4[0] + d4[2] * dEi + d4[1] * dEj;
                    q[4] = d5[0] + d5[2] * dEi + d5[1] * dEj;
                    q[5] = d1[2];
                    q[6] = d2[2];
                    q[7] = d1[1];
                    q[8] = d2[1];
                    rho += cj * g0[2];
                    const double gp = q[0] * cj * fcoef;
                    gx += gp * pg.dx;
                    gy += gp * pg.dy;
                    gz += gp * pg.dz;
                    gai += ai * q[1] * cj;
                }
#pragma unroll
                for (int t = 0; t < kXaQ; ++t) slots[(size_t)t * lds + slot] = q[t];
            }
            ddE[ci_idx] = rho > rho_max ? 0.0 : ddemb5(A, ei, rho);
            fc[ci_idx] = fcoef;
            const double fai = gai - 3.0 * A.kBT / ai;
            f[fa_off + ci_idx] = fai;
            acc[0] += fai * fai;
        }
        f[3 * (size_t)i] = gx;
        f[3 * (size_t)i + 1] = gy;
        f[3 * (size_t)i + 2] = gz;
        acc[0] += gx * gx + gy * gy + gz * gz;
    }
    block_sum_store<1>(acc, part + (size_t)blk * kPartW);
}

// the periodic difference x_i - x_j of a list entry (as pair_geo)
__device__ inline void xa_dx(const double4& xi, const double4& xj, uint32_t code, const double* __restrict__ h,
                             double& dx, double& dy, double& dz)
{
#pragma clang fp contract(off)
    dpair_dx(xi, xj, code, h, dx, dy, dz);
}

// J, first pass (src/ff_eam_dmd.cpp:691-756): the pair terms with F'(rho)
// fixed and, per component, the change of F'(rho): dde_i = F''(rho_i)
// sum_j c_j (A D_r g2 + B D_a g2), A = dx_ij . (dx_i - dx_j),
// B = alpha_i dalpha_i + alpha_j dalpha_j
__global__ __launch_bounds__(kBlock) void k_dmd_jxa1(int n, const double4* __restrict__ x,
                                                     const double* __restrict__ al, const double* __restrict__ cc,
                                                     const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                     int ldl, int C, BoxDev bd, const double* __restrict__ slots,
                                                     size_t lds, const double* __restrict__ ddE,
                                                     const double* __restrict__ fc, const double* __restrict__ in,
                                                     double* __restrict__ out, double* __restrict__ dde)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 xi = x[i];
    const size_t fa_off = (size_t)3 * n;
    const double dxi0 = in[3 * (size_t)i], dxi1 = in[3 * (size_t)i + 1], dxi2 = in[3 * (size_t)i + 2];
    double ox = 0.0, oy = 0.0, oz = 0.0;
    for (int ic = 0; ic < C; ++ic) {
        const int ci_idx = i * C + ic;
        const double ci = cc[ci_idx];
        if (ci < 0.0) {
            dde[ci_idx] = 0.0;
            out[fa_off + ci_idx] = 0.0;
            continue;
        }
        const double ai = al[ci_idx], dai = in[fa_off + ci_idx], fci = fc[ci_idx];
        double oa = 0.0, dd = 0.0;
        const int cnt = nn[ci_idx];
        for (int k = 0; k < cnt; ++k) {
            const size_t slot = (size_t)k * ldl + ci_idx;
            const uint32_t e = list[slot];
            const int jc = (int)(e & kIndexMask);
            const int js = jc / C;
            const double fr = slots[slot], fa = slots[lds + slot], frr = slots[2 * lds + slot];
            const double fra = slots[3 * lds + slot], faa = slots[4 * lds + slot];
            const double r2 = slots[5 * lds + slot], a2 = slots[6 * lds + slot];
            double dx, dy, dz;
            xa_dx(xi, x[js], e >> kIndexBits, bd.h, dx, dy, dz);
            const double ex = dxi0 - in[3 * (size_t)js], ey = dxi1 - in[3 * (size_t)js + 1],
                         ez = dxi2 - in[3 * (size_t)js + 2];
            const double aj = al[jc], daj = in[fa_off + jc], cj = cc[jc];
            const double Av = dx * ex + dy * ey + dz * ez;
            const double Bv = ai * dai + aj * daj;
            const double df = -(fra * Bv + frr * Av), fp = -fr;
            const double pa0 = -(fra * Av + faa * Bv), pa1 = -fa;
            dd += (Av * r2 + Bv * a2) * cj;
            const double w = cj * fci;
            ox += (df * dx + fp * ex) * w;
            oy += (df * dy + fp * ey) * w;
            oz += (df * dz + fp * ez) * w;
            oa += cj * (pa0 * ai + pa1 * dai);
        }
        dde[ci_idx] = dd * ddE[ci_idx];
        out[fa_off + ci_idx] = oa;
    }
    out[3 * (size_t)i] = ox;
    out[3 * (size_t)i + 1] = oy;
    out[3 * (size_t)i + 2] = oz;
}

// J, second pass (src/ff_eam_dmd.cpp:758-830): the pair terms of the change
// of F'(rho), then -3 kT dalpha / alpha^2 per present component
__global__ __launch_bounds__(kBlock) void k_dmd_jxa2(int n, const double4* __restrict__ x,
                                                     const double* __restrict__ al, const double* __restrict__ cc,
                                                     const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                     int ldl, int C, BoxDev bd, double kBT,
                                                     const double* __restrict__ slots, size_t lds,
                                                     const double* __restrict__ fc, const double* __restrict__ dde,
                                                     const double* __restrict__ in, double* __restrict__ out)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 xi = x[i];
    const size_t fa_off = (size_t)3 * n;
    double ox = 0.0, oy = 0.0, oz = 0.0;
    for (int ic = 0; ic < C; ++ic) {
        const int ci_idx = i * C + ic;
        const double ci = cc[ci_idx];
        if (ci < 0.0) continue;
        const double ai = al[ci_idx], dai = in[fa_off + ci_idx], fci = fc[ci_idx], ddi = dde[ci_idx];
        double oa = 0.0;
        const int cnt = nn[ci_idx];
        for (int k = 0; k < cnt; ++k) {
            const size_t slot = (size_t)k * ldl + ci_idx;
            const uint32_t e = list[slot];
            const int jc = (int)(e & kIndexMask);
            const double r2 = slots[5 * lds + slot], a2 = slots[6 * lds + slot];
            const double r1 = slots[7 * lds + slot], a1 = slots[8 * lds + slot];
            const double ddj = dde[jc], cj = cc[jc];
            double dx, dy, dz;
            xa_dx(xi, x[jc / C], e >> kIndexBits, bd.h, dx, dy, dz);
            const double df = -(r2 * ddi + r1 * ddj) * cj * fci;
            ox += df * dx;
            oy += df * dy;
            oz += df * dz;
            oa += ai * -(a2 * ddi + a1 * ddj) * cj;
        }
        out[fa_off + ci_idx] += oa - 3.0 * kBT * dai / (ai * ai);
    }
    out[3 * (size_t)i] += ox;
    out[3 * (size_t)i + 1] += oy;
    out[3 * (size_t)i + 2] += oz;
}

// (x, alpha) in caller order -> one flat slot-order vector, and back
__global__ __launch_bounds__(kBlock) void k_xa_scatter(int n, int C, const int* __restrict__ slot_of,
                                                       const double* __restrict__ xh, const double* __restrict__ ah,
                                                       double* __restrict__ v)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    for (int d = 0; d < 3; ++d) v[3 * (size_t)s + d] = xh[3 * (size_t)id + d];
    for (int k = 0; k < C; ++k) v[(size_t)3 * n + (size_t)s * C + k] = ah[(size_t)id * C + k];
}

__global__ __launch_bounds__(kBlock) void k_xa_gather(int n, int C, const int* __restrict__ slot_of,
                                                      const double* __restrict__ v, double* __restrict__ xh,
                                                      double* __restrict__ ah)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    for (int d = 0; d < 3; ++d) xh[3 * (size_t)id + d] = v[3 * (size_t)s + d];
    for (int k = 0; k < C; ++k) ah[(size_t)id * C + k] = v[(size_t)3 * n + (size_t)s * C + k];
}

// MinDMDHandler::max_alpha_lcl_alpha (src/min_cg_dmd.h:347-360) with
// max_dalpha infinite: the largest r <= 1 with alpha + r h_alpha >= 0 for
// every component, minimum per workgroup into part[blk * kPartW]
__global__ __launch_bounds__(kBlock) void k_xa_limit(int nc, const double* __restrict__ al,
                                                     const double* __restrict__ ha, double* __restrict__ part)
{
    __shared__ double sh[kBlock];
    const int k = blockIdx.x * kBlock + threadIdx.x;
    double r = 1.0;
    if (k < nc) {
        const double h = ha[k];
        if (h < 0.0) r = fmin(-al[k] / h, r);
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * kPartW] = sh[0];
}

// x += t h_x, alpha += t h_alpha (the state in slot order)
__global__ __launch_bounds__(kBlock) void k_xa_apply(int n, int nc, double t, const double* __restrict__ h,
                                                     double4* __restrict__ x, double* __restrict__ al)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        double4 v = x[k];
        v.x += t * h[3 * (size_t)k];
        v.y += t * h[3 * (size_t)k + 1];
        v.z += t * h[3 * (size_t)k + 2];
        x[k] = v;
    }
    if (k < nc) al[k] += t * h[(size_t)3 * n + k];
}

// host <-> slot order
__global__ __launch_bounds__(kBlock) void k_dscatter(int n, int C, const double* __restrict__ xh,
                                                     const double* __restrict__ ah, const double* __restrict__ ch,
                                                     const int* __restrict__ slot_of, double4* __restrict__ x,
                                                     double* __restrict__ al, double* __restrict__ cc)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    x[s] = make_double4(xh[3 * (size_t)id], xh[3 * (size_t)id + 1], xh[3 * (size_t)id + 2], 0.0);
    for (int k = 0; k < C; ++k) {
        al[s * C + k] = ah[(size_t)id * C + k];
        cc[s * C + k] = ch[(size_t)id * C + k];
    }
}

__global__ __launch_bounds__(kBlock) void k_dgather(int n, int C, const int* __restrict__ slot_of,
                                                    const double4* __restrict__ x, const double* __restrict__ a,
                                                    const double* __restrict__ b, double* __restrict__ xo,
                                                    double* __restrict__ ao, double* __restrict__ bo)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    const double4 v = x[s];
    xo[3 * (size_t)id] = v.x;
    xo[3 * (size_t)id + 1] = v.y;
    xo[3 * (size_t)id + 2] = v.z;
    for (int k = 0; k < C; ++k) {
        if (ao) ao[(size_t)id * C + k] = a[s * C + k];
        if (bo) bo[(size_t)id * C + k] = b[s * C + k];
    }
}

// per-component array from slot order to caller order
__global__ __launch_bounds__(kBlock) void k_dgather_comp(int n, int C, const int* __restrict__ slot_of,
                                                         const double* __restrict__ a, double* __restrict__ ao)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    for (int k = 0; k < C; ++k) ao[(size_t)id * C + k] = a[s * C + k];
}

// per-component array from caller order to slot order
__global__ __launch_bounds__(kBlock) void k_dscatter_comp(int n, int C, const int* __restrict__ slot_of,
                                                          const double* __restrict__ a, double* __restrict__ ao)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    for (int k = 0; k < C; ++k) ao[s * C + k] = a[(size_t)id * C + k];
}

__global__ __launch_bounds__(kBlock) void k_smear_points(DmdArgs A, int tt, int npts, const double* __restrict__ r,
                                                         const double* __restrict__ a, double* __restrict__ out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npts) return;
    double rp[3], dr[3], da[3];
    smear_quad<true>(A, tt, r[i], a[i], rp, dr, da);
    for (int k = 0; k < 3; ++k) {
        out[(size_t)i * 9 + k] = rp[k];
        out[(size_t)i * 9 + 3 + k] = dr[k];
        out[(size_t)i * 9 + 6 + k] = da[k];
    }
}

// Surrogate requests of every pair the passes will evaluate, in the order the
// passes meet them (SurrCursor). COUNT: the number of requests per (network,
// site) and, in the last segment, of pairs left to the quadrature; else: the
// requests themselves at the scanned offsets.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_surr_stage(int n, const double4* __restrict__ x,
                                                       const double* __restrict__ al, const double* __restrict__ cc,
                                                       const int* __restrict__ el, const int* __restrict__ nn,
                                                       const uint32_t* __restrict__ list, int ldl, DmdArgs A,
                                                       BoxDev bd, SurrBox box, int* __restrict__ cnt,
                                                       const int* __restrict__ offs, SurrRequest* __restrict__ req)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int C = A.C, ne = A.ne, nt = ne * ne;
    int pos[kMaxSurrNets + 1];  // requests per network, then the slots left to the quadrature
    for (int t = 0; t <= nt; ++t) pos[t] = COUNT ? 0 : offs[(size_t)t * n + i];
    const double sr = 2.0 / ((double)box.r_hi - box.r_lo), sa = 2.0 / ((double)box.a_hi - box.a_lo);
    const double4 xi = x[i];
    for (int ic = 0; ic < C; ++ic) {
        const int ci_idx = i * C + ic;
        if (cc[ci_idx] < 0.0) continue;
        const double ai = al[ci_idx];
        const int ei = el[ci_idx];
        const int cntk = nn[ci_idx];
        for (int k = 0; k < cntk; ++k) {
            const uint32_t e = list[(size_t)k * ldl + ci_idx];
            const int jc = (int)(e & kIndexMask);
            const double4 xj = x[jc / C];
            const double aj = al[jc];
            const int ej = el[jc];
            const PairGeo pg = pair_geo(xi, xj, e >> kIndexBits, bd.h, ai, aj, A.xmax, A.cut[ei * ne + ej]);
            if (!pg.live) continue;  // the passes skip it too
            const double r = pg.r;
            if (!surrogate_in_box(box, r, ai, aj)) {
                ++pos[nt];
                continue;
            }
            const float xr = (float)((r - box.r_lo) * sr - 1.0);
            const float xa = (float)((ai - box.a_lo) * sa - 1.0), xb = (float)((aj - box.a_lo) * sa - 1.0);
            if (ei == ej) {
                const int t = ei * ne + ei;
                if (!COUNT) req[pos[t]] = SurrRequest{xr, xa, xb, 0.0f};
                ++pos[t];
            } else {
                const int t_in = ej * ne + ei, t_out = ei * ne + ej;
                if (!COUNT) {
                    req[pos[t_in]] = SurrRequest{xr, xb, xa, 0.0f};
                    req[pos[t_out]] = SurrRequest{xr, xa, xb, 0.0f};
                }
                ++pos[t_in];
                ++pos[t_out];
            }
        }
    }
    if (COUNT)
        for (int t = 0; t <= nt; ++t) cnt[(size_t)t * n + i] = pos[t];
}

}  // namespace

// ---------------- DMDSystem ----------------

// physicists' Gauss-Hermite nodes/weights by Newton on H_n(x) with the
// three-term recurrence (Golub-Welsch is not needed for n <= 16)
void gauss_hermite(int n, double* x, double* w)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        // initial guesses (Numerical Recipes style asymptotics)
        double z;
        if (i == 0) z = std::sqrt(2.0 * n + 1.0) - 1.85575 * std::pow(2.0 * n + 1.0, -0.16667);
        else if (i == 1) z = x[n - 1] - 1.14 * std::pow((double)n, 0.426) / x[n - 1];
        else if (i == 2) z = 1.86 * x[n - 2] - 0.86 * x[n - 1];
        else if (i == 3) z = 1.91 * x[n - 3] - 0.91 * x[n - 2];
        else z = 2.0 * x[n - i] - x[n - i + 1];
        double pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            // orthonormal recurrence: p_k = sqrt(2/k) z p_{k-1} - sqrt((k-1)/k) p_{k-2}
            double p1 = std::pow(pi, -0.25), p2 = 0.0;
            for (int k = 1; k <= n; ++k) {
                const double p3 = p2;
                p2 = p1;
                p1 = z * std::sqrt(2.0 / k) * p2 - std::sqrt((k - 1.0) / k) * p3;
            }
            pp = std::sqrt(2.0 * n) * p2;
            const double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) <= 1e-15 * std::max(1.0, std::fabs(z))) break;
        }
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = w[i] = 2.0 / (pp * pp);
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}


DMDSystem::DMDSystem(std::shared_ptr<EamForceField> ff, int nsites, int c_dim, int nquad, const double* H,
                     const double* x, const uint8_t* elem, const double* alpha, const double* c, double skin,
                     double kB, double T, double hP, const double* masses)
    : ff_(std::move(ff)), n_(nsites), C_(c_dim), nc_(nsites * c_dim), skin_(skin)
{
    const EamTables& t = ff_->tables();
    MAPP_REQUIRE(t.dmd, ERR_ARG, "the DMD system needs a force field built with dmd=1 (quartic F, r*rho tables)");
    MAPP_REQUIRE(nsites >= 1 && c_dim >= 1 && c_dim <= kMaxComp, ERR_ARG, "1 to 4 components per site");
    MAPP_REQUIRE(nquad >= 1 && nquad <= kMaxQuad, ERR_ARG, "1 to 16 Hermite points");
    MAPP_REQUIRE((uint64_t)nsites * c_dim <= kIndexMask, ERR_CAPACITY, "too many components");
    MAPP_REQUIRE(kB > 0.0 && T > 0.0 && hP > 0.0, ERR_ARG, "kB, T and hP must be positive");
    elem_host_.assign(elem, elem + (size_t)nsites * c_dim);
    for (int k = 0; k < nsites * c_dim; ++k) {
        MAPP_REQUIRE(elem[k] < t.ne, ERR_ARG, "element index out of range");
        MAPP_REQUIRE(c[k] < 0.0 || alpha[k] > 0.0, ERR_ARG, "present components need alpha > 0");
    }
    MAPP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    box_ = make_box(H);
    // quadrature and thermodynamic constants (src/ff_dmd.cpp:157-204, src/atoms_dmd.cpp:19)
    A_.ne = t.ne;
    A_.nr = t.nr;
    A_.nrho = t.nrho;
    A_.N = nquad;
    A_.C = c_dim;
    A_.dr_inv = 1.0 / t.dr;
    A_.drho_inv = 1.0 / t.drho;
    A_.rho_max = (double)t.nrho * t.drho;
    A_.kBT = kB * T;
    gauss_hermite(nquad, A_.xi, A_.wi);
    A_.xmax = A_.xi[nquad - 1];
    scl_ = A_.xmax;
    for (int e = 0; e < t.ne; ++e) {
        const double lambda = hP / std::sqrt(2.0 * M_PI * A_.kBT * masses[e]);
        A_.mu0[e] = A_.kBT * 1.5 * (2.0 * std::log(lambda) - 1.0 - std::log(M_PI));
    }
    masses_.assign(masses, masses + t.ne);
    for (int k = 0; k < t.ne * t.ne; ++k) A_.cut[k] = t.rc[k];
    // FP64 device tables
    {
        const int ne = t.ne, nr = t.nr;
        std::vector<double4> T4((size_t)ne * ne * nr * 3);
        auto d4 = [](const double* c) { return make_double4(c[0], c[1], c[2], c[3]); };
        for (int a = 0; a < ne; ++a)
            for (int b = 0; b < ne; ++b)
                for (int m = 0; m < nr; ++m) {
                    const size_t base = (((size_t)a * ne + b) * nr + m) * 3;
                    T4[base] = d4(t.rphi_at(a, b) + (size_t)m * 4);
                    T4[base + 1] = d4(t.rho_at(a, b) + (size_t)m * 4);
                    T4[base + 2] = d4(t.rho_at(b, a) + (size_t)m * 4);
                }
        tab_.reserve(T4.size());
        MAPP_HIP(hipMemcpy(tab_.get(), T4.data(), T4.size() * sizeof(double4), hipMemcpyHostToDevice));
        std::vector<float4> Tf(T4.size());
        for (size_t k = 0; k < T4.size(); ++k)
            Tf[k] = make_float4((float)T4[k].x, (float)T4[k].y, (float)T4[k].z, (float)T4[k].w);
        tabf_.reserve(Tf.size());
        MAPP_HIP(hipMemcpy(tabf_.get(), Tf.data(), Tf.size() * sizeof(float4), hipMemcpyHostToDevice));
        A_.tabf = tabf_.get();
        F5_.reserve(t.F.size());
        MAPP_HIP(hipMemcpy(F5_.get(), t.F.data(), t.F.size() * sizeof(double), hipMemcpyHostToDevice));
        A_.tab = tab_.get();
        A_.F5 = F5_.get();
    }
    plan_knots();
    double max_alpha = 0.0;
    for (int k = 0; k < nsites * c_dim; ++k)
        if (c[k] != -1.0) max_alpha = std::max(max_alpha, alpha[k]);
    max_cut_ = t.max_rc() + skin_ + scl_ * std::sqrt(2.0) * max_alpha;
    setup_grid();
    x_.reserve(n_); x_alt_.reserve(n_); s_.reserve(n_); x0_.reserve(n_); s_tmp_.reserve(n_);
    al_.reserve(nc_); al_alt_.reserve(nc_); c_.reserve(nc_); c_alt_.reserve(nc_); al0_.reserve(nc_);
    el_.reserve(nc_); el_alt_.reserve(nc_);
    id_.reserve(n_); id_alt_.reserve(n_); slot_of_.reserve(n_);
    f_.reserve(n_); fa_.reserve(nc_); mu_.reserve(nc_); E_.reserve(nc_); dE_.reserve(nc_);
    keys_.reserve(n_); keys_alt_.reserve(n_); vals_.reserve(n_); vals_alt_.reserve(n_);
    nn_.reserve(nc_);
    flags_.reserve(4);
    part_.reserve((size_t)2 * grid_for(std::max(n_, nc_)) * kPartW + 64);
    red_.reserve(64);
    red2_.reserve(256 * kPartW);
    std::vector<double4> xh(n_);
    std::vector<int> idh(n_), elh(nc_);
    for (int i = 0; i < n_; ++i) {
        xh[i] = make_double4(x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], 0.0);
        idh[i] = i;
    }
    for (int k = 0; k < nc_; ++k) elh[k] = elem[k];
    MAPP_HIP(hipMemcpyAsync(x_.get(), xh.data(), sizeof(double4) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(id_.get(), idh.data(), sizeof(int) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(al_.get(), alpha, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(c_.get(), c, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(el_.get(), elh.data(), sizeof(int) * nc_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    rebuild();
}

DMDSystem::~DMDSystem()
{
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void DMDSystem::setup_grid()
{
    int ncells = 1;
    for (int d = 0; d < 3; ++d) {
        const double cs = max_cut_ * box_.depth_inv[d];
        grid_.cut_s[d] = cs;
        int nc = (int)(1.0 / cs);
        if (nc < 1) nc = 1;
        MAPP_REQUIRE(nc <= 1024, ERR_CAPACITY, "at most 1024 cells per dimension");
        grid_.nc[d] = nc;
        int R = (int)std::ceil(cs * nc - 1e-12);
        if (R < 1) R = 1;
        MAPP_REQUIRE(R <= 2, ERR_CAPACITY, "box narrower than half the list cutoff; replicate the cell first");
        grid_.R[d] = R;
        ncells *= nc;
    }
    grid_.ncells = ncells;
    std::vector<std::pair<uint32_t, int>> code(ncells);
    for (int c2 = 0, c = 0; c2 < grid_.nc[2]; ++c2)
        for (int c1 = 0; c1 < grid_.nc[1]; ++c1)
            for (int c0 = 0; c0 < grid_.nc[0]; ++c0, ++c) code[c] = {morton3(c0, c1, c2), c};
    std::sort(code.begin(), code.end());
    std::vector<int> rank(ncells);
    for (int r = 0; r < ncells; ++r) rank[code[r].second] = r;
    cell_rank_.reserve(ncells);
    cell_start_.reserve(ncells + 1);
    MAPP_HIP(hipMemcpy(cell_rank_.get(), rank.data(), sizeof(int) * ncells, hipMemcpyHostToDevice));
    const double dens = n_ / box_.vol;
    const double sphere = 4.0 / 3.0 * M_PI * max_cut_ * max_cut_ * max_cut_;
    maxn_ = std::max(maxn_, (int)(1.25 * dens * sphere * C_) + 16);
}

void DMDSystem::reduce_rows(const double* part, int nrows, int ncols, double* out)
{
    const int g = std::min(256, std::max(1, (nrows + kBlock - 1) / kBlock));
    const int chunk = (nrows + g - 1) / g;
    hipLaunchKernelGGL(k_reduce_rows, dim3(g), dim3(kBlock), 0, stream_, nrows, ncols, part, red2_.get(), chunk);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kBlock), 0, stream_, g, ncols, red2_.get(), out, g);
}

void DMDSystem::rebuild()
{
    ++ver_;
    const int nb = grid_for(n_);
    DGrid g{};
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = grid_.nc[d];
        g.R[d] = grid_.R[d];
    }
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    prof_.begin(K_CELL_SORT, stream_);
    hipLaunchKernelGGL(k_dcell_keys, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), bd, g, cell_rank_.get(),
                       s_tmp_.get(), keys_.get(), vals_.get());
    int end_bit = 1;
    while ((1 << end_bit) < grid_.ncells + 1) ++end_bit;
    size_t bytes = 0;
    MAPP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_.get(), keys_alt_.get(), vals_.get(),
                                                vals_alt_.get(), n_, 0, end_bit, stream_));
    if (bytes > sort_tmp_bytes_) {
        sort_tmp_.reserve(bytes);
        sort_tmp_bytes_ = bytes;
    }
    MAPP_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp_.get(), bytes, keys_.get(), keys_alt_.get(), vals_.get(),
                                                vals_alt_.get(), n_, 0, end_bit, stream_));
    hipLaunchKernelGGL(k_dpermute, dim3(nb), dim3(kBlock), 0, stream_, n_, C_, vals_alt_.get(), s_tmp_.get(), bd,
                       al_.get(), c_.get(), el_.get(), id_.get(), x_alt_.get(), s_.get(), x0_.get(), al_alt_.get(),
                       c_alt_.get(), el_alt_.get(), al0_.get(), id_alt_.get(), slot_of_.get());
    std::swap(x_.p, x_alt_.p);
    std::swap(al_.p, al_alt_.p);
    std::swap(c_.p, c_alt_.p);
    std::swap(el_.p, el_alt_.p);
    std::swap(id_.p, id_alt_.p);
    hipLaunchKernelGGL(k_dcell_start, dim3((grid_.ncells + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_,
                       grid_.ncells, keys_alt_.get(), cell_start_.get());
    prof_.end(stream_);
    build_list();
    ++nrebuild_;
}

void DMDSystem::build_list()
{
    const int nb = grid_for(n_);
    DGrid g{};
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = grid_.nc[d];
        g.R[d] = grid_.R[d];
    }
    const EamTables& t = ff_->tables();
    g.ne = t.ne;
    g.C = C_;
    g.scl = scl_;
    for (int k = 0; k < t.ne * t.ne; ++k) g.cut_sk[k] = t.rc[k] + skin_;
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    ldl_ = nc_;
    for (int attempt = 0; attempt < 2; ++attempt) {
        list_.reserve((size_t)maxn_ * ldl_);
        MAPP_HIP(hipMemsetAsync(flags_.get(), 0, sizeof(int), stream_));
        prof_.begin(K_BUILD_LIST, stream_);
        hipLaunchKernelGGL(k_dbuild_list, dim3(nb), dim3(kBlock), 0, stream_, n_, s_.get(), bd, g, al_.get(),
                           c_.get(), el_.get(), cell_rank_.get(), cell_start_.get(), maxn_, ldl_, list_.get(),
                           nn_.get(), flags_.get());
        prof_.end(stream_);
        int over = 0;
        MAPP_HIP(hipMemcpyAsync(&over, flags_.get(), sizeof(int), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        if (over <= maxn_) return;
        maxn_ = over + over / 8 + 8;
    }
    throw Error(ERR_STATE, "DMD neighbor list capacity did not converge");
}

bool DMDSystem::decide()
{
    const int nb = grid_for(n_);
    MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
    hipLaunchKernelGGL(k_dmd_decide, dim3(nb), dim3(kBlock), 0, stream_, n_, C_, x_.get(), x0_.get(), al_.get(),
                       al0_.get(), scl_, 0.5 * skin_, flags_.get());
    int flag = 0;
    MAPP_HIP(hipMemcpyAsync(&flag, flags_.get() + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    return flag == 0;
}

bool DMDSystem::update()
{
    if (decide()) return false;
    rebuild();
    return true;
}

void DMDSystem::use_surrogate(std::shared_ptr<Surrogate> s)
{
    if (s) MAPP_REQUIRE(s->nnets() == A_.ne * A_.ne, ERR_ARG, "the surrogate was built for another element count");
    MAPP_REQUIRE(A_.ne * A_.ne <= kMaxSurrNets, ERR_CAPACITY, "too many elements for the surrogate");
    surr_ = std::move(s);
}

void DMDSystem::surrogate_coverage(int64_t* by_net, int64_t* by_quadrature)
{
    if (by_net) *by_net = s_net_;
    if (by_quadrature) *by_quadrature = s_quad_;
}

// Requests for every slot, scanned into one array grouped by network, then
// one inference launch per network writing the per-slot outputs.
void DMDSystem::stage_surrogate()
{
    const int nt = A_.ne * A_.ne;
    MAPP_REQUIRE((size_t)ldl_ * maxn_ < (1ull << 30), ERR_CAPACITY, "too many list slots for 32-bit request offsets");
    const size_t m = (size_t)(nt + 1) * n_;
    s_cnt_.reserve(m);
    s_off_.reserve(m);
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    const SurrBox sb = surr_->box();
    const int nb = (n_ + kBlock - 1) / kBlock;
    prof_.begin(K_SURR_INFER, stream_);
    hipLaunchKernelGGL(k_surr_stage<true>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(),
                       el_.get(), nn_.get(), list_.get(), ldl_, A_, bd, sb, s_cnt_.get(), s_off_.get(),
                       s_req_.get());
    size_t bytes = 0;
    MAPP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, s_cnt_.get(), s_off_.get(), (int)m, stream_));
    if (bytes > s_scan_bytes_) {
        s_scan_tmp_.reserve(bytes);
        s_scan_bytes_ = bytes;
    }
    MAPP_HIP(hipcub::DeviceScan::ExclusiveSum(s_scan_tmp_.get(), bytes, s_cnt_.get(), s_off_.get(), (int)m,
                                              stream_));
    std::vector<int> first(nt + 1), last_off(nt + 1), last_cnt(nt + 1);
    for (int t = 0; t <= nt; ++t) {
        MAPP_HIP(hipMemcpyAsync(&first[t], s_off_.get() + (size_t)t * n_, sizeof(int), hipMemcpyDeviceToHost,
                                stream_));
        MAPP_HIP(hipMemcpyAsync(&last_off[t], s_off_.get() + (size_t)t * n_ + n_ - 1, sizeof(int),
                                hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipMemcpyAsync(&last_cnt[t], s_cnt_.get() + (size_t)t * n_ + n_ - 1, sizeof(int),
                                hipMemcpyDeviceToHost, stream_));
    }
    MAPP_HIP(hipStreamSynchronize(stream_));
    const int total = last_off[nt - 1] + last_cnt[nt - 1];  // requests end where the quadrature count starts
    s_req_.reserve(std::max(1, total));
    s_y_.reserve((size_t)std::max(1, total) * kSurrOut);
    hipLaunchKernelGGL(k_surr_stage<false>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(),
                       el_.get(), nn_.get(), list_.get(), ldl_, A_, bd, sb, s_cnt_.get(), s_off_.get(),
                       s_req_.get());
    s_net_ = 0;
    for (int t = 0; t < nt; ++t) {
        const int cnt = last_off[t] + last_cnt[t] - first[t];
        s_net_ += cnt;
        surr_->infer_dev(t, cnt, s_req_.get() + first[t], s_y_.get() + (size_t)first[t] * kSurrOut, stream_);
    }
    s_quad_ = last_off[nt] + last_cnt[nt] - first[nt];
    prof_.end(stream_);
}

// Workgroup sizes of the LDS passes: all waves of the one workgroup per CU
// fit the register file without spills (pass 1 of two entries fits 128
// VGPRs, of four 256). Pass 2 takes two entries in 512 threads: 768 would
// fit its registers too, but a persistent walk is as long as its busiest
// workgroup, and 262,144 sites over 256 x 768 threads leave a third of the
// workgroups a second chunk (measured 1.49 against 1.13 ms); one entry in
// 1024 threads measured 1.42 ms, four entries spill.
constexpr int knot_nt1(int B) { return B == 4 ? 512 : 1024; }
constexpr int kKnotNT2 = 512;

void DMDSystem::pass1(bool energy_only, int f32)
{
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    if (f32 == 1 && knots_ok_ && !surr_ && batch_ != 1) {
        const int nb = std::min(knot_grid_, grid_for(n_));
        ac_.reserve(nc_);
        ed_.reserve(nc_);
        prof_.begin(energy_only ? K_DMD_ENERGY : K_DMD_PASS1, stream_);
        hipLaunchKernelGGL(k_pack_ac, dim3((nc_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, nc_, al_.get(),
                           c_.get(), ac_.get());
        if (batch_ == 4)
            hipLaunchKernelGGL((k_dmd_pass1k<4, knot_nt1(4)>), dim3(nb), dim3(knot_nt1(4)), knot_lds_, stream_, n_,
                               x_.get(), al_.get(), c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, A_, bd,
                               kn_, ac_.get(), ed_.get(), energy_only ? 1 : 0, E_.get(), dE_.get(), part_.get());
        else
            hipLaunchKernelGGL((k_dmd_pass1k<2, knot_nt1(2)>), dim3(nb), dim3(knot_nt1(2)), knot_lds_, stream_, n_,
                               x_.get(), al_.get(), c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, A_, bd,
                               kn_, ac_.get(), ed_.get(), energy_only ? 1 : 0, E_.get(), dE_.get(), part_.get());
        prof_.end(stream_);
        reduce_rows(part_.get(), nb, 2, red_.get());
        ed_ver_ = ver_;
        return;
    }
    ed_ver_ = -1;
    const int nb = grid_for(n_);
    SurrogateArgs S{};
    if (surr_) {
        stage_surrogate();
        S = SurrogateArgs{s_y_.get(), s_off_.get(), n_, surr_->box()};
    }
    prof_.begin(energy_only ? K_DMD_ENERGY : K_DMD_PASS1, stream_);
    if (surr_)
        hipLaunchKernelGGL(k_dmd_pass1<true>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(),
                           el_.get(), nn_.get(), list_.get(), ldl_, A_, bd, S, energy_only ? 1 : 0, E_.get(),
                           dE_.get(), part_.get());
    else if (batch_ == 1)
        hipLaunchKernelGGL(k_dmd_pass1<false>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(),
                           el_.get(), nn_.get(), list_.get(), ldl_, A_, bd, S, energy_only ? 1 : 0, E_.get(),
                           dE_.get(), part_.get());
    else {
#define MAPP_P1(B, F) hipLaunchKernelGGL((k_dmd_pass1q<B, F>), dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), \
                                         al_.get(), c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, A_, bd, \
                                         energy_only ? 1 : 0, E_.get(), dE_.get(), part_.get())
        if (batch_ == 2) {
            if (f32) MAPP_P1(2, true); else MAPP_P1(2, false);
        } else {
            if (f32) MAPP_P1(4, true); else MAPP_P1(4, false);
        }
#undef MAPP_P1
    }
    prof_.end(stream_);
    reduce_rows(part_.get(), nb, 2, red_.get());
}

void DMDSystem::pass2(bool unit, int f32)
{
    const int nb = grid_for(n_);
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    double* p2 = part_.get() + (size_t)nb * kPartW;
    if (f32 == 1 && knots_ok_ && !surr_ && !unit && batch_ != 1) {
        // the packed records of pass 1 of this state (pass 2 follows pass 1)
        MAPP_REQUIRE(ed_ver_ == ver_, ERR_STATE, "the FP32-knot pass 2 needs pass 1 of the same state");
        const int nbk = std::min(knot_grid_, nb);
        prof_.begin(K_DMD_PASS2, stream_);
        hipLaunchKernelGGL((k_dmd_pass2k<2, kKnotNT2>), dim3(nbk), dim3(kKnotNT2), knot_lds_, stream_, n_, x_.get(),
                           al_.get(), c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, A_, bd, kn_, ac_.get(),
                           ed_.get(), E_.get(), dE_.get(), f_.get(), fa_.get(), mu_.get(), p2);
        prof_.end(stream_);
        reduce_rows(p2, nbk, 6, red_.get() + 8);
        return;
    }
    SurrogateArgs S{};
    if (surr_) S = SurrogateArgs{s_y_.get(), s_off_.get(), n_, surr_->box()};
    prof_.begin(K_DMD_PASS2, stream_);
    if (surr_)
        hipLaunchKernelGGL(k_dmd_pass2<true>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(),
                           el_.get(), nn_.get(), list_.get(), ldl_, A_, bd, S, E_.get(), dE_.get(), f_.get(),
                           fa_.get(), mu_.get(), p2);
    else if (unit) {
        gu_.reserve(n_);
        gau_.reserve(nc_);
        j_reserve();
        hipLaunchKernelGGL((k_dmd_pass2q<4, true, false>), dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(),
                           c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, A_, bd, E_.get(), dE_.get(),
                           f_.get(), fa_.get(), mu_.get(), p2, gu_.get(), gau_.get(), jstate());
        gu_ver_ = ver_;
        jslot_ver_ = ver_;
    } else if (batch_ == 1)
        hipLaunchKernelGGL(k_dmd_pass2<false>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(),
                           el_.get(), nn_.get(), list_.get(), ldl_, A_, bd, S, E_.get(), dE_.get(), f_.get(),
                           fa_.get(), mu_.get(), p2);
    else {
#define MAPP_P2(B, F) hipLaunchKernelGGL((k_dmd_pass2q<B, false, F>), dim3(nb), dim3(kBlock), 0, stream_, n_, \
                                         x_.get(), al_.get(), c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, \
                                         A_, bd, E_.get(), dE_.get(), f_.get(), fa_.get(), mu_.get(), p2, nullptr, \
                                         nullptr, JState{})
        if (batch_ == 2) {
            if (f32) MAPP_P2(2, true); else MAPP_P2(2, false);
        } else {
            if (f32) MAPP_P2(4, true); else MAPP_P2(4, false);
        }
#undef MAPP_P2
    }
    prof_.end(stream_);
    reduce_rows(p2, nb, 6, red_.get() + 8);
}

void DMDSystem::set_f32(int mode)
{
    MAPP_REQUIRE(mode >= 0 && mode <= 2, ERR_ARG, "table mode must be 0 (FP64 rows), 1 (FP32 knots) or 2 (FP32 rows)");
    f32_ = mode;
}

// FP32 knots for the LDS passes. Each table function's interval cubic must
// be the Hermite cubic of its two knots' values and slopes (what the spline
// builder makes); a table that is not stays on the FP32 rows. Equal functions
// (the density of one element seen from every pair of an eam/alloy table)
// are kept once, and the window is the top knots that fit the workgroup LDS
// beside the passes' static arrays.
void DMDSystem::plan_knots()
{
    knots_ok_ = false;
    const EamTables& t = ff_->tables();
    const int ne = t.ne, nr = t.nr;
    if (nr < 3) return;
    std::vector<std::vector<double>> uniq;
    std::vector<int> fn((size_t)ne * ne * 3);
    for (int a = 0; a < ne; ++a)
        for (int b = 0; b < ne; ++b)
            for (int k = 0; k < 3; ++k) {
                const double* c = k == 0 ? t.rphi_at(a, b) : (k == 1 ? t.rho_at(a, b) : t.rho_at(b, a));
                std::vector<double> ys;
                if (!hermite_knots(c, nr, ys)) return;
                int id = -1;
                for (size_t q = 0; q < uniq.size() && id < 0; ++q)
                    if (uniq[q] == ys) id = (int)q;
                if (id < 0) {
                    id = (int)uniq.size();
                    uniq.push_back(std::move(ys));
                }
                fn[((size_t)a * ne + b) * 3 + k] = id;
            }
    const int nf = (int)uniq.size();
    int dev = 0, lds_max = 0, ncu = 0;
    MAPP_HIP(hipGetDevice(&dev));
    MAPP_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    MAPP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // room for the passes' static LDS: the slot map and the block reduction
    const long budget = (long)lds_max - 4096;
    const int nwin = (int)std::min<long>(nr, budget / (8L * nf));
    if (nwin < 2 || ncu < 1) return;
    std::vector<float2> g((size_t)nf * nr);
    for (int q = 0; q < nf; ++q)
        for (int m = 0; m < nr; ++m)
            g[(size_t)q * nr + m] = make_float2((float)uniq[q][2 * (size_t)m], (float)uniq[q][2 * (size_t)m + 1]);
    knots_.reserve(g.size());
    MAPP_HIP(hipMemcpy(knots_.get(), g.data(), g.size() * sizeof(float2), hipMemcpyHostToDevice));
    knot_fn_.reserve(fn.size());
    MAPP_HIP(hipMemcpy(knot_fn_.get(), fn.data(), fn.size() * sizeof(int), hipMemcpyHostToDevice));
    kn_ = KnotArgs{knots_.get(), knot_fn_.get(), nf, nr, nwin, nr - nwin, (int)fn.size()};
    knot_lds_ = (size_t)nf * nwin * sizeof(float2);
    knot_grid_ = std::max(1, ncu / kXcd) * kXcd;
    const void* fns[] = {(const void*)k_dmd_pass1k<2, knot_nt1(2)>, (const void*)k_dmd_pass1k<4, knot_nt1(4)>,
                         (const void*)k_dmd_pass2k<2, kKnotNT2>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)knot_lds_) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
    knots_ok_ = true;
}

void DMDSystem::set_batch(int b)
{
    MAPP_REQUIRE(b == 1 || b == 2 || b == 4, ERR_ARG, "DMD batch must be 1, 2 or 4");
    batch_ = b;
}

double DMDSystem::c_d(const double* r_crd, const double* zeta, double* c_d_host)
{
    const EamTables& t = ff_->tables();
    CdArgs P{};
    for (int e = 0; e < t.ne; ++e) {
        MAPP_REQUIRE(r_crd[e] > 0.0 && zeta[e] > 0.0, ERR_ARG, "r_crd and C must be positive");
        MAPP_REQUIRE(r_crd[e] <= t.rc[(size_t)e * t.ne + e], ERR_ARG, "r_crd must not exceed r_c of the element");
        P.rsq_crd[e] = r_crd[e] * r_crd[e];
        P.zeta[e] = zeta[e];
        // c_1 of ForceFieldEAMDMD::set_temp (src/ff_eam_dmd.cpp:1062-1066)
        P.c1[e] = std::sqrt(0.5 * A_.kBT / masses_[e]) / M_PI;
    }
    P.beta = 1.0 / A_.kBT;
    // calc_mu: rho, E, dE, the raw chemical potentials and the gradient per unit c
    pass1(false);
    pass2(!surr_);
    // the free energy, entropy and virial sums of these passes, for force_calc_static
    static_.reserve(14);
    MAPP_HIP(hipMemcpyAsync(static_.get(), red_.get(), sizeof(double) * 14, hipMemcpyDeviceToDevice, stream_));
    static_ver_ = ver_;
    // what J reads later at this state: the parameters, mu and dE (the
    // slot values and F'' follow on the first J call)
    jmu_.reserve(nc_);
    jdE_.reserve(nc_);
    MAPP_HIP(hipMemcpyAsync(jmu_.get(), mu_.get(), sizeof(double) * nc_, hipMemcpyDeviceToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(jdE_.get(), dE_.get(), sizeof(double) * nc_, hipMemcpyDeviceToDevice, stream_));
    jP_ = P;
    jver_ = surr_ ? -1 : ver_;
    jrp_ver_ = -1;
    cd_.reserve(nc_);
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    const int nb = grid_for(nc_);
    prof_.begin(K_DMD_CD, stream_);
    hipLaunchKernelGGL(k_dmd_cd, dim3(nb), dim3(kBlock), 0, stream_, nc_, x_.get(), al_.get(), c_.get(), el_.get(),
                       nn_.get(), list_.get(), ldl_, C_, bd, P, mu_.get(), cd_.get(), part_.get());
    hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, nb, 1, 0, part_.get(), red_.get() + 40);
    prof_.end(stream_);
    double norm = 0.0;
    MAPP_HIP(hipMemcpyAsync(&norm, red_.get() + 40, sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (c_d_host) {
        DevBuf<double> tc;
        tc.reserve(nc_);
        hipLaunchKernelGGL(k_dgather_comp, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_,
                           slot_of_.get(), cd_.get(), tc.get());
        MAPP_HIP(hipMemcpyAsync(c_d_host, tc.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
    }
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    return norm;
}

JState DMDSystem::jstate()
{
    return JState{jidx_.get(), jr2_.get(), jr1_.get(), jt0_.get(), jcnt_.get(), jddE_.get(),
                  jidx2_.get(), jf0_.get(), jf1_.get(), jf2_.get(), jS_.get(), jcnt2_.get()};
}

void DMDSystem::j_reserve()
{
    const size_t slots = (size_t)std::max(maxn_, 1) * ldl_;
    jidx_.reserve(slots);
    jr2_.reserve(slots);
    jr1_.reserve(slots);
    jt0_.reserve(slots);
    jidx2_.reserve(slots);
    jf0_.reserve(slots);
    jf1_.reserve(slots);
    jf2_.reserve(slots);
    jcnt_.reserve(nc_);
    jcnt2_.reserve(nc_);
    jddE_.reserve(nc_);
    jS_.reserve((size_t)3 * nc_);
    ja1_.reserve(nc_);
    jb0_.reserve(nc_);
    jxs_.reserve(n_);
}

void DMDSystem::J_prepare()
{
    MAPP_REQUIRE(jver_ >= 0 && jver_ == ver_, ERR_STATE,
                 "J needs c_d at the current state (quadrature, no surrogate) first");
    if (jrp_ver_ == ver_) return;
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    j_reserve();
    prof_.begin(K_DMD_JSTATE, stream_);
    if (jslot_ver_ != ver_)  // the slot values come with c_d's second pass; else a pass of their own
        hipLaunchKernelGGL(k_dmd_jrp<4>, dim3(grid_for(n_)), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(),
                           c_.get(), el_.get(), nn_.get(), list_.get(), ldl_, maxn_, A_, bd, jdE_.get(), jstate());
    hipLaunchKernelGGL(k_dmd_jex, dim3(grid_for(nc_)), dim3(kBlock), 0, stream_, nc_, x_.get(), al_.get(), c_.get(),
                       el_.get(), nn_.get(), list_.get(), ldl_, C_, bd, jP_, jmu_.get(), jstate());
    prof_.end(stream_);
    jrp_ver_ = ver_;
}

void DMDSystem::J(const double* x_host, double* ax_host)
{
    J_prepare();
    jx_.reserve(nc_);
    jax_.reserve(nc_);
    DevBuf<double> th;
    th.reserve(nc_);
    MAPP_HIP(hipMemcpyAsync(th.get(), x_host, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_dscatter_comp, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_,
                       slot_of_.get(), th.get(), jx_.get());
    J_dev(jx_.get(), jax_.get());
    hipLaunchKernelGGL(k_dgather_comp, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_,
                       slot_of_.get(), jax_.get(), th.get());
    MAPP_HIP(hipMemcpyAsync(ax_host, th.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
}

void DMDSystem::J_entries(int64_t* live, int64_t* exchange)
{
    MAPP_REQUIRE(jrp_ver_ >= 0 && jrp_ver_ == ver_, ERR_STATE, "J_entries needs a J call at the current state");
    std::vector<int> a(nc_), b(nc_);
    MAPP_HIP(hipMemcpyAsync(a.data(), jcnt_.get(), sizeof(int) * nc_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipMemcpyAsync(b.data(), jcnt2_.get(), sizeof(int) * nc_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    *live = *exchange = 0;
    for (int i = 0; i < nc_; ++i) {
        *live += a[i];
        *exchange += b[i];
    }
}

void DMDSystem::J_dev(const double* x, double* ax)
{
    const int nbc = grid_for(nc_);
    const JState js = jstate();
    prof_.begin(K_DMD_J, stream_);
    hipLaunchKernelGGL(k_dmd_ja1, dim3(nbc), dim3(kBlock), 0, stream_, nc_, C_, c_.get(), ldl_, js, x, ja1_.get(),
                       jxs_.get());
    hipLaunchKernelGGL(k_dmd_jb0, dim3(nbc), dim3(kBlock), 0, stream_, nc_, ldl_, js, ja1_.get(), x, jb0_.get());
    hipLaunchKernelGGL(k_dmd_jax, dim3(nbc), dim3(kBlock), 0, stream_, nc_, C_, ldl_, js, jb0_.get(), x, jxs_.get(),
                       ax);
    prof_.end(stream_);
}

// __GMRES::solve and GMRES<V>::solve (src/gmres.h:40-60, src/new_min_vec.h:1129-1146)
// from zero on device vectors of length nv: one classical Gram-Schmidt pass
// per product, the Givens rotations and the back substitution on the host.
// A(in, out) applies the operator; b is left unchanged. True when the
// residual falls below tol within m products; norm = |y|.
template <class Op>
bool DMDSystem::gmres(GmresWork& w, size_t nv, int m, Op&& A, const double* b, double tol, double* x, double& norm)
{
    const int nb = (int)((nv + kBlock - 1) / kBlock);
    w.Q.reserve((size_t)(m + 1) * nv);
    w.v.reserve(nv);
    w.red.reserve(2 * kPartW);
    w.part.reserve((size_t)nb * kPartW + kPartW);
    double* Q = w.Q.get();
    double* part = w.part.get();
    const size_t ld = nv;
    const int len = (int)nv;
    auto fetch = [&](int count, double* out) {
        MAPP_HIP(hipMemcpyAsync(out, w.red.get(), sizeof(double) * count, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    std::vector<double> H((size_t)m * (m + 1) / 2), b_hat(m + 1), cs(2 * (m + 1));
    auto solve_y = [&](int nvec) {
        GmY y{};
        for (int i = nvec - 1; i > -1; --i) {
            y.y[i] = b_hat[i];
            for (int j = i + 1; j < nvec; ++j) y.y[i] -= H[(size_t)j * (j + 1) / 2 + i] * y.y[j];
            y.y[i] /= H[(size_t)i * (i + 1) / 2 + i];
        }
        hipLaunchKernelGGL(k_gm_combine, dim3(nb), dim3(kBlock), 0, stream_, len, nvec, y, Q, ld, x);
        double s2 = 0.0;
        for (int k = 0; k < nvec; ++k) s2 += y.y[k] * y.y[k];
        norm = std::sqrt(s2);
    };
    prof_.begin(K_DMD_NEWTON, stream_);
    hipLaunchKernelGGL(k_gm_dots, dim3(nb), dim3(kBlock), 0, stream_, len, 1, b, b, ld, part);
    reduce_rows(part, nb, 1, w.red.get());
    prof_.end(stream_);
    double bb;
    fetch(1, &bb);
    const double bn = std::sqrt(bb);
    if (!(bn > 0.0)) {  // already a root: no step (the reference would divide by zero here)
        hipLaunchKernelGGL(k_gm_scale, dim3(nb), dim3(kBlock), 0, stream_, len, 0.0, b, x);
        norm = 0.0;
        return true;
    }
    hipLaunchKernelGGL(k_gm_scale, dim3(nb), dim3(kBlock), 0, stream_, len, 1.0 / bn, b, Q);
    b_hat[0] = bn;
    for (int it = 0; it < m; ++it) {
        const int iv = it + 1;
        A(Q + (size_t)it * ld, w.v.get());
        prof_.begin(K_DMD_NEWTON, stream_);
        hipLaunchKernelGGL(k_gm_dots, dim3(nb), dim3(kBlock), 0, stream_, len, iv, w.v.get(), Q, ld, part);
        reduce_rows(part, nb, iv, w.red.get());
        hipLaunchKernelGGL(k_gm_orth, dim3(nb), dim3(kBlock), 0, stream_, len, iv, w.v.get(), Q, ld, w.red.get(),
                           Q + (size_t)iv * ld, part);
        reduce_rows(part, nb, 1, w.red.get() + kPartW);
        prof_.end(stream_);
        double hr[2 * kPartW];
        fetch(2 * kPartW, hr);
        double* h = H.data() + (size_t)it * (it + 1) / 2;
        for (int j = 0; j < iv; ++j) h[j] = hr[j];
        const double nrm = std::sqrt(hr[kPartW]);
        hipLaunchKernelGGL(k_gm_scale, dim3(nb), dim3(kBlock), 0, stream_, len, 1.0 / nrm, Q + (size_t)iv * ld,
                           Q + (size_t)iv * ld);
        for (int i = 0; i < it; ++i) {
            const double t = cs[2 * i] * h[i] - cs[2 * i + 1] * h[i + 1];
            h[i + 1] = cs[2 * i + 1] * h[i] + cs[2 * i] * h[i + 1];
            h[i] = t;
        }
        const double t = std::sqrt(h[it] * h[it] + nrm * nrm);
        cs[2 * it] = h[it] / t;
        cs[2 * it + 1] = -nrm / t;
        h[it] = cs[2 * it] * h[it] - cs[2 * it + 1] * nrm;
        b_hat[it + 1] = cs[2 * it + 1] * b_hat[it];
        b_hat[it] *= cs[2 * it];
        if (std::fabs(b_hat[it + 1]) < tol) {
            solve_y(it + 1);
            return true;
        }
    }
    solve_y(m);
    return false;
}

bool DMDSystem::newton(const double* r_crd, const double* zeta, double beta, const double* a_host, double a_tol,
                       double err_fac, int max_newton, int max_gmres, int* iters)
{
    MAPP_REQUIRE(beta > 0.0 && a_tol > 0.0 && err_fac > 0.0, ERR_ARG, "beta, a_tol and err_fac must be positive");
    MAPP_REQUIRE(max_newton >= 0, ERR_ARG, "max_nnewton_iters must be >= 0");
    MAPP_REQUIRE(max_gmres >= 1 && max_gmres <= kPartW, ERR_ARG, "max_ngmres_iters must be in [1, 8]");
    MAPP_REQUIRE(!surr_, ERR_STATE, "the implicit step runs on the quadrature (no surrogate)");
    const int m = max_gmres;
    const int nb = (nc_ + kBlock - 1) / kBlock;
    nt_a_.reserve(nc_);
    nt_F_.reserve(nc_);
    nt_del_.reserve(nc_);
    nt_c0_.reserve(nc_);
    nt_red_.reserve(16);
    double* part = part_.get();
    {
        DevBuf<double> th;
        th.reserve(nc_);
        MAPP_HIP(hipMemcpyAsync(th.get(), a_host, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_dscatter_comp, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_,
                           slot_of_.get(), th.get(), nt_a_.get());
        MAPP_HIP(hipStreamSynchronize(stream_));
    }
    MAPP_HIP(hipMemcpyAsync(nt_c0_.get(), c_.get(), sizeof(double) * nc_, hipMemcpyDeviceToDevice, stream_));
    const double beta_inv = 1.0 / beta;
    const double a_tol_sqrt = a_tol * std::sqrt((double)nc_);  // nc_dof = natms c_dim (src/ff_dmd.cpp:75)
    auto F_norm = [&] {
        c_d(r_crd, zeta, nullptr);
        hipLaunchKernelGGL(k_newton_f, dim3(nb), dim3(kBlock), 0, stream_, nc_, beta_inv, c_.get(), nt_a_.get(),
                           cd_.get(), nt_F_.get());
    };
    auto fetch = [&](int count, double* out) {
        MAPP_HIP(hipMemcpyAsync(out, nt_red_.get(), sizeof(double) * count, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    // (I / beta - J) applied to in: J x, then -J x + x / beta
    auto op = [&](const double* in, double* out) {
        J_prepare();
        J_dev(in, out);
        prof_.begin(K_DMD_NEWTON, stream_);
        hipLaunchKernelGGL(k_newton_jac, dim3(nb), dim3(kBlock), 0, stream_, nc_, beta_inv, in, out);
        prof_.end(stream_);
    };
    auto gmres_step = [&](double tol, double& norm) { gmres(nt_gm_, nc_, m, op, nt_F_.get(), tol, nt_del_.get(), norm); };
    // DAEImplicit::newton (src/dae_imp.cpp:146-206)
    const double res_tol = 0.005 * a_tol_sqrt / err_fac;
    const double denom = 0.1 * a_tol_sqrt / err_fac;
    double norm = 1.0, delta = 0.0, delta_prev = 0.0, ratio = 1.0, R = 1.0;
    int iter = 0;
    bool converge = false, diverge = false;
    F_norm();
    while (iter < max_newton && !converge && !diverge) {
        gmres_step(res_tol, norm);
        // update_c
        prof_.begin(K_DMD_NEWTON, stream_);
        hipLaunchKernelGGL(k_newton_r, dim3(nb), dim3(kBlock), 0, stream_, nc_, c_.get(), nt_del_.get(), part);
        hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, nb, 0, 1, part, nt_red_.get());
        prof_.end(stream_);
        double r;
        fetch(1, &r);
        hipLaunchKernelGGL(k_newton_apply, dim3(nb), dim3(kBlock), 0, stream_, nc_, r, nt_del_.get(), c_.get());
        ++ver_;
        ratio = r;
        delta = std::fabs(ratio * norm / denom);
        if (iter) R = std::max(0.3 * R, delta / delta_prev);
        if (iter > 1 && R * delta < 1.0) {
            F_norm();
            converge = true;
            continue;
        }
        if (iter > 1 && delta / delta_prev > 2.0) {
            diverge = true;
            continue;
        }
        delta_prev = delta;
        F_norm();
        iter++;
    }
    if (!converge) {
        MAPP_HIP(hipMemcpyAsync(c_.get(), nt_c0_.get(), sizeof(double) * nc_, hipMemcpyDeviceToDevice, stream_));
        ++ver_;
    }
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    if (iters) *iters = iter;
    return converge;
}

// the prep pass of ForceFieldEAMDMD::__prepJ_n_res at the current state (after
// pass 1 for F'(rho)): slot values, F'' and fcoef per component and the
// residual f in xa_f_; returns |f|
double DMDSystem::xa_prepare()
{
    MAPP_REQUIRE(!surr_, ERR_STATE, "the second-derivative pass runs on the quadrature (no surrogate)");
    pass1(false, false);
    xa_lds_ = (size_t)maxn_ * ldl_;
    xa_slots_.reserve(xa_lds_ * kXaQ);
    xa_ddE_.reserve(nc_);
    xa_fc_.reserve(nc_);
    xa_dde_.reserve(nc_);
    xa_f_.reserve((size_t)3 * n_ + nc_);
    const int nb = grid_for(n_);
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    prof_.begin(K_DMD_PREPJ, stream_);
    hipLaunchKernelGGL(k_dmd_prepj, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(), el_.get(),
                       nn_.get(), list_.get(), ldl_, A_, bd, dE_.get(), xa_slots_.get(), xa_lds_, xa_ddE_.get(),
                       xa_fc_.get(), xa_f_.get(), part_.get());
    reduce_rows(part_.get(), nb, 1, red_.get() + 48);
    prof_.end(stream_);
    double v = 0.0;
    MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 48, sizeof(double), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    xa_ver_ = ver_;
    return std::sqrt(v);
}

// J applied to a flat slot-order vector (the state of the last xa_prepare)
void DMDSystem::Jxa_dev(const double* in, double* out)
{
    const int nb = grid_for(n_);
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    prof_.begin(K_DMD_JXA, stream_);
    hipLaunchKernelGGL(k_dmd_jxa1, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(), nn_.get(),
                       list_.get(), ldl_, C_, bd, xa_slots_.get(), xa_lds_, xa_ddE_.get(), xa_fc_.get(), in, out,
                       xa_dde_.get());
    hipLaunchKernelGGL(k_dmd_jxa2, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), al_.get(), c_.get(), nn_.get(),
                       list_.get(), ldl_, C_, bd, A_.kBT, xa_slots_.get(), xa_lds_, xa_fc_.get(), xa_dde_.get(), in,
                       out);
    prof_.end(stream_);
}

void DMDSystem::Jxa(const double* dx, const double* da, double* jdx, double* jda, double* f, double* fa)
{
    update();
    xa_prepare();
    const size_t nv = (size_t)3 * n_ + nc_;
    DevBuf<double> hx, ha, tin, tout;
    hx.reserve((size_t)3 * n_);
    ha.reserve(nc_);
    tin.reserve(nv);
    tout.reserve(nv);
    const int ns = (n_ + kBlock - 1) / kBlock;
    MAPP_HIP(hipMemcpyAsync(hx.get(), dx, sizeof(double) * 3 * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(ha.get(), da, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_xa_scatter, dim3(ns), dim3(kBlock), 0, stream_, n_, C_, slot_of_.get(), hx.get(), ha.get(),
                       tin.get());
    Jxa_dev(tin.get(), tout.get());
    auto out = [&](const double* v, double* xh, double* ah) {
        hipLaunchKernelGGL(k_xa_gather, dim3(ns), dim3(kBlock), 0, stream_, n_, C_, slot_of_.get(), v, hx.get(),
                           ha.get());
        if (xh) MAPP_HIP(hipMemcpyAsync(xh, hx.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
        if (ah) MAPP_HIP(hipMemcpyAsync(ah, ha.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    out(tout.get(), jdx, jda);
    if (f || fa) out(xa_f_.get(), f, fa);
    prof_.flush();
}

// DAE::min_error_false after its CG stage (src/dae.cpp:294-325): while |f| >
// a_tol sqrt(3 n + n c_dim) and fewer than max_newton steps, GMRES solves
// J h = f from zero (tolerance 0.005 a_tol sqrt(...)); x and alpha take h
// whole when no alpha would turn negative, else half the largest fraction
// that keeps them non-negative; then the list update and f at the new state
int DMDSystem::newton_xa(double a_tol, int max_newton, int max_gmres, double* res_out)
{
    MAPP_REQUIRE(a_tol > 0.0, ERR_ARG, "a_tol must be positive");
    MAPP_REQUIRE(max_newton >= 0, ERR_ARG, "max_nnewton_iters must be >= 0");
    MAPP_REQUIRE(max_gmres >= 1 && max_gmres <= kPartW, ERR_ARG, "max_ngmres_iters must be in [1, 8]");
    const size_t nv = (size_t)3 * n_ + nc_;
    const double tol_s = a_tol * std::sqrt((double)nv);
    xa_h_.reserve(nv);
    update();
    double res = xa_prepare();
    const int nbe = (nc_ + kBlock - 1) / kBlock;
    const int nba = (std::max(n_, nc_) + kBlock - 1) / kBlock;
    int istep = 0;
    for (; istep < max_newton && res / tol_s > 1.0; ++istep) {
        double norm;
        gmres(xa_gm_, nv, max_gmres, [&](const double* in, double* out) { Jxa_dev(in, out); }, xa_f_.get(),
              0.005 * tol_s, xa_h_.get(), norm);
        prof_.begin(K_DMD_NEWTON, stream_);
        hipLaunchKernelGGL(k_xa_limit, dim3(nbe), dim3(kBlock), 0, stream_, nc_, al_.get(),
                           xa_h_.get() + (size_t)3 * n_, part_.get());
        hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, nbe, 0, 1, part_.get(),
                           red_.get() + 52);
        prof_.end(stream_);
        double r = 1.0;
        MAPP_HIP(hipMemcpyAsync(&r, red_.get() + 52, sizeof(double), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        const double t = r == 1.0 ? 1.0 : 0.5 * r;
        hipLaunchKernelGGL(k_xa_apply, dim3(nba), dim3(kBlock), 0, stream_, n_, nc_, t, xa_h_.get(), x_.get(),
                           al_.get());
        ++ver_;
        update();
        res = xa_prepare();
    }
    prof_.flush();
    if (res_out) *res_out = res;
    return istep;
}

DMDReport DMDSystem::derivative(bool with_c, const double* ext_mu)
{
    pass1(false, f32_);
    pass2(false, f32_);
    if (with_c) {
        MAPP_REQUIRE(ext_mu != nullptr, ERR_ARG, "ext_mu is required with c degrees of freedom");
        DevBuf<double> em;
        em.reserve(kMaxElems);
        MAPP_HIP(hipMemcpyAsync(em.get(), ext_mu, sizeof(double) * A_.ne, hipMemcpyHostToDevice, stream_));
        const int nb = grid_for(n_);
        double* p3 = part_.get() + (size_t)nb * kPartW;
        hipLaunchKernelGGL(k_dmd_chem, dim3(nb), dim3(kBlock), 0, stream_, n_, C_, al_.get(), c_.get(), el_.get(),
                           A_.kBT, A_, em.get(), mu_.get(), p3);
        reduce_rows(p3, nb, 1, red_.get() + 16);
        MAPP_HIP(hipStreamSynchronize(stream_));
    }
    double h[24];
    MAPP_HIP(hipMemcpyAsync(h, red_.get(), sizeof(double) * 24, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    DMDReport r{};
    r.s = h[1];
    r.fe = h[0] + A_.kBT * h[1] - (with_c ? h[16] : 0.0);
    for (int k = 0; k < 6; ++k) r.FH[k] = h[8 + k];
    return r;
}

double DMDSystem::value(bool with_c, const double* ext_mu)
{
    pass1(true, f32_);
    double h[2];
    MAPP_HIP(hipMemcpyAsync(h, red_.get(), sizeof(double) * 2, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    double fe = h[0] + A_.kBT * h[1];
    if (with_c) {
        MAPP_REQUIRE(ext_mu != nullptr, ERR_ARG, "ext_mu is required with c degrees of freedom");
        std::vector<double> c(nc_);
        std::vector<int> e(nc_);
        MAPP_HIP(hipMemcpy(c.data(), c_.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost));
        MAPP_HIP(hipMemcpy(e.data(), el_.get(), sizeof(int) * nc_, hipMemcpyDeviceToHost));
        double s = 0.0;
        for (int k = 0; k < nc_; ++k)
            if (c[k] >= 0.0) s += c[k] * ext_mu[e[k]];
        fe -= s;
    }
    return fe;
}

void DMDSystem::set_state(const double* x, const double* alpha, const double* c)
{
    ++ver_;
    DevBuf<double> tx, ta, tc;
    tx.reserve((size_t)3 * n_);
    ta.reserve(nc_);
    tc.reserve(nc_);
    MAPP_HIP(hipMemcpyAsync(tx.get(), x, sizeof(double) * 3 * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(ta.get(), alpha, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(tc.get(), c, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_dscatter, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_, tx.get(),
                       ta.get(), tc.get(), slot_of_.get(), x_.get(), al_.get(), c_.get());
    MAPP_HIP(hipStreamSynchronize(stream_));
}

// A new box and, unless x is null, new positions in caller order (alpha and c
// kept): the grid, then a rebuild. A box the grid cannot hold leaves the
// system as it was.
void DMDSystem::set_box(const double* H, const double* x)
{
    for (int i = 0; i < 3; ++i)
        MAPP_REQUIRE(H[i * 3 + i] > 0.0, ERR_ARG, "H must be lower triangular with a positive diagonal");
    MAPP_REQUIRE(H[1] == 0.0 && H[2] == 0.0 && H[5] == 0.0, ERR_ARG, "H must be lower triangular");
    const Box old = box_;
    box_ = make_box(H);
    try {
        setup_grid();
    } catch (...) {
        box_ = old;
        setup_grid();
        throw;
    }
    if (x) {
        std::vector<double> xo((size_t)3 * n_), a(nc_), c(nc_);
        get_state(xo.data(), a.data(), c.data());
        set_state(x, a.data(), c.data());
    }
    rebuild();
}

// ForceFieldDMD::force_calc_static (src/ff_dmd.cpp:429-445,
// src/ff_eam_dmd.cpp:1284-1340): the free energy, virial and entropy of the
// state the last c_d left, from the sums its passes already formed
DMDReport DMDSystem::force_calc_static()
{
    MAPP_REQUIRE(static_ver_ >= 0 && static_ver_ == ver_, ERR_STATE,
                 "force_calc_static reads the state of the last c_d call; x, alpha or c changed since");
    double h[14];
    MAPP_HIP(hipMemcpyAsync(h, static_.get(), sizeof(double) * 14, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    DMDReport r{};
    r.s = h[1];
    r.fe = h[0] + A_.kBT * h[1];
    for (int k = 0; k < 6; ++k) r.FH[k] = h[8 + k];
    return r;
}

void DMDSystem::export_cfg(const char* path, int nelems, const double* masses, const std::vector<std::string>& names)
{
    MAPP_REQUIRE(nelems >= ff_->tables().ne, ERR_ARG, "one name and mass per element of the force field");
    std::vector<double> x((size_t)3 * n_), a(nc_), c(nc_);
    get_state(x.data(), a.data(), c.data());
    double H[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H[i * 3 + j] = box_.H[i][j];
    write_cfg_dmd(path, n_, C_, H, x.data(), elem_host_.data(), a.data(), c.data(), nelems, masses, names);
}

void DMDSystem::set_c(const double* c)
{
    ++ver_;
    DevBuf<double> tc;
    tc.reserve(nc_);
    MAPP_HIP(hipMemcpyAsync(tc.get(), c, sizeof(double) * nc_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_dscatter_comp, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_,
                       slot_of_.get(), tc.get(), c_.get());
    MAPP_HIP(hipStreamSynchronize(stream_));
}

double DMDSystem::unit_grad_err_sq()
{
    MAPP_REQUIRE(gu_ver_ >= 0 && gu_ver_ == ver_, ERR_STATE, "the gradient per unit c needs c_d at the current state");
    const int nb = grid_for(std::max(n_, nc_));
    hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(kBlock), 0, stream_, n_, nc_, gu_.get(), gau_.get(), gu_.get(),
                       gau_.get(), part_.get());
    reduce_rows(part_.get(), nb, 1, red_.get() + 24);
    double v = 0.0;
    MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 24, sizeof(double), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    return v;
}

void DMDSystem::unit_gradient(double* g_host, double* ga_host)
{
    MAPP_REQUIRE(gu_ver_ >= 0 && gu_ver_ == ver_, ERR_STATE, "the gradient per unit c needs c_d at the current state");
    DevBuf<double> tx, ta;
    tx.reserve((size_t)3 * n_);
    ta.reserve(nc_);
    hipLaunchKernelGGL(k_dgather, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_, slot_of_.get(),
                       gu_.get(), gau_.get(), (const double*)nullptr, tx.get(), ta.get(), (double*)nullptr);
    if (g_host) MAPP_HIP(hipMemcpyAsync(g_host, tx.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
    if (ga_host) MAPP_HIP(hipMemcpyAsync(ga_host, ta.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
}

double DMDSystem::grad_err_sq()
{
    const int nb = grid_for(std::max(n_, nc_));
    hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(kBlock), 0, stream_, n_, nc_, f_.get(), fa_.get(), f_.get(), fa_.get(),
                       part_.get());
    reduce_rows(part_.get(), nb, 1, red_.get() + 24);
    double v = 0.0;
    MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 24, sizeof(double), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    return v;
}

void DMDSystem::get_state(double* x, double* alpha, double* c)
{
    DevBuf<double> tx, ta, tc;
    tx.reserve((size_t)3 * n_);
    ta.reserve(nc_);
    tc.reserve(nc_);
    hipLaunchKernelGGL(k_dgather, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_, slot_of_.get(),
                       x_.get(), al_.get(), c_.get(), tx.get(), ta.get(), tc.get());
    if (x) MAPP_HIP(hipMemcpyAsync(x, tx.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
    if (alpha) MAPP_HIP(hipMemcpyAsync(alpha, ta.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
    if (c) MAPP_HIP(hipMemcpyAsync(c, tc.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
}

void DMDSystem::get_forces(double* f, double* f_alpha, double* mu)
{
    DevBuf<double> tx, ta, tc;
    tx.reserve((size_t)3 * n_);
    ta.reserve(nc_);
    tc.reserve(nc_);
    hipLaunchKernelGGL(k_dgather, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_, slot_of_.get(),
                       f_.get(), fa_.get(), mu_.get(), tx.get(), ta.get(), tc.get());
    if (f) MAPP_HIP(hipMemcpyAsync(f, tx.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
    if (f_alpha) MAPP_HIP(hipMemcpyAsync(f_alpha, ta.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
    if (mu) MAPP_HIP(hipMemcpyAsync(mu, tc.get(), sizeof(double) * nc_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
}

// (i, j, n) with n != 0 or site(i) < site(j) in caller indices
// (src/neighbor_dmd.cpp:150-167 keeps i < j over local sites and every ghost)
int64_t DMDSystem::half_pairs_count()
{
    std::vector<int> nn(nc_), id(n_);
    std::vector<uint32_t> list((size_t)maxn_ * ldl_);
    MAPP_HIP(hipMemcpy(nn.data(), nn_.get(), sizeof(int) * nc_, hipMemcpyDeviceToHost));
    MAPP_HIP(hipMemcpy(id.data(), id_.get(), sizeof(int) * n_, hipMemcpyDeviceToHost));
    MAPP_HIP(hipMemcpy(list.data(), list_.get(), sizeof(uint32_t) * list.size(), hipMemcpyDeviceToHost));
    int64_t cnt = 0;
    for (int q = 0; q < nc_; ++q)
        for (int k = 0; k < nn[q]; ++k) {
            const uint32_t e = list[(size_t)k * ldl_ + q];
            const int jc = (int)(e & kIndexMask);
            if ((e >> kIndexBits) != (uint32_t)kZeroImage || id[q / C_] < id[jc / C_]) ++cnt;
        }
    return cnt;
}

void DMDSystem::half_pairs(int32_t* pi, int32_t* pj, int8_t* img)
{
    std::vector<int> nn(nc_), id(n_);
    std::vector<uint32_t> list((size_t)maxn_ * ldl_);
    MAPP_HIP(hipMemcpy(nn.data(), nn_.get(), sizeof(int) * nc_, hipMemcpyDeviceToHost));
    MAPP_HIP(hipMemcpy(id.data(), id_.get(), sizeof(int) * n_, hipMemcpyDeviceToHost));
    MAPP_HIP(hipMemcpy(list.data(), list_.get(), sizeof(uint32_t) * list.size(), hipMemcpyDeviceToHost));
    int64_t c = 0;
    for (int q = 0; q < nc_; ++q)
        for (int k = 0; k < nn[q]; ++k) {
            const uint32_t e = list[(size_t)k * ldl_ + q];
            const int jc = (int)(e & kIndexMask);
            const uint32_t code = e >> kIndexBits;
            if (code != (uint32_t)kZeroImage || id[q / C_] < id[jc / C_]) {
                pi[c] = id[q / C_] * C_ + q % C_;
                pj[c] = id[jc / C_] * C_ + jc % C_;
                img[3 * c + 2] = (int8_t)((int)(code / 25) - 2);
                img[3 * c + 1] = (int8_t)((int)((code % 25) / 5) - 2);
                img[3 * c + 0] = (int8_t)((int)(code % 5) - 2);
                ++c;
            }
        }
}

void DMDSystem::smear_dev(int tt, int npts, const double* r, const double* a, double* out)
{
    MAPP_REQUIRE(tt >= 0 && tt < A_.ne * A_.ne, ERR_ARG, "table index out of range");
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_smear_points, dim3((npts + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, A_, tt, npts,
                       r, a, out);
    MAPP_HIP(hipGetLastError());
    MAPP_HIP(hipStreamSynchronize(stream_));
}

void DMDSystem::smear(int ei, int ej, int npts, const double* r, const double* a, double* out)
{
    MAPP_REQUIRE(ei >= 0 && ej >= 0 && ei < A_.ne && ej < A_.ne, ERR_ARG, "element index out of range");
    DevBuf<double> dr, da, dout;
    dr.reserve(npts);
    da.reserve(npts);
    dout.reserve((size_t)npts * 9);
    MAPP_HIP(hipMemcpy(dr.get(), r, sizeof(double) * npts, hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(da.get(), a, sizeof(double) * npts, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_smear_points, dim3((npts + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, A_,
                       ei * A_.ne + ej, npts, dr.get(), da.get(), dout.get());
    MAPP_HIP(hipStreamSynchronize(stream_));
    MAPP_HIP(hipMemcpy(out, dout.get(), sizeof(double) * npts * 9, hipMemcpyDeviceToHost));
}

// MinCGDMD::__run over x and alpha (src/min_cg_dmd.h:633-700) with the
// backtracking line search (src/ls.h:711-770: c = 0.4, rho = 0.5, the
// initial step from the previous decrease) and the step limits of
// MinDMDHandler::ls_prep (max_dx, max_dalpha, alpha kept positive). Stops on
// the energy decrement e_tol like the reference, or when the largest
// gradient component drops below grad_tol (0 disables).
// The DMD minimisers' status for a failed line search, the same for CG and
// L-BFGS: 1 zero gradient (LS_F_GRAD0), 2 not downhill (LS_F_DOWNHILL), 4 any
// other failure (minimum alpha reached, or a bracketing failure of the golden
// or Brent search).
static int dmd_min_status(int ls)
{
    if (ls == LS_F_GRAD0) return 1;
    if (ls == LS_F_DOWNHILL) return 2;
    return 4;
}

CGReport DMDSystem::cg(int max_iter, double e_tol, double max_dx, double max_dalpha, double grad_tol,
                       const LineSearchParams* ls)
{
    ++ver_;
    const int nb = grid_for(std::max(n_, nc_));
    const int nbe = (std::max(n_, nc_) + kBlock - 1) / kBlock;
    cg_x0_.reserve(n_); cg_h_.reserve(n_); cg_f0_.reserve(n_);
    cg_a0_.reserve(nc_); cg_ha_.reserve(nc_); cg_fa0_.reserve(nc_);
    double* part = part_.get();
    auto dotv = [&](const double4* ax, const double* aa, const double4* bx, const double* ba) {
        hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(kBlock), 0, stream_, n_, nc_, ax, aa, bx, ba, part);
        reduce_rows(part, nb, 1, red_.get() + 24);
        double v;
        MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 24, sizeof(double), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        return v;
    };
    // extrema over the grid: at most 256 workgroups, each row of the partial buffer once
    const int gx = std::min(256, grid_for(std::max(n_, nc_)));
    auto gmax = [&]() {
        hipLaunchKernelGGL(k_gmax, dim3(gx), dim3(kBlock), 0, stream_, n_, nc_, f_.get(), fa_.get(), part);
        hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, gx, 1, 0, part, red_.get() + 28);
        double v;
        MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 28, sizeof(double), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        return v;
    };
    auto copy_state = [&](DevBuf<double4>& dx, DevBuf<double>& da, const double4* sx, const double* sa) {
        MAPP_HIP(hipMemcpyAsync(dx.get(), sx, sizeof(double4) * n_, hipMemcpyDeviceToDevice, stream_));
        MAPP_HIP(hipMemcpyAsync(da.get(), sa, sizeof(double) * nc_, hipMemcpyDeviceToDevice, stream_));
    };
    // the list update of a trial point; after a rebuild the CG vectors follow
    // their sites into the new slot order
    auto move_update = [&]() {
        if (!update()) return;
        DevBuf<double4> x0n, hn, f0n;
        DevBuf<double> a0n, han, fa0n;
        x0n.reserve(n_); hn.reserve(n_); f0n.reserve(n_);
        a0n.reserve(nc_); han.reserve(nc_); fa0n.reserve(nc_);
        hipLaunchKernelGGL(k_dcg_follow, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, C_,
                           vals_alt_.get(), x_.get(), x_alt_.get(), cg_x0_.get(), cg_h_.get(), cg_f0_.get(),
                           cg_a0_.get(), cg_ha_.get(), cg_fa0_.get(), x0n.get(), hn.get(), f0n.get(), a0n.get(),
                           han.get(), fa0n.get());
        auto take = [](auto& a, auto& b) {
            std::swap(a.p, b.p);
            std::swap(a.n, b.n);
        };
        take(cg_x0_, x0n); take(cg_h_, hn); take(cg_f0_, f0n);
        take(cg_a0_, a0n); take(cg_ha_, han); take(cg_fa0_, fa0n);
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    CGReport rep{};
    const int reb0 = nrebuild_;
    // MinDMDHandler::F / F_reset (src/min_cg_dmd.h:462-475)
    auto f_trial = [&](double t) {
        hipLaunchKernelGGL(k_cg_move, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, t, cg_x0_.get(), cg_h_.get(),
                           cg_a0_.get(), cg_ha_.get(), c_.get(), x_.get(), al_.get());
        move_update();
        ++rep.evals;
        return value(false, nullptr);
    };
    auto f_reset = [&]() {
        hipLaunchKernelGGL(k_cg_move, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, 0.0, cg_x0_.get(), cg_h_.get(),
                           cg_a0_.get(), cg_ha_.get(), c_.get(), x_.get(), al_.get());
        move_update();
    };
    // LineSearchBackTrack's defaults (src/ls.cpp:456-461) unless the caller chose a search
    const LineSearchParams lp = ls ? *ls : LineSearchParams{};
    LineSearcher<decltype(f_trial), decltype(f_reset)> searcher(lp, f_trial, f_reset);
    DMDReport d = derivative(false, nullptr);
    rep.evals = 1;
    rep.fe0 = d.fe;
    double e_curr = d.fe;
    copy_state(cg_h_, cg_ha_, f_.get(), fa_.get());
    double f0f0 = dotv(f_.get(), fa_.get(), f_.get(), fa_.get());
    int status = 0, it = 0;
    for (; it < max_iter; ++it) {
        if (f0f0 == 0.0) { status = 1; break; }
        if (grad_tol > 0.0 && gmax() < grad_tol) { status = 3; break; }
        copy_state(cg_x0_, cg_a0_, x_.get(), al_.get());
        copy_state(cg_f0_, cg_fa0_, f_.get(), fa_.get());
        const double e_prev = e_curr;
        double fh = dotv(f_.get(), fa_.get(), cg_h_.get(), cg_ha_.get());
        if (fh < 0.0) {
            copy_state(cg_h_, cg_ha_, f_.get(), fa_.get());
            fh = f0f0;
        }
        // ls_prep: largest step within max_dx, max_dalpha and alpha > 0
        hipLaunchKernelGGL(k_cg_limits, dim3(gx), dim3(kBlock), 0, stream_, n_, nc_, cg_h_.get(), cg_ha_.get(),
                           al_.get(), c_.get(), part);
        hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, gx, 2, 1, part, red_.get() + 32);
        double lim[3];
        MAPP_HIP(hipMemcpyAsync(lim, red_.get() + 32, sizeof(double) * 3, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        double max_a = std::numeric_limits<double>::infinity();
        if (lim[0] > 0.0) max_a = std::min(max_a, std::fabs(max_dx / lim[0]));
        if (lim[1] > 0.0) max_a = std::min(max_a, std::fabs(max_dalpha / lim[1]));
        max_a = std::min(max_a, lim[2]);
        const double dfa = -fh;
        if (dfa >= 0.0 || !(max_a > 0.0)) { status = 2; break; }
        // LineSearchBackTrack::min with init_flag 1 (src/ls.h:711-769)
        double t = 0.0;
        const int ls = searcher.min(e_curr, t, dfa, max_a, 1);
        if (ls != LS_S) {
            status = dmd_min_status(ls);
            break;
        }
        d = derivative(false, nullptr);
        ++rep.evals;
        if (e_prev - e_curr < e_tol) { status = 5; ++it; break; }
        const double ff = dotv(f_.get(), fa_.get(), f_.get(), fa_.get());
        const double ff0 = dotv(f_.get(), fa_.get(), cg_f0_.get(), cg_fa0_.get());
        const double ratio = (ff - ff0) / f0f0;
        hipLaunchKernelGGL(k_cg_dir, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, ratio, f_.get(), fa_.get(),
                           cg_h_.get(), cg_ha_.get());
        f0f0 = ff;
    }
    rep.iters = it;
    rep.status = status;
    rep.fe = e_curr;
    rep.grad_norm = std::sqrt(dotv(f_.get(), fa_.get(), f_.get(), fa_.get()));
    rep.grad_max = gmax();
    rep.rebuilds = nrebuild_ - reb0;
    return rep;
}

CGReport DMDSystem::lbfgs(int m, int max_iter, double e_tol, double max_dx, double max_dalpha, double grad_tol,
                          const LineSearchParams* ls)
{
    MAPP_REQUIRE(m >= 0 && m <= 64, ERR_ARG, "L-BFGS memory must lie in [0, 64]");
    ++ver_;
    const int nb = grid_for(std::max(n_, nc_));
    const int nbe = (std::max(n_, nc_) + kBlock - 1) / kBlock;
    cg_x0_.reserve(n_); cg_h_.reserve(n_); cg_f0_.reserve(n_);
    cg_a0_.reserve(nc_); cg_ha_.reserve(nc_); cg_fa0_.reserve(nc_);
    // the stored pairs (s, y), position and alpha parts
    std::vector<std::unique_ptr<DevBuf<double4>>> sx, yx;
    std::vector<std::unique_ptr<DevBuf<double>>> sa, ya;
    for (int i = 0; i < m; ++i) {
        sx.push_back(std::make_unique<DevBuf<double4>>());
        yx.push_back(std::make_unique<DevBuf<double4>>());
        sa.push_back(std::make_unique<DevBuf<double>>());
        ya.push_back(std::make_unique<DevBuf<double>>());
        sx[i]->reserve(n_); yx[i]->reserve(n_); sa[i]->reserve(nc_); ya[i]->reserve(nc_);
    }
    double* part = part_.get();
    auto dotv = [&](const double4* ax, const double* aa, const double4* bx, const double* ba) {
        hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(kBlock), 0, stream_, n_, nc_, ax, aa, bx, ba, part);
        reduce_rows(part, nb, 1, red_.get() + 24);
        double v;
        MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 24, sizeof(double), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        return v;
    };
    const int gx = std::min(256, grid_for(std::max(n_, nc_)));
    auto gmax = [&]() {
        hipLaunchKernelGGL(k_gmax, dim3(gx), dim3(kBlock), 0, stream_, n_, nc_, f_.get(), fa_.get(), part);
        hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, gx, 1, 0, part, red_.get() + 28);
        double v;
        MAPP_HIP(hipMemcpyAsync(&v, red_.get() + 28, sizeof(double), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        return v;
    };
    auto copy_state = [&](DevBuf<double4>& dx, DevBuf<double>& da, const double4* src_x, const double* src_a) {
        MAPP_HIP(hipMemcpyAsync(dx.get(), src_x, sizeof(double4) * n_, hipMemcpyDeviceToDevice, stream_));
        MAPP_HIP(hipMemcpyAsync(da.get(), src_a, sizeof(double) * nc_, hipMemcpyDeviceToDevice, stream_));
    };
    auto axpy = [&](double a, const DevBuf<double4>* ux, const DevBuf<double>* ua, double g) {
        hipLaunchKernelGGL(k_dlb_axpy, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, a, ux ? ux->get() : nullptr,
                           ua ? ua->get() : nullptr, g, cg_h_.get(), cg_ha_.get());
    };
    auto take = [](auto& a, auto& b) {
        std::swap(a.p, b.p);
        std::swap(a.n, b.n);
    };
    // the list update of a trial point; after a rebuild every vector follows
    // its site into the new slot order (the origin x0 with the site's wrap)
    auto move_update = [&]() {
        if (!update()) return;
        DevBuf<double4> x0n, hn, f0n;
        DevBuf<double> a0n, han, fa0n;
        x0n.reserve(n_); hn.reserve(n_); f0n.reserve(n_);
        a0n.reserve(nc_); han.reserve(nc_); fa0n.reserve(nc_);
        const dim3 g((n_ + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_dcg_follow, g, dim3(kBlock), 0, stream_, n_, C_, vals_alt_.get(), x_.get(),
                           x_alt_.get(), cg_x0_.get(), cg_h_.get(), cg_f0_.get(), cg_a0_.get(), cg_ha_.get(),
                           cg_fa0_.get(), x0n.get(), hn.get(), f0n.get(), a0n.get(), han.get(), fa0n.get());
        take(cg_x0_, x0n); take(cg_h_, hn); take(cg_f0_, f0n);
        take(cg_a0_, a0n); take(cg_ha_, han); take(cg_fa0_, fa0n);
        for (int i = 0; i < m; ++i) {
            DevBuf<double4> tx;
            DevBuf<double> ta;
            tx.reserve(n_);
            ta.reserve(nc_);
            hipLaunchKernelGGL(k_dlb_follow, g, dim3(kBlock), 0, stream_, n_, C_, vals_alt_.get(), sx[i]->get(),
                               sa[i]->get(), tx.get(), ta.get());
            take(*sx[i], tx); take(*sa[i], ta);
            hipLaunchKernelGGL(k_dlb_follow, g, dim3(kBlock), 0, stream_, n_, C_, vals_alt_.get(), yx[i]->get(),
                               ya[i]->get(), tx.get(), ta.get());
            take(*yx[i], tx); take(*ya[i], ta);
            MAPP_HIP(hipStreamSynchronize(stream_));  // tx, ta are freed at the end of this scope
        }
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    CGReport rep{};
    const int reb0 = nrebuild_;
    // MinDMDHandler::F / F_reset (src/min_cg_dmd.h:462-475)
    auto f_trial = [&](double t) {
        hipLaunchKernelGGL(k_cg_move, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, t, cg_x0_.get(), cg_h_.get(),
                           cg_a0_.get(), cg_ha_.get(), c_.get(), x_.get(), al_.get());
        move_update();
        ++rep.evals;
        return value(false, nullptr);
    };
    auto f_reset = [&]() {
        hipLaunchKernelGGL(k_cg_move, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, 0.0, cg_x0_.get(), cg_h_.get(),
                           cg_a0_.get(), cg_ha_.get(), c_.get(), x_.get(), al_.get());
        move_update();
    };
    // LineSearchBackTrack's defaults (src/ls.cpp:456-461) unless the caller chose a search
    const LineSearchParams lp = ls ? *ls : LineSearchParams{};
    LineSearcher<decltype(f_trial), decltype(f_reset)> searcher(lp, f_trial, f_reset);

    // MinLBFGSDMD::__run (src/min_l-bfgs_dmd.h:113-210)
    DMDReport d = derivative(false, nullptr);
    rep.evals = 1;
    rep.fe0 = d.fe;
    double e_curr = d.fe;
    std::vector<double> rho(m, 0.0), al(m, 0.0);
    std::vector<int> ord(m);
    for (int i = 0; i < m; ++i) ord[i] = i;
    int k = 0, status = 0, it = 0;
    double gamma = 1.0;
    for (; it < max_iter; ++it) {
        if (grad_tol > 0.0 && gmax() < grad_tol) { status = 3; break; }
        copy_state(cg_x0_, cg_a0_, x_.get(), al_.get());
        copy_state(cg_f0_, cg_fa0_, f_.get(), fa_.get());
        copy_state(cg_h_, cg_ha_, f_.get(), fa_.get());
        for (int i = 0; i < k; ++i) {
            const int b = ord[i];
            al[i] = -rho[i] * dotv(sx[b]->get(), sa[b]->get(), cg_h_.get(), cg_ha_.get());
            axpy(al[i], yx[b].get(), ya[b].get(), 1.0);
        }
        if (gamma != 1.0) axpy(0.0, nullptr, nullptr, gamma);
        for (int i = k - 1; i > -1; --i) {
            const int b = ord[i];
            const double scl0 = -al[i] - rho[i] * dotv(yx[b]->get(), ya[b]->get(), cg_h_.get(), cg_ha_.get());
            axpy(scl0, sx[b].get(), sa[b].get(), 1.0);
        }
        const double e_prev = e_curr;
        double f_h = dotv(f_.get(), fa_.get(), cg_h_.get(), cg_ha_.get());
        if (f_h < 0.0) {
            copy_state(cg_h_, cg_ha_, f_.get(), fa_.get());
            k = 0;
            f_h = dotv(f_.get(), fa_.get(), f_.get(), fa_.get());
        }
        // MinDMDHandler::ls_prep (src/min_cg_dmd.h:432-458)
        const double hh = dotv(cg_h_.get(), cg_ha_.get(), cg_h_.get(), cg_ha_.get());
        double dfa = -f_h, max_a = 0.0;
        if (hh == 0.0) {
            dfa = 0.0;
        } else if (dfa >= 0.0) {
            dfa = 1.0;
        } else {
            hipLaunchKernelGGL(k_cg_limits, dim3(gx), dim3(kBlock), 0, stream_, n_, nc_, cg_h_.get(), cg_ha_.get(),
                               al_.get(), c_.get(), part);
            hipLaunchKernelGGL(k_extrema_rows, dim3(1), dim3(kBlock), 0, stream_, gx, 2, 1, part, red_.get() + 32);
            double lim[3];
            MAPP_HIP(hipMemcpyAsync(lim, red_.get() + 32, sizeof(double) * 3, hipMemcpyDeviceToHost, stream_));
            MAPP_HIP(hipStreamSynchronize(stream_));
            max_a = std::numeric_limits<double>::infinity();
            if (lim[0] > 0.0) max_a = std::min(max_a, std::fabs(max_dx / lim[0]));
            if (lim[1] > 0.0) max_a = std::min(max_a, std::fabs(max_dalpha / lim[1]));
            max_a = std::min(max_a, lim[2]);
        }
        double alpha_m = 0.0;
        const int ls = searcher.min(e_curr, alpha_m, dfa, max_a, 0);
        if (ls != LS_S) {
            status = dmd_min_status(ls);
            break;
        }
        d = derivative(false, nullptr);
        ++rep.evals;
        if (e_prev - e_curr < e_tol) { status = 5; ++it; break; }
        if (m) {
            if (k != m) ++k;
            std::rotate(ord.begin(), ord.begin() + (k - 1), ord.begin() + k);
            for (int i = m - 1; i > 0; --i) rho[i] = rho[i - 1];
            const int b = ord[0];
            hipLaunchKernelGGL(k_dlb_pair, dim3(nbe), dim3(kBlock), 0, stream_, n_, nc_, alpha_m, cg_h_.get(),
                               cg_ha_.get(), cg_f0_.get(), cg_fa0_.get(), f_.get(), fa_.get(), sx[b]->get(),
                               sa[b]->get(), yx[b]->get(), ya[b]->get());
            const double inner0 = dotv(sx[b]->get(), sa[b]->get(), yx[b]->get(), ya[b]->get());
            const double inner1 = dotv(yx[b]->get(), ya[b]->get(), yx[b]->get(), ya[b]->get());
            gamma = inner0 / inner1;
            rho[0] = 1.0 / inner0;
        } else {
            const double hf0 = dotv(cg_h_.get(), cg_ha_.get(), cg_f0_.get(), cg_fa0_.get());
            const double hf = dotv(cg_h_.get(), cg_ha_.get(), f_.get(), fa_.get());
            const double ff = dotv(f_.get(), fa_.get(), f_.get(), fa_.get());
            const double f0f0 = dotv(cg_f0_.get(), cg_fa0_.get(), cg_f0_.get(), cg_fa0_.get());
            const double ff0 = dotv(f_.get(), fa_.get(), cg_f0_.get(), cg_fa0_.get());
            gamma = alpha_m * (hf0 - hf) / (ff + f0f0 - 2.0 * ff0);
        }
    }
    rep.iters = it;
    rep.status = status;
    rep.fe = e_curr;
    rep.grad_norm = std::sqrt(dotv(f_.get(), fa_.get(), f_.get(), fa_.get()));
    rep.grad_max = gmax();
    rep.rebuilds = nrebuild_ - reb0;
    return rep;
}

}  // namespace mapp
