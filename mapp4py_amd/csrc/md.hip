// md.hip -- device-resident MD: cell sort, neighbor list, two-pass EAM,
// velocity-Verlet / Nose-Hoover-chain NVT. Kernels are written for gfx950
// (wave64, 8 XCDs): one thread per atom, lists stored slot-major so a wave
// reads 256 contiguous bytes per neighbor rank, and workgroup ids remapped
// so that each XCD (and its L2) works on one contiguous, spatially compact
// range of cell-sorted atoms.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "md.hpp"
#include "kern.hpp"

namespace mapp {

namespace {

struct GridArgs {
    int nc[3];
    int R[3];
    int ne;
    double cut_sk_sq[kMaxElems * kMaxElems];
};
// ---------------- neighbor build ----------------

// wrap into the box, remember s, sort key = Morton rank of the cell
__global__ __launch_bounds__(kBlock) void k_cell_keys(int n, const double4* __restrict__ x, BoxDev bd, GridArgs g,
                                                      const int* __restrict__ cell_rank, double4* __restrict__ s_tmp,
                                                      uint32_t* __restrict__ keys, int* __restrict__ vals)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 xi = x[i];
    double s0, s1, s2;
    x_to_s(bd.b, xi.x, xi.y, xi.z, s0, s1, s2);
    s_tmp[i] = make_double4(s0, s1, s2, xi.w);
    const int c = cell_coord(s0, g.nc[0]) + g.nc[0] * (cell_coord(s1, g.nc[1]) + g.nc[1] * cell_coord(s2, g.nc[2]));
    keys[i] = (uint32_t)cell_rank[c];
    vals[i] = i;
}

// gather every per-atom array into the new cell order; x is re-expressed from
// the wrapped s exactly as a rebuild does (src/cell.h:208)
__global__ __launch_bounds__(kBlock) void k_permute(int n, const int* __restrict__ perm, const double4* __restrict__ s_tmp,
                                                    const double4* __restrict__ v_old, const int* __restrict__ id_old,
                                                    BoxDev bd, double4* __restrict__ x_new, double4* __restrict__ v_new,
                                                    int* __restrict__ id_new, double4* __restrict__ s4,
                                                    double4* __restrict__ x0, int* __restrict__ slot_of)
{
    const int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= n) return;
    const int o = perm[p];
    const double4 s = s_tmp[o];
    double x0_, x1_, x2_;
    s_to_x(bd.h, s.x, s.y, s.z, x0_, x1_, x2_);
    const double4 xn = make_double4(x0_, x1_, x2_, s.w);
    x_new[p] = xn;
    x0[p] = xn;
    s4[p] = s;
    v_new[p] = v_old[o];
    const int id = id_old[o];
    id_new[p] = id;
    slot_of[id] = p;
}

// cell_start[r] = first sorted slot whose key >= r (r = 0..ncells)
__global__ __launch_bounds__(kBlock) void k_cell_start(int n, int ncells, const uint32_t* __restrict__ keys,
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

// Full neighbor list with the reference's FP64 predicate: x_i = s_i H and the
// image of j at (s_j + n) H, rsq = dx0^2 + (dx1^2 + dx2^2) < (r_c + skin)^2.
__global__ __launch_bounds__(kBlock) void k_build_list(int n, const double4* __restrict__ s4, BoxDev bd, GridArgs g,
                                                       const int* __restrict__ cell_rank,
                                                       const int* __restrict__ cell_start, int maxn, int ldl,
                                                       uint32_t* __restrict__ list, int* __restrict__ nn,
                                                       int* __restrict__ flags)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 si = s4[i];
    const int ti = (int)si.w;
    double xi0, xi1, xi2;
    s_to_x(bd.h, si.x, si.y, si.z, xi0, xi1, xi2);
    const int c0 = cell_coord(si.x, g.nc[0]), c1 = cell_coord(si.y, g.nc[1]), c2 = cell_coord(si.z, g.nc[2]);
    int cnt = 0;
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
                    const double rsq = rsq_exact(xi0, xi1, xi2, xj0, xj1, xj2);
                    if (rsq < g.cut_sk_sq[ti * g.ne + (int)sj.w]) {
                        if (cnt < maxn) list[(size_t)cnt * ldl + i] = (uint32_t)j | code;
                        ++cnt;
                    }
                }
            }
        }
    }
    nn[i] = cnt;
    if (cnt > maxn) atomicMax(&flags[0], cnt);
}

// ---------------- EAM ----------------

__device__ inline void pair_dx(const double4& xi, const double4& xj, uint32_t code, const double* __restrict__ h,
                               double& dx, double& dy, double& dz)
{
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


// knot interval and offset with the reference's clamps (src/ff_eam.cpp:117-128)
__device__ inline int knot(double v, double inv, int nknots, float& p)
{
    const double pm = v * inv;
    int m = (int)pm;
    m = m < nknots - 2 ? m : nknots - 2;
    m = m > 0 ? m : 0;
    const double f = pm - (double)m;
    p = (float)(f < 1.0 ? f : 1.0);
    return m;
}

__device__ inline void embed(const EamArgs& P, int t, double rho, double& F, double& dF)
{
    const double pm = rho * P.drho_inv;
    int m = (int)pm;
    m = m < P.nrho - 2 ? m : P.nrho - 2;
    m = m > 0 ? m : 0;
    double p = pm - (double)m;
    p = p < 1.0 ? p : 1.0;
    const double* c = P.Ftab + ((size_t)t * P.nrho + m) * 4;
    dF = ((3.0 * c[3] * p + 2.0 * c[2]) * p + c[1]) * P.drho_inv;
    F = ((c[3] * p + c[2]) * p + c[1]) * p + c[0];
    if (rho > P.rho_max) F += dF * (rho - P.rho_max);
}

// A batch of B neighbor entries is resolved in three stages -- all list
// entries, then all neighbor positions, then all spline rows -- so each thread
// has B independent loads in flight per stage instead of one dependent chain
// per neighbor (the passes are latency bound, not bandwidth bound).
struct PairGeo {
    double dx, dy, dz, r;
    int row;   // spline row (pair type * nr + knot), 0 when outside the cutoff
    float p;
    bool in;
};

// pass-2 record: position with (type, F'(rho)) packed into the fourth slot,
// so one gather brings everything pass 2 needs about a neighbor
__device__ inline double pack_type_dF(int t, float dF)
{
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)t << 32) | __float_as_uint(dF)));
}
__device__ inline int rec_type(double w) { return (int)((unsigned long long)__double_as_longlong(w) >> 32); }
__device__ inline float rec_dF(double w) { return __uint_as_float((unsigned)(__double_as_longlong(w) & 0xffffffffull)); }

__device__ inline PairGeo pair_geo(const double4& xi, int ti, uint32_t e, const double4& xj, int tj, const EamArgs& P,
                                   const BoxDev& bd)
{
    PairGeo g;
    pair_dx(xi, xj, e >> kIndexBits, bd.h, g.dx, g.dy, g.dz);
    const double rsq = g.dx * g.dx + g.dy * g.dy + g.dz * g.dz;
    const int tt = ti * P.ne + tj;
    g.in = rsq < P.cut_sq[tt];
    g.r = root64(rsq);
    const int m = knot(g.r, P.dr_inv, P.nr, g.p);
    g.row = g.in ? tt * P.nr + m : 0;
    return g;
}

// separation of two 16-byte pass-2 records with the image shift of the list entry
__device__ inline void rec16_dx(const uint4& ri, const uint4& rj, uint32_t code, const Rec16Frame& R, double& dx,
                                double& dy, double& dz)
{
    unsigned sx = 0u, sy = 0u, sz = 0u;
    if (code != (uint32_t)kZeroImage) {
        const int n2 = (int)(code / 25u) - 2;
        const int rem = (int)(code % 25u);
        const unsigned n1 = (unsigned)(rem / 5 - 2), n0 = (unsigned)(rem % 5 - 2), m2 = (unsigned)n2;
        sx = n0 * R.cH[0] + n1 * R.cH[3] + m2 * R.cH[6];
        sy = n1 * R.cH[4] + m2 * R.cH[7];
        sz = m2 * R.cH[8];
    }
    dx = (double)(int)(ri.x - rj.x - sx) * R.q;
    dy = (double)(int)(ri.y - rj.y - sy) * R.q;
    dz = (double)(int)(ri.z - rj.z - sz) * R.q;
}

// 16-byte position records for the density pass: the coordinates on the
// grid of the pass-2 records (32-bit integers modulo 2^32) and the element
// in the fourth word. Record differences are exact integers (rec16_dx).
__global__ __launch_bounds__(kBlock) void k_make_records(int n, const double4* __restrict__ x, Rec16Frame R,
                                                         uint4* __restrict__ rec)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 xi = x[i];
    rec[i] = make_uint4((unsigned)__double2ll_rn(xi.x * R.inv_q), (unsigned)__double2ll_rn(xi.y * R.inv_q),
                        (unsigned)__double2ll_rn(xi.z * R.inv_q), (unsigned)(int)xi.w);
}

// pass 1: rho_i = sum_j rho_{j->i}(r_ij) (FP64 sum of FP32 spline values),
// F(rho_i) and F'(rho_i); embedding energy reduced per workgroup. Neighbor
// positions come from the 16-byte records (one gather per neighbor).
template <int B>
__global__ __launch_bounds__(kBlock) void k_eam_rho(int n, const double4* __restrict__ x, const uint4* __restrict__ rec,
                                                    const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                    int ldl, EamArgs P, double* __restrict__ rho_out,
                                                    double4* __restrict__ xf, uint4* __restrict__ xr, Rec16Frame R16,
                                                    double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        const uint4 ri = rec[i];
        const int ti = (int)ri.w;
        const int cnt = nn[i];
        double rho = 0.0;
        for (int k0 = 0; k0 < cnt; k0 += B) {
            uint32_t e[B];
            uint4 rj[B];
            float4 c[B];
            PairGeo g[B];
#pragma unroll
            for (int b = 0; b < B; ++b) e[b] = k0 + b < cnt ? list[(size_t)(k0 + b) * ldl + i] : (uint32_t)i | (kZeroImage << kIndexBits);
#pragma unroll
            for (int b = 0; b < B; ++b) rj[b] = rec[e[b] & kIndexMask];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                rec16_dx(ri, rj[b], e[b] >> kIndexBits, R16, g[b].dx, g[b].dy, g[b].dz);
                const double rsq = g[b].dx * g[b].dx + g[b].dy * g[b].dy + g[b].dz * g[b].dz;
                const int tt = ti * P.ne + (int)rj[b].w;
                g[b].in = rsq < P.cut_sq[tt] && (k0 + b < cnt);
                g[b].r = root64(rsq);
                const int m = knot(g[b].r, P.dr_inv, P.nr, g[b].p);
                g[b].row = g[b].in ? tt * P.nr + m : 0;
            }
#pragma unroll
            for (int b = 0; b < B; ++b) c[b] = P.tabA[g[b].in ? g[b].row : 0];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const float p = g[b].p;
                const float v = ((c[b].w * p + c[b].z) * p + c[b].y) * p + c[b].x;
                if (g[b].in) rho += (double)v;
            }
        }
        double F, dF;
        embed(P, ti, rho, F, dF);
        rho_out[i] = rho;
        if (xr) {
            xr[i] = make_uint4(ri.x, ri.y, ri.z, __float_as_uint((float)dF));
        } else {
            const double4 xi = x[i];
            xf[i] = make_double4(xi.x, xi.y, xi.z, pack_type_dF(ti, (float)dF));
        }
        acc[0] = F;
    }
    block_sum_store<1>(acc, part + (size_t)blk * kPartW);
}

// second half kick of velocity Verlet, fused into the force pass: the owner
// of f_i updates v_i and adds m v (x) v to the kinetic tensor

// pass 2: pair + embedding forces on the owner only (no scatter), half of
// each full-list pair's energy and virial
template <int B, bool ONE, bool R16>
__global__ __launch_bounds__(kBlock) void k_eam_force(int n, const double4* __restrict__ x,
                                                      const uint4* __restrict__ xr, Rec16Frame RF,
                                                      const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                      int ldl, EamArgs P, BoxDev bd,
                                                      float4* __restrict__ f_out, double* __restrict__ part,
                                                      KickArgs K, double4* __restrict__ v, double* __restrict__ part_mvv)
{
    // x (or xr, for single-element tables) holds the pass-2 records written by the density pass
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    double mvv[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
        uint4 ri = make_uint4(0u, 0u, 0u, 0u);
        int ti = 0;
        float dFi;
        if (R16) {
            ri = xr[i];
            dFi = __uint_as_float(ri.w);
        } else {
            xi = x[i];
            ti = rec_type(xi.w);
            dFi = rec_dF(xi.w);
        }
        const int cnt = nn[i];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        for (int k0 = 0; k0 < cnt; k0 += B) {
            uint32_t e[B];
            double4 xj[B];
            uint4 rj[B];
            float dFj[B];
            float4 ca[B], cb[B], cp[B];
            PairGeo g[B];
#pragma unroll
            for (int b = 0; b < B; ++b) e[b] = k0 + b < cnt ? list[(size_t)(k0 + b) * ldl + i] : (uint32_t)i | (kZeroImage << kIndexBits);
            if (R16) {
#pragma unroll
                for (int b = 0; b < B; ++b) rj[b] = xr[e[b] & kIndexMask];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    rec16_dx(ri, rj[b], e[b] >> kIndexBits, RF, g[b].dx, g[b].dy, g[b].dz);
                    const double rsq = g[b].dx * g[b].dx + g[b].dy * g[b].dy + g[b].dz * g[b].dz;
                    g[b].in = rsq < P.cut_sq[0] && (k0 + b < cnt);
                    g[b].r = root64(rsq);
                    const int m = knot(g[b].r, P.dr_inv, P.nr, g[b].p);
                    g[b].row = g[b].in ? m : 0;
                    dFj[b] = __uint_as_float(rj[b].w);
                }
            } else {
#pragma unroll
                for (int b = 0; b < B; ++b) xj[b] = x[e[b] & kIndexMask];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    g[b] = pair_geo(xi, ti, e[b], xj[b], rec_type(xj[b].w), P, bd);
                    g[b].in = g[b].in && (k0 + b < cnt);
                    dFj[b] = rec_dF(xj[b].w);
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (ONE) {
                    // one element: rho_{j->i} = rho_{i->j}; row = (rho c1..c3, rphi c0), (rphi c1..c3, -)
                    const float4* c = P.tabC + (size_t)(g[b].in ? g[b].row : 0) * 2;
                    const float4 u = c[0], v = c[1];
                    ca[b] = make_float4(0.f, u.x, u.y, u.z);
                    cb[b] = ca[b];
                    cp[b] = make_float4(u.w, v.x, v.y, v.z);
                } else {
                    const float4* c = P.tabB + (size_t)(g[b].in ? g[b].row : 0) * 3;
                    ca[b] = c[0];
                    cb[b] = c[1];
                    cp[b] = c[2];
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (!g[b].in) continue;
                const float p = g[b].p;
                const float rinv = 1.0f / (float)g[b].r;
                const float d_in = ((3.f * ca[b].w * p + 2.f * ca[b].z) * p + ca[b].y) * P.dr_invf;
                const float d_out = ((3.f * cb[b].w * p + 2.f * cb[b].z) * p + cb[b].y) * P.dr_invf;
                const float z = ((cp[b].w * p + cp[b].z) * p + cp[b].y) * p + cp[b].x;
                const float zp = ((3.f * cp[b].w * p + 2.f * cp[b].z) * p + cp[b].y) * P.dr_invf;
                const float phi = z * rinv;
                const float dphi = (zp - phi) * rinv;
                const float fpair = -(dphi + dFi * d_in + dFj[b] * d_out) * rinv;
                const double dx = g[b].dx, dy = g[b].dy, dz = g[b].dz;
                fx += fpair * (float)dx;
                fy += fpair * (float)dy;
                fz += fpair * (float)dz;
                const double hf = 0.5 * (double)fpair;
                acc[0] += 0.5 * (double)phi;
                acc[1] += hf * dx * dx;
                acc[2] += hf * dx * dy;
                acc[3] += hf * dx * dz;
                acc[4] += hf * dy * dy;
                acc[5] += hf * dy * dz;
                acc[6] += hf * dz * dz;
            }
        }
        f_out[i] = make_float4(fx, fy, fz, 0.f);
        if (K.on) {
            double4 vi = v[i];
            const double m = K.mass[ti];
            vi.x += (double)fx * K.dt2 / m;
            vi.y += (double)fy * K.dt2 / m;
            vi.z += (double)fz * K.dt2 / m;
            v[i] = vi;
            mvv[0] = vi.x * vi.x * m;
            mvv[1] = vi.x * vi.y * m;
            mvv[2] = vi.x * vi.z * m;
            mvv[3] = vi.y * vi.y * m;
            mvv[4] = vi.y * vi.z * m;
            mvv[5] = vi.z * vi.z * m;
        }
    }
    block_sum_store<7>(acc, part + (size_t)blk * kPartW);
    if (K.on) {
        __syncthreads();
        block_sum_store<6>(mvv, part_mvv + (size_t)blk * kPartW);
    }
}

// value(): rho and pair energy in one pass, then F(rho)
__global__ __launch_bounds__(kBlock) void k_eam_energy(int n, const double4* __restrict__ x,
                                                       const int* __restrict__ nn, const uint32_t* __restrict__ list,
                                                       int ldl, EamArgs P, BoxDev bd, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        const double4 xi = x[i];
        const int ti = (int)xi.w;
        const int cnt = nn[i];
        double rho = 0.0, ep = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const uint32_t e = list[(size_t)k * ldl + i];
            const double4 xj = x[e & kIndexMask];
            double dx, dy, dz;
            pair_dx(xi, xj, e >> kIndexBits, bd.h, dx, dy, dz);
            const double rsq = dx * dx + dy * dy + dz * dz;
            const int tt = ti * P.ne + (int)xj.w;
            if (rsq < P.cut_sq[tt]) {
                const double r = root64(rsq);
                float p;
                const int m = knot(r, P.dr_inv, P.nr, p);
                const float4* c = P.tabB + ((size_t)tt * P.nr + m) * 3;
                const float4 ca = c[0], cp = c[2];
                rho += (double)(((ca.w * p + ca.z) * p + ca.y) * p + ca.x);
                ep += 0.5 * (double)((((cp.w * p + cp.z) * p + cp.y) * p + cp.x) / (float)r);
            }
        }
        double F, dF;
        embed(P, ti, rho, F, dF);
        acc[0] = F + ep;
    }
    block_sum_store<1>(acc, part + (size_t)blk * kPartW);
}

// ---------------- NVT ----------------

// nhc state layout (doubles)
enum { NH_FAC = 0, NH_T = 1, NH_ETA = 2, NH_MVV = 5, NH_SIZE = 16 };

struct NvtArgs {
    double dt, dt2, kB, T, ddt, ddt2, freq_sq;
    int ndof, natms;
    double mass[kMaxElems];
};

// v <- v*fac + f dt/2 / m ; x <- x + v dt ; partial sums of v dt
__global__ __launch_bounds__(kBlock) void k_kick_drift(int n, double4* __restrict__ x, double4* __restrict__ v,
                                                       const float4* __restrict__ f, NvtArgs A,
                                                       const double* __restrict__ nh, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double fac = nh[NH_FAC];
        double4 xi = x[i];
        double4 vi = v[i];
        const float4 fi = f[i];
        const double m = A.mass[(int)xi.w];
        vi.x = vi.x * fac + (double)fi.x * A.dt2 / m;
        vi.y = vi.y * fac + (double)fi.y * A.dt2 / m;
        vi.z = vi.z * fac + (double)fi.z * A.dt2 / m;
        acc[0] = vi.x * A.dt;
        acc[1] = vi.y * A.dt;
        acc[2] = vi.z * A.dt;
        xi.x += vi.x * A.dt;
        xi.y += vi.y * A.dt;
        xi.z += vi.z * A.dt;
        x[i] = xi;
        v[i] = vi;
    }
    block_sum_store<3>(acc, part + (size_t)blk * kPartW);
}

// remove the centre-of-mass drift of the step, then the skin check of
// DynamicMD::decide (|x - x0|^2 > skin^2 / 4 anywhere -> rebuild)
__global__ __launch_bounds__(kBlock) void k_shift_decide(int n, double4* __restrict__ x, const double4* __restrict__ x0,
                                                         const double* __restrict__ com, double inv_n, double skin_sq4,
                                                         int* __restrict__ flags)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    double4 xi = x[i];
    xi.x -= com[0] * inv_n;
    xi.y -= com[1] * inv_n;
    xi.z -= com[2] * inv_n;
    x[i] = xi;
    const double4 a = x0[i];
    if (rsq_exact(a.x, a.y, a.z, xi.x, xi.y, xi.z) > skin_sq4) atomicOr(&flags[1], 1);
}

// second half kick and the kinetic tensor sum m v (x) v
__global__ __launch_bounds__(kBlock) void k_kick2(int n, const double4* __restrict__ x, double4* __restrict__ v,
                                                  const float4* __restrict__ f, NvtArgs A, int kick,
                                                  double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double4 vi = v[i];
        const double m = A.mass[(int)x[i].w];
        if (kick) {
            const float4 fi = f[i];
            vi.x += (double)fi.x * A.dt2 / m;
            vi.y += (double)fi.y * A.dt2 / m;
            vi.z += (double)fi.z * A.dt2 / m;
            v[i] = vi;
        }
        acc[0] = vi.x * vi.x * m;
        acc[1] = vi.x * vi.y * m;
        acc[2] = vi.x * vi.z * m;
        acc[3] = vi.y * vi.y * m;
        acc[4] = vi.y * vi.z * m;
        acc[5] = vi.z * vi.z * m;
    }
    block_sum_store<6>(acc, part + (size_t)blk * kPartW);
}

__device__ inline double nhc_phi(double x)
{
    if (fabs(x) < 0.015625) return 1.0 + x / 2.0 + x * x / 6.0 + x * x * x / 24.0;
    return (exp(x) - 1.0) / x;
}

// One Nose-Hoover-chain (L=3, one iteration) half step on the scaled
// temperature; returns the velocity scale (src/thermostat.cpp:47-143)
__device__ inline double nhc_half(double* e, const NvtArgs& A, double T_r)
{
    double rescale = 1.0, prev, tmp;
    e[2] += A.ddt2 * (e[1] * e[1] - A.freq_sq);
    prev = -e[2];
    e[1] += A.ddt2 * nhc_phi(prev * A.ddt2) * (prev * e[1] + e[0] * e[0] - A.freq_sq);
    prev = -e[1];
    e[0] += A.ddt2 * nhc_phi(prev * A.ddt2) * (prev * e[0] + A.freq_sq * (T_r - 1.0));
    prev = -e[0];
    tmp = exp(A.ddt * prev);
    rescale *= tmp;
    T_r *= tmp * tmp;
    prev = (T_r - 1.0) * A.freq_sq;
    e[0] += A.ddt2 * nhc_phi(-e[1] * A.ddt2) * (-e[1] * e[0] + prev);
    prev = A.ndof * e[0] * e[0] - A.freq_sq;
    e[1] += A.ddt2 * nhc_phi(-e[2] * A.ddt2) * (-e[2] * e[1] + prev);
    prev = e[1] * e[1] - A.freq_sq;
    e[2] += A.ddt2 * prev;
    return rescale;
}

// mode 0: initial T from mvv, fac = 1; mode 1: thermostat at the start of a
// step (fac *= s); mode 2: T from the new mvv, thermostat after the step (fac = s)
__global__ void k_nhc(double* __restrict__ nh, const double* __restrict__ mvv, NvtArgs A, int mode)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (mode == 0 || mode == 2) {
        for (int k = 0; k < 6; ++k) nh[NH_MVV + k] = mvv[k];
        nh[NH_T] = (nh[NH_MVV] + (nh[NH_MVV + 3] + nh[NH_MVV + 5])) / (A.ndof * A.kB);
        if (mode == 0) {
            nh[NH_FAC] = 1.0;
            return;
        }
    }
    const double s = nhc_half(nh + NH_ETA, A, nh[NH_T] / A.T);
    if (mode == 1) nh[NH_FAC] *= s; else nh[NH_FAC] = s;
    const double s2 = s * s;
    for (int k = 0; k < 6; ++k) nh[NH_MVV + k] *= s2;
    nh[NH_T] *= s2;
}

__global__ __launch_bounds__(kBlock) void k_scale_v(int n, double4* __restrict__ v, const double* __restrict__ nh)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double fac = nh[NH_FAC];
    double4 vi = v[i];
    vi.x *= fac;
    vi.y *= fac;
    vi.z *= fac;
    v[i] = vi;
}

// ---------------- host <-> slot order ----------------

__global__ __launch_bounds__(kBlock) void k_scatter_x(int n, const double* __restrict__ xh, const int* __restrict__ slot_of,
                                                      double4* __restrict__ x, int keep_w)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const int s = slot_of[id];
    double4 o = x[s];
    o.x = xh[3 * (size_t)id];
    o.y = xh[3 * (size_t)id + 1];
    o.z = xh[3 * (size_t)id + 2];
    if (!keep_w) o.w = 0.0;
    x[s] = o;
}

__global__ __launch_bounds__(kBlock) void k_gather_d4(int n, const double4* __restrict__ a, const int* __restrict__ slot_of,
                                                      double* __restrict__ out)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const double4 o = a[slot_of[id]];
    out[3 * (size_t)id] = o.x;
    out[3 * (size_t)id + 1] = o.y;
    out[3 * (size_t)id + 2] = o.z;
}

__global__ __launch_bounds__(kBlock) void k_gather_f4(int n, const float4* __restrict__ a, const int* __restrict__ slot_of,
                                                      double* __restrict__ out)
{
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= n) return;
    const float4 o = a[slot_of[id]];
    out[3 * (size_t)id] = o.x;
    out[3 * (size_t)id + 1] = o.y;
    out[3 * (size_t)id + 2] = o.z;
}


}  // namespace

// ---------------- EamForceField ----------------

EamForceField::EamForceField(EamTables t) : t_(std::move(t))
{
    const int ne = t_.ne, nr = t_.nr, nrho = t_.nrho;
    MAPP_REQUIRE(ne >= 1 && ne <= kMaxElems, ERR_ARG, "1 to 8 elements are supported");
    std::vector<float4> A((size_t)ne * ne * nr), B((size_t)ne * ne * nr * 3);
    auto f4 = [](const double* c) { return make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[3]); };
    for (int a = 0; a < ne; ++a)
        for (int b = 0; b < ne; ++b) {
            const size_t tt = (size_t)a * ne + b;
            for (int m = 0; m < nr; ++m) {
                const float4 in = f4(t_.rho_at(b, a) + (size_t)m * 4);   // density at a due to b
                const float4 out = f4(t_.rho_at(a, b) + (size_t)m * 4);  // density at b due to a
                A[tt * nr + m] = in;
                B[(tt * nr + m) * 3 + 0] = in;
                B[(tt * nr + m) * 3 + 1] = out;
                B[(tt * nr + m) * 3 + 2] = f4(t_.rphi_at(a, b) + (size_t)m * 4);
            }
        }
    // one element: the density pair is symmetric, so a row needs only the
    // rho derivative coefficients and the four r*phi coefficients (32 B)
    std::vector<float4> Cc(ne == 1 ? (size_t)nr * 2 : 1);
    if (ne == 1)
        for (int m = 0; m < nr; ++m) {
            const double* r = t_.rho_at(0, 0) + (size_t)m * 4;
            const double* p = t_.rphi_at(0, 0) + (size_t)m * 4;
            Cc[(size_t)m * 2] = make_float4((float)r[1], (float)r[2], (float)r[3], (float)p[0]);
            Cc[(size_t)m * 2 + 1] = make_float4((float)p[1], (float)p[2], (float)p[3], 0.f);
        }
    tabC_.reserve(Cc.size());
    MAPP_HIP(hipMemcpy(tabC_.get(), Cc.data(), Cc.size() * sizeof(float4), hipMemcpyHostToDevice));
    std::vector<double> F((size_t)ne * nrho * 4);
    for (int a = 0; a < ne; ++a)
        for (int m = 0; m < nrho; ++m)
            for (int q = 0; q < 4; ++q) F[((size_t)a * nrho + m) * 4 + q] = t_.F_at(a)[(size_t)m * t_.nf() + q];
    tabA_.reserve(A.size());
    tabB_.reserve(B.size());
    Ftab_.reserve(F.size());
    MAPP_HIP(hipMemcpy(tabA_.get(), A.data(), A.size() * sizeof(float4), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(tabB_.get(), B.data(), B.size() * sizeof(float4), hipMemcpyHostToDevice));
    MAPP_HIP(hipMemcpy(Ftab_.get(), F.data(), F.size() * sizeof(double), hipMemcpyHostToDevice));
}

EamArgs EamForceField::args() const
{
    EamArgs P{};
    P.ne = t_.ne;
    P.nr = t_.nr;
    P.nrho = t_.nrho;
    P.dr_inv = 1.0 / t_.dr;
    P.drho_inv = 1.0 / t_.drho;
    P.rho_max = (double)t_.nrho * t_.drho;
    P.dr_invf = (float)P.dr_inv;
    for (int k = 0; k < t_.ne * t_.ne; ++k) P.cut_sq[k] = t_.rc[k] * t_.rc[k];
    P.tabA = tabA_.get();
    P.tabB = tabB_.get();
    P.tabC = tabC_.get();
    P.Ftab = Ftab_.get();
    return P;
}

// ---------------- MDSystem ----------------

MDSystem::MDSystem(std::shared_ptr<EamForceField> ff, int natms, const double* H, const double* x,
                   const uint8_t* elem, double skin)
    : ff_(std::move(ff)), n_(natms), skin_(skin)
{
    MAPP_REQUIRE(natms >= 1, ERR_ARG, "at least one atom is required");
    MAPP_REQUIRE((uint32_t)natms <= kIndexMask, ERR_CAPACITY, "at most 33554431 atoms per system");
    MAPP_REQUIRE(skin >= 0.0, ERR_ARG, "skin must be non-negative");
    for (int i = 0; i < 3; ++i)
        MAPP_REQUIRE(H[i * 3 + i] > 0.0, ERR_ARG, "H must be lower triangular with a positive diagonal");
    MAPP_REQUIRE(H[1] == 0.0 && H[2] == 0.0 && H[5] == 0.0, ERR_ARG, "H must be lower triangular");
    const int ne = ff_->tables().ne;
    for (int i = 0; i < natms; ++i) MAPP_REQUIRE(elem[i] < ne, ERR_ARG, "element index out of range");
    MAPP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    box_ = make_box(H);
    max_cut_ = ff_->tables().max_rc() + skin_;
    setup_grid();
    x_.reserve(n_); x_alt_.reserve(n_); s_.reserve(n_); x0_.reserve(n_);
    v_.reserve(n_); v_alt_.reserve(n_); id_.reserve(n_); id_alt_.reserve(n_); slot_of_.reserve(n_);
    f_.reserve(n_); rho_.reserve(n_); xf_.reserve(n_); xr_.reserve(n_); rec_.reserve(n_);
    keys_.reserve(n_); keys_alt_.reserve(n_); vals_.reserve(n_); vals_alt_.reserve(n_); s_tmp_.reserve(n_);
    nn_.reserve(n_);
    flags_.reserve(4);
    part_.reserve((size_t)3 * grid_for(n_) * kPartW + 64);  // rho, force, integrator rows
    red_.reserve(64);
    red2_.reserve(256 * kPartW);
    std::vector<double4> xh(n_);
    std::vector<int> idh(n_);
    for (int i = 0; i < n_; ++i) {
        xh[i] = make_double4(x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], (double)elem[i]);
        idh[i] = i;
    }
    MAPP_HIP(hipMemcpyAsync(x_.get(), xh.data(), sizeof(double4) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(id_.get(), idh.data(), sizeof(int) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemsetAsync(v_.get(), 0, sizeof(double4) * n_, stream_));
    MAPP_HIP(hipMemsetAsync(f_.get(), 0, sizeof(float4) * n_, stream_));
    rebuild();
}

MDSystem::~MDSystem()
{
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void MDSystem::sync() { MAPP_HIP(hipStreamSynchronize(stream_)); }

// two-level fixed-order sum of partial rows: up to 256 blocks, then one
void MDSystem::reduce_rows(const double* part, int nrows, int ncols, double* out)
{
    const int g = std::min(256, std::max(1, (nrows + kBlock - 1) / kBlock));
    const int chunk = (nrows + g - 1) / g;
    hipLaunchKernelGGL(k_reduce_rows, dim3(g), dim3(kBlock), 0, stream_, nrows, ncols, part, red2_.get(), chunk);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kBlock), 0, stream_, g, ncols, red2_.get(), out, g);
}

void MDSystem::set_batch(int rho, int force)
{
    MAPP_REQUIRE(rho == 1 || rho == 4 || rho == 8, ERR_ARG, "rho batch must be 1, 4 or 8");
    MAPP_REQUIRE(force == 1 || force == 2 || force == 4 || force == 8, ERR_ARG, "force batch must be 1, 2, 4 or 8");
    batch_rho_ = rho;
    batch_force_ = force;
}

// cells of at least max_cut in every perpendicular width; boxes thinner than
// max_cut are searched over +/-R images (R <= 2 fits the 7-bit image code)
void MDSystem::setup_grid()
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
        MAPP_REQUIRE(R <= 2, ERR_CAPACITY,
                     "box narrower than half the cutoff+skin in some direction; replicate the cell first");
        grid_.R[d] = R;
        ncells *= nc;
    }
    grid_.ncells = ncells;
    // 16-byte position records: separations up to 4 (cut + skin) stay below 2^31 q
    r16_.q = std::ldexp(1.0, (int)std::ceil(std::log2(4.0 * max_cut_ / std::ldexp(1.0, 31))));
    r16_.inv_q = 1.0 / r16_.q;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d)
            r16_.cH[3 * k + d] = (unsigned)(unsigned long long)std::llrint(box_.H[k][d] * r16_.inv_q);
    // Morton rank of every cell so that cell-sorted atoms are spatially compact
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
    // neighbors expected in the search sphere at this density, with margin
    const double dens = n_ / box_.vol;
    const double sphere = 4.0 / 3.0 * M_PI * max_cut_ * max_cut_ * max_cut_;
    maxn_ = std::max(maxn_, (int)(1.25 * dens * sphere) + 16);
}

void MDSystem::rebuild()
{
    const int nb = grid_for(n_);
    GridArgs g{};
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = grid_.nc[d];
        g.R[d] = grid_.R[d];
    }
    const EamTables& t = ff_->tables();
    g.ne = t.ne;
    for (int a = 0; a < t.ne; ++a)
        for (int b = 0; b < t.ne; ++b) {
            const double c = t.rc[(size_t)a * t.ne + b] + skin_;
            g.cut_sk_sq[a * t.ne + b] = c * c;
        }
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    prof_.begin(K_CELL_SORT, stream_);
    hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), bd, g, cell_rank_.get(),
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
    hipLaunchKernelGGL(k_permute, dim3(nb), dim3(kBlock), 0, stream_, n_, vals_alt_.get(), s_tmp_.get(), v_.get(),
                       id_.get(), bd, x_alt_.get(), v_alt_.get(), id_alt_.get(), s_.get(), x0_.get(), slot_of_.get());
    std::swap(x_.p, x_alt_.p);
    std::swap(v_.p, v_alt_.p);
    std::swap(id_.p, id_alt_.p);
    hipLaunchKernelGGL(k_cell_start, dim3((grid_.ncells + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_,
                       grid_.ncells, keys_alt_.get(), cell_start_.get());
    prof_.end(stream_);
    build_list();
    ++nrebuild_;
}

void MDSystem::build_list()
{
    const int nb = grid_for(n_);
    GridArgs g{};
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = grid_.nc[d];
        g.R[d] = grid_.R[d];
    }
    const EamTables& t = ff_->tables();
    g.ne = t.ne;
    for (int a = 0; a < t.ne; ++a)
        for (int b = 0; b < t.ne; ++b) {
            const double c = t.rc[(size_t)a * t.ne + b] + skin_;
            g.cut_sk_sq[a * t.ne + b] = c * c;
        }
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    ldl_ = n_;
    for (int attempt = 0; attempt < 2; ++attempt) {
        list_.reserve((size_t)maxn_ * ldl_);
        MAPP_HIP(hipMemsetAsync(flags_.get(), 0, sizeof(int), stream_));
        prof_.begin(K_BUILD_LIST, stream_);
        hipLaunchKernelGGL(k_build_list, dim3(nb), dim3(kBlock), 0, stream_, n_, s_.get(), bd, g, cell_rank_.get(),
                           cell_start_.get(), maxn_, ldl_, list_.get(), nn_.get(), flags_.get());
        prof_.end(stream_);
        int over = 0;
        MAPP_HIP(hipMemcpyAsync(&over, flags_.get(), sizeof(int), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        if (over <= maxn_) return;
        maxn_ = over + over / 8 + 8;
    }
    throw Error(ERR_STATE, "neighbor list capacity did not converge");
}

void MDSystem::force_passes(const KickArgs* kick)
{
    const int nb = grid_for(n_);
    const EamArgs P = ff_->args();
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    MAPP_HIP(hipMemsetAsync(flags_.get() + 2, 0, sizeof(int), stream_));
    prof_.begin(K_EAM_RHO, stream_);
    hipLaunchKernelGGL(k_make_records, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), r16_, rec_.get());
    const bool one = P.ne == 1;
    uint4* xr = one ? xr_.get() : nullptr;
    switch (batch_rho_) {
        case 1: hipLaunchKernelGGL(k_eam_rho<1>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), rec_.get(), nn_.get(), list_.get(), ldl_, P, rho_.get(), xf_.get(), xr, r16_, part_.get()); break;
        case 4: hipLaunchKernelGGL(k_eam_rho<4>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), rec_.get(), nn_.get(), list_.get(), ldl_, P, rho_.get(), xf_.get(), xr, r16_, part_.get()); break;
        default: hipLaunchKernelGGL(k_eam_rho<8>, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), rec_.get(), nn_.get(), list_.get(), ldl_, P, rho_.get(), xf_.get(), xr, r16_, part_.get()); break;
    }
    prof_.end(stream_);
    prof_.begin(K_EAM_FORCE, stream_);
    double* p2 = part_.get() + (size_t)nb * kPartW;
    KickArgs K{};
    if (kick) {
        K = *kick;
        K.on = 1;
    }
    double* p3 = part_.get() + (size_t)2 * nb * kPartW;
#define MAPP_FORCE(B, ONE) hipLaunchKernelGGL((k_eam_force<B, ONE, ONE>), dim3(nb), dim3(kBlock), 0, stream_, n_, \
                                              xf_.get(), xr_.get(), r16_, nn_.get(), list_.get(), ldl_, P, bd, \
                                              f_.get(), p2, K, v_.get(), p3)
    switch (batch_force_) {
        case 1: if (one) MAPP_FORCE(1, true); else MAPP_FORCE(1, false); break;
        case 2: if (one) MAPP_FORCE(2, true); else MAPP_FORCE(2, false); break;
        case 8: if (one) MAPP_FORCE(8, true); else MAPP_FORCE(8, false); break;
        default: if (one) MAPP_FORCE(4, true); else MAPP_FORCE(4, false); break;
    }
#undef MAPP_FORCE
    prof_.end(stream_);
    reduce_rows(part_.get(), 2 * nb, 7, red_.get());
}

MDReport MDSystem::reduce_report(int)
{
    MDReport r{};
    MAPP_HIP(hipMemcpyAsync(r.e, red_.get(), sizeof(double) * 7, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    return r;
}

MDReport MDSystem::force_calc()
{
    force_passes(nullptr);
    MDReport r = reduce_report(7);
    prof_.flush();
    return r;
}

double MDSystem::value()
{
    const int nb = grid_for(n_);
    const EamArgs P = ff_->args();
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    prof_.begin(K_EAM_ENERGY, stream_);
    hipLaunchKernelGGL(k_eam_energy, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), nn_.get(), list_.get(), ldl_,
                       P, bd, part_.get());
    prof_.end(stream_);
    reduce_rows(part_.get(), nb, 1, red_.get());
    double e = 0.0;
    MAPP_HIP(hipMemcpyAsync(&e, red_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    return e;
}

bool MDSystem::decide()
{
    const int nb = grid_for(n_);
    MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
    // no centre-of-mass shift outside the integrator: com sums are zero
    MAPP_HIP(hipMemsetAsync(red_.get() + 16, 0, sizeof(double) * 3, stream_));
    hipLaunchKernelGGL(k_shift_decide, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), x0_.get(), red_.get() + 16,
                       1.0 / n_, 0.25 * skin_ * skin_, flags_.get());
    int flag = 0;
    MAPP_HIP(hipMemcpyAsync(&flag, flags_.get() + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    return flag == 0;
}

bool MDSystem::update()
{
    if (decide()) return false;
    rebuild();
    return true;
}

void MDSystem::set_x(const double* xh)
{
    DevBuf<double> tmp;
    tmp.reserve((size_t)3 * n_);
    MAPP_HIP(hipMemcpyAsync(tmp.get(), xh, sizeof(double) * 3 * n_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_scatter_x, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, tmp.get(),
                       slot_of_.get(), x_.get(), 1);
    sync();
}

void MDSystem::get_x(double* xh)
{
    DevBuf<double> tmp;
    tmp.reserve((size_t)3 * n_);
    hipLaunchKernelGGL(k_gather_d4, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, x_.get(),
                       slot_of_.get(), tmp.get());
    MAPP_HIP(hipMemcpyAsync(xh, tmp.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
    sync();
}

void MDSystem::set_v(const double* vh)
{
    DevBuf<double> tmp;
    tmp.reserve((size_t)3 * n_);
    MAPP_HIP(hipMemcpyAsync(tmp.get(), vh, sizeof(double) * 3 * n_, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_scatter_x, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, tmp.get(),
                       slot_of_.get(), v_.get(), 0);
    sync();
}

void MDSystem::get_v(double* vh)
{
    DevBuf<double> tmp;
    tmp.reserve((size_t)3 * n_);
    hipLaunchKernelGGL(k_gather_d4, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, v_.get(),
                       slot_of_.get(), tmp.get());
    MAPP_HIP(hipMemcpyAsync(vh, tmp.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
    sync();
}

void MDSystem::get_f(double* fh)
{
    DevBuf<double> tmp;
    tmp.reserve((size_t)3 * n_);
    hipLaunchKernelGGL(k_gather_f4, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, n_, f_.get(),
                       slot_of_.get(), tmp.get());
    MAPP_HIP(hipMemcpyAsync(fh, tmp.get(), sizeof(double) * 3 * n_, hipMemcpyDeviceToHost, stream_));
    sync();
}

void MDSystem::get_rho(double* rh)
{
    std::vector<double> r(n_);
    std::vector<int> id(n_);
    MAPP_HIP(hipMemcpyAsync(r.data(), rho_.get(), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipMemcpyAsync(id.data(), id_.get(), sizeof(int) * n_, hipMemcpyDeviceToHost, stream_));
    sync();
    for (int s = 0; s < n_; ++s) rh[id[s]] = r[s];
}

void MDSystem::list_stats(int* maxn, int64_t* nfull)
{
    std::vector<int> nn(n_);
    MAPP_HIP(hipMemcpyAsync(nn.data(), nn_.get(), sizeof(int) * n_, hipMemcpyDeviceToHost, stream_));
    sync();
    int m = 0;
    int64_t s = 0;
    for (int v : nn) {
        m = std::max(m, v);
        s += v;
    }
    *maxn = m;
    *nfull = s;
}

namespace {
struct HostList {
    std::vector<int> nn, id;
    std::vector<uint32_t> list;
};
HostList fetch_list(const int* nn_d, const int* id_d, const uint32_t* list_d, int n, int maxn, int ldl, hipStream_t st)
{
    HostList h;
    h.nn.resize(n);
    h.id.resize(n);
    h.list.resize((size_t)maxn * ldl);
    MAPP_HIP(hipMemcpyAsync(h.nn.data(), nn_d, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    MAPP_HIP(hipMemcpyAsync(h.id.data(), id_d, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    MAPP_HIP(hipMemcpyAsync(h.list.data(), list_d, sizeof(uint32_t) * h.list.size(), hipMemcpyDeviceToHost, st));
    MAPP_HIP(hipStreamSynchronize(st));
    return h;
}
inline void decode_img(uint32_t code, int8_t* img)
{
    img[2] = (int8_t)((int)(code / 25) - 2);
    img[1] = (int8_t)((int)((code % 25) / 5) - 2);
    img[0] = (int8_t)((int)(code % 5) - 2);
}
}  // namespace

// The reference keeps i<j for pairs inside the box and every ghost pair
// (src/neighbor_md.cpp:135): from the full list keep (i,j,n) when n != 0 or
// id_i < id_j.
int64_t MDSystem::half_pairs_count()
{
    HostList h = fetch_list(nn_.get(), id_.get(), list_.get(), n_, maxn_, ldl_, stream_);
    int64_t c = 0;
    for (int s = 0; s < n_; ++s)
        for (int k = 0; k < h.nn[s]; ++k) {
            const uint32_t e = h.list[(size_t)k * ldl_ + s];
            const uint32_t code = e >> kIndexBits;
            if (code != (uint32_t)kZeroImage || h.id[s] < h.id[e & kIndexMask]) ++c;
        }
    return c;
}

void MDSystem::half_pairs(int32_t* pi, int32_t* pj, int8_t* img)
{
    HostList h = fetch_list(nn_.get(), id_.get(), list_.get(), n_, maxn_, ldl_, stream_);
    int64_t c = 0;
    for (int s = 0; s < n_; ++s)
        for (int k = 0; k < h.nn[s]; ++k) {
            const uint32_t e = h.list[(size_t)k * ldl_ + s];
            const uint32_t code = e >> kIndexBits;
            const int j = h.id[e & kIndexMask];
            if (code != (uint32_t)kZeroImage || h.id[s] < j) {
                pi[c] = h.id[s];
                pj[c] = j;
                decode_img(code, img + 3 * c);
                ++c;
            }
        }
}

int MDSystem::nvt_run(const double* masses, double kB, double T, double dt, int nsteps, MDReport* rep, double* T_part)
{
    MAPP_REQUIRE(nsteps >= 0 && dt > 0.0 && T > 0.0 && kB > 0.0, ERR_ARG, "invalid NVT parameters");
    const int nb = grid_for(n_);
    const EamTables& t = ff_->tables();
    NvtArgs A{};
    A.dt = dt;
    A.dt2 = dt / 2.0;
    A.kB = kB;
    A.T = T;
    // ThermostatNHC(dt/2, t_relax = 100 dt, L = 3, niters = 1) (src/md_nvt.cpp:19)
    A.ddt = dt / 2.0;
    A.ddt2 = 0.5 * (dt / 2.0);
    A.freq_sq = 1.0 / ((100.0 * dt) * (100.0 * dt));
    A.ndof = 3 * n_;
    A.natms = n_;
    for (int a = 0; a < t.ne; ++a) {
        MAPP_REQUIRE(masses[a] > 0.0, ERR_ARG, "masses must be positive");
        A.mass[a] = masses[a];
    }
    double* nh = red_.get() + 32;  // NH_SIZE doubles of thermostat state
    double* com = red_.get() + 16;
    double* mvv = red_.get() + 8;
    MAPP_HIP(hipMemsetAsync(nh, 0, sizeof(double) * NH_SIZE, stream_));
    const int rebuilds0 = nrebuild_;
    force_passes(nullptr);
    hipLaunchKernelGGL(k_kick2, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), A, 0,
                       part_.get() + (size_t)2 * nb * kPartW);
    reduce_rows(part_.get() + (size_t)2 * nb * kPartW, nb, 6, mvv);
    hipLaunchKernelGGL(k_nhc, dim3(1), dim3(64), 0, stream_, nh, mvv, A, 0);
    const double skin_sq4 = 0.25 * skin_ * skin_;
    KickArgs kick{};
    kick.dt2 = A.dt2;
    for (int a = 0; a < t.ne; ++a) kick.mass[a] = A.mass[a];
    int flag = 0;
    for (int step = 0; step < nsteps; ++step) {
        prof_.begin(K_INTEGRATE, stream_);
        hipLaunchKernelGGL(k_nhc, dim3(1), dim3(64), 0, stream_, nh, mvv, A, 1);
        hipLaunchKernelGGL(k_kick_drift, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), A, nh,
                           part_.get() + (size_t)2 * nb * kPartW);
        reduce_rows(part_.get() + (size_t)2 * nb * kPartW, nb, 3, com);
        MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
        hipLaunchKernelGGL(k_shift_decide, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), x0_.get(), com,
                           1.0 / n_, skin_sq4, flags_.get());
        prof_.end(stream_);
        MAPP_HIP(hipMemcpyAsync(&flag, flags_.get() + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        if (flag) rebuild();
        force_passes(&kick);
        prof_.begin(K_INTEGRATE, stream_);
        reduce_rows(part_.get() + (size_t)2 * nb * kPartW, nb, 6, mvv);
        hipLaunchKernelGGL(k_nhc, dim3(1), dim3(64), 0, stream_, nh, mvv, A, 2);
        prof_.end(stream_);
    }
    hipLaunchKernelGGL(k_scale_v, dim3(nb), dim3(kBlock), 0, stream_, n_, v_.get(), nh);
    double nhh[NH_SIZE];
    MAPP_HIP(hipMemcpyAsync(nhh, nh, sizeof(double) * NH_SIZE, hipMemcpyDeviceToHost, stream_));
    MDReport r = reduce_report(7);
    prof_.flush();
    if (rep) *rep = r;
    if (T_part) *T_part = nhh[NH_T];
    return nrebuild_ - rebuilds0;
}

}  // namespace mapp
