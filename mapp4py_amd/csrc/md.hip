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

#include <mutex>

#include "md.hpp"
#include "cfg.hpp"
#include "kern.hpp"

// Scratch rows of MDSystem::value_noalloc: the forces (float4, first
// kNoAllocMaxAtoms * 16 bytes) and densities (double, after them) that the two
// EAM passes write and nobody reads. It is a device symbol, so every GPU
// holds its own copy, at its own address, from the library's load to the end
// of the process; host code must use that address (scratch_rows below), never
// this name.
static __device__ __align__(256) unsigned char scratch[(size_t)mapp::kNoAllocMaxAtoms * (16 + 8)];

namespace mapp {

namespace {

struct GridArgs {
    int nc[3];
    int R[3];
    int ne;
    double cut_sk_sq[kMaxElems * kMaxElems];
    // on the record grid: pairs with rsq below rec_in_sq are inside the list
    // cut-off for certain, above rec_out_sq outside (the band between holds
    // every pair the record quantisation could misplace)
    double rec_in_sq[kMaxElems * kMaxElems], rec_out_sq[kMaxElems * kMaxElems];
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
                                                    BoxDev bd, Rec16Frame R, double4* __restrict__ x_new,
                                                    double4* __restrict__ v_new, int* __restrict__ id_new,
                                                    double4* __restrict__ s4, double4* __restrict__ x0,
                                                    int* __restrict__ slot_of, uint4* __restrict__ rec)
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
    // the 16-byte record of the new slot (coordinates on the record grid, element)
    rec[p] = make_uint4((unsigned)__double2ll_rn(xn.x * R.inv_q), (unsigned)__double2ll_rn(xn.y * R.inv_q),
                        (unsigned)__double2ll_rn(xn.z * R.inv_q), (unsigned)(int)xn.w);
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

// The stencil of every cell in increasing Morton rank of its 27 neighbor cells
// (a 3x3x3 stencil of distinct cells): order[cell][q] is the stencil index
// t = (o0+1) + 3 (o1+1) + 9 (o2+1) of the q-th cell to visit, 27 bytes in a
// 32-byte row. It depends only on the grid, so it is made once per grid.
__global__ __launch_bounds__(kBlock) void k_cell_order(int ncells, GridArgs g, const int* __restrict__ cell_rank,
                                                       uint4* __restrict__ order)
{
    const int w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= ncells) return;
    const int c0 = w % g.nc[0], c1 = (w / g.nc[0]) % g.nc[1], c2 = w / (g.nc[0] * g.nc[1]);
    // key = rank * 32 + stencil index; 27 rounds of "smallest key above the last"
    uint32_t key[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int w0 = (c0 + t % 3 - 1 + g.nc[0]) % g.nc[0];
        const int w1 = (c1 + t / 3 % 3 - 1 + g.nc[1]) % g.nc[1];
        const int w2 = (c2 + t / 9 - 1 + g.nc[2]) % g.nc[2];
        key[t] = ((uint32_t)cell_rank[w0 + g.nc[0] * (w1 + g.nc[1] * w2)] << 5) | (uint32_t)t;
    }
    uint32_t word[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    uint32_t last = 0u;
#pragma unroll
    for (int round = 0; round < 27; ++round) {
        uint32_t best = 0xffffffffu;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const uint32_t k = key[t];
            best = ((round == 0 || k > last) && k < best) ? k : best;
        }
        last = best;
        word[round >> 2] |= (best & 31u) << (8 * (round & 3));
    }
    order[2 * (size_t)w] = make_uint4(word[0], word[1], word[2], word[3]);
    order[2 * (size_t)w + 1] = make_uint4(word[4], word[5], word[6], word[7]);
}

// Full neighbor list with the reference's FP64 predicate: x_i = s_i H and the
// image of j at (s_j + n) H, rsq = dx0^2 + (dx1^2 + dx2^2) < (r_c + skin)^2.
// Layout: entries k .. k+3 of slot i form one 16-byte block at
// list4[(k / 4) * ldl + i], so a wave's 64 atoms read or write one block
// each as 1 KB of contiguous memory. Each lane gathers its entries in
// registers and stores whole blocks: a block is written once, by one 16-byte
// store, instead of four scattered 4-byte stores that leave lines partly
// written in L2. With the per-cell stencil order (order != nullptr) the
// cells are visited in increasing Morton rank, so every row lists j in
// increasing slot order: the k-th entries of a wave's 64 spatially adjacent
// atoms then lie close together, and the passes' neighbor gathers touch few
// cache lines. Candidates are screened on the 16-byte records: one gather
// each, exact integer separations. Only a pair within the quantisation band
// of the cut-off (a few 1e-8 A wide) is decided by the FP64 predicate on s,
// so the neighbor set is that of the predicate for every pair.
// ONE: a single element, so the cut-offs are kernel-argument scalars; with
// several, the element pairs' cut-offs are read from LDS (an indexed
// kernel-argument array would cost a memory round trip per candidate).
template <bool ONE>
__global__ __launch_bounds__(kBlock) void k_build_list(int n, const double4* __restrict__ s4,
                                                       const uint4* __restrict__ rec, Rec16Frame R, BoxDev bd,
                                                       GridArgs g, const int* __restrict__ cell_rank,
                                                       const int* __restrict__ cell_start,
                                                       const uint4* __restrict__ order, int maxb, int ldl,
                                                       uint4* __restrict__ list4, int* __restrict__ nn,
                                                       int* __restrict__ flags)
{
    __shared__ double cut_lds[3][kMaxElems * kMaxElems];
    if (!ONE) {
        for (int k = threadIdx.x; k < g.ne * g.ne; k += kBlock) {
            cut_lds[0][k] = g.rec_in_sq[k];
            cut_lds[1][k] = g.rec_out_sq[k];
            cut_lds[2][k] = g.cut_sk_sq[k];
        }
        __syncthreads();
    }
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 si = s4[i];
    const uint4 ri = rec[i];
    const int ti = (int)si.w;
    double xi0, xi1, xi2;
    s_to_x(bd.h, si.x, si.y, si.z, xi0, xi1, xi2);
    const int c0 = cell_coord(si.x, g.nc[0]), c1 = cell_coord(si.y, g.nc[1]), c2 = cell_coord(si.z, g.nc[2]);
    int cnt = 0;
    uint32_t p0 = 0u, p1 = 0u, p2 = 0u, p3 = 0u;  // the block being filled
    // candidates of the stencil cell at offset (o0, o1, o2)
    auto scan = [&](int o0, int o1, int o2) {
        const int u2 = c2 + o2, n2 = floordiv(u2, g.nc[2]), w2 = u2 - n2 * g.nc[2];
        const int u1 = c1 + o1, n1 = floordiv(u1, g.nc[1]), w1 = u1 - n1 * g.nc[1];
        const int u0 = c0 + o0, n0 = floordiv(u0, g.nc[0]), w0 = u0 - n0 * g.nc[0];
        const int rank = cell_rank[w0 + g.nc[0] * (w1 + g.nc[1] * w2)];
        const int jb = cell_start[rank], je = cell_start[rank + 1];
        const bool self_img = (n0 | n1 | n2) == 0;
        const uint32_t img = (uint32_t)((n0 + 2) + 5 * (n1 + 2) + 25 * (n2 + 2));
        const uint32_t code = img << kIndexBits;
        const uint4 sh = R.shift[img];
        const uint4 rs = make_uint4(ri.x - sh.x, ri.y - sh.y, ri.z - sh.z, 0u);
        // the next candidate's record is loaded one candidate ahead
        uint4 rn = rec[jb < je ? jb : i];
        for (int j = jb; j < je; ++j) {
            const uint4 rj = rn;
            rn = rec[j + 1 < je ? j + 1 : j];
            if (self_img && j == i) continue;
            const double dx = (double)(int)(rs.x - rj.x), dy = (double)(int)(rs.y - rj.y),
                         dz = (double)(int)(rs.z - rj.z);
            const double rsq_u = dx * dx + dy * dy + dz * dz;
            const int tt = ONE ? 0 : ti * g.ne + (int)rj.w;
            bool in = rsq_u < (ONE ? g.rec_in_sq[0] : cut_lds[0][tt]);
            if (!in && rsq_u <= (ONE ? g.rec_out_sq[0] : cut_lds[1][tt])) {
                const double4 sj = s4[j];
                double xj0, xj1, xj2;
                s_to_x(bd.h, shift_unit(sj.x, n0), shift_unit(sj.y, n1), shift_unit(sj.z, n2), xj0, xj1, xj2);
                in = rsq_exact(xi0, xi1, xi2, xj0, xj1, xj2) < (ONE ? g.cut_sk_sq[0] : cut_lds[2][tt]);
            }
            if (in) {
                const uint32_t e = (uint32_t)j | code;
                const int q = cnt & 3;
                p0 = q == 0 ? e : p0;
                p1 = q == 1 ? e : p1;
                p2 = q == 2 ? e : p2;
                p3 = q == 3 ? e : p3;
                ++cnt;
                if (q == 3 && (cnt >> 2) <= maxb) list4[(size_t)((cnt >> 2) - 1) * ldl + i] = make_uint4(p0, p1, p2, p3);
            }
        }
    };
    if (order) {
        const size_t w = (size_t)c0 + (size_t)g.nc[0] * (c1 + (size_t)g.nc[1] * c2);
        const uint4 oa = order[2 * w], ob = order[2 * w + 1];
        uint32_t cur = oa.x;
        for (int q = 0; q < 27; ++q) {
            const int t = (int)(cur & 31u);
            scan(t % 3 - 1, t / 3 % 3 - 1, t / 9 - 1);
            cur >>= 8;
            if ((q & 3) == 3) {
                const int nq = (q >> 2) + 1;
                cur = nq == 1 ? oa.y : nq == 2 ? oa.z : nq == 3 ? oa.w : nq == 4 ? ob.x : nq == 5 ? ob.y : ob.z;
            }
        }
    } else {
        for (int o2 = -g.R[2]; o2 <= g.R[2]; ++o2)
            for (int o1 = -g.R[1]; o1 <= g.R[1]; ++o1)
                for (int o0 = -g.R[0]; o0 <= g.R[0]; ++o0) scan(o0, o1, o2);
    }
    // the last, partly filled block (its entries past cnt are never read)
    if ((cnt & 3) != 0 && (cnt >> 2) < maxb) list4[(size_t)(cnt >> 2) * ldl + i] = make_uint4(p0, p1, p2, p3);
    nn[i] = cnt;
    if (cnt > 4 * maxb) atomicMax(&flags[0], cnt);
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
    // v >= 0 here (a distance), so the reference's lower clamp m >= 0 holds
    const double pm = v * inv;
    int m = (int)pm;
    m = m < nknots - 2 ? m : nknots - 2;
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

// A batch of B list entries is resolved in three stages -- all list entries,
// then all neighbor records, then all spline rows -- so each thread has B
// independent loads in flight per stage instead of one dependent chain per
// neighbor. Every lane runs whole batches without branches: entries past the
// end of its row resolve to the atom itself and their contributions are
// selected away, and the passes are bound by instruction issue, so the pair
// arithmetic is kept to what the result needs (FP32 splines and reciprocal
// root, FP64 where a sum or the spline abscissa needs it).

// entries k0 .. k0+G-1 of slot i, G = 1, 2 or a multiple of 4 and k0 a
// multiple of G: the G/4 16-byte blocks, or the part of one block that holds
// them. The loads are unconditional (a block past the capacity maxb is never
// read: the last block is read again instead) and entries past the row end
// become the atom itself, zero image.
template <int G>
__device__ inline void list_entries(const uint4* __restrict__ list4, int ldl, int maxb, int k0, int i, int cnt,
                                    uint32_t (&e)[G])
{
    const uint32_t self = (uint32_t)i | ((uint32_t)kZeroImage << kIndexBits);
    if constexpr (G >= 4) {
#pragma unroll
        for (int q = 0; q < G / 4; ++q) {
            const int kb = (k0 >> 2) + q;
            const uint4 v = list4[(size_t)(kb < maxb ? kb : maxb - 1) * ldl + i];
            const int k = k0 + 4 * q;
            e[4 * q] = k < cnt ? v.x : self;
            e[4 * q + 1] = k + 1 < cnt ? v.y : self;
            e[4 * q + 2] = k + 2 < cnt ? v.z : self;
            e[4 * q + 3] = k + 3 < cnt ? v.w : self;
        }
    } else {
        const int kb = (k0 >> 2) < maxb ? (k0 >> 2) : maxb - 1;
        const uint32_t* blk = (const uint32_t*)(list4 + (size_t)kb * ldl + i) + (k0 & 3);
        if constexpr (G == 2) {
            const uint2 v = *(const uint2*)blk;
            e[0] = k0 < cnt ? v.x : self;
            e[1] = k0 + 1 < cnt ? v.y : self;
        } else {
            e[0] = k0 < cnt ? blk[0] : self;
        }
    }
}

// whether any lane of the wave holds a shifted periodic image among B entries
template <int B>
__device__ inline bool wave_has_image(const uint32_t (&e)[B])
{
    bool img = false;
#pragma unroll
    for (int b = 0; b < B; ++b) img |= (e[b] >> kIndexBits) != (uint32_t)kZeroImage;
    return __ballot(img) != 0ull;
}

// Geometry of one list entry on the record grid: the separation in grid
// units (exact integers), 1/r (FP32), and the spline interval and offset of r.
// Every length is that of Angstrom scaled by the power of two 1/q, so the
// cut-off test and the interval are those of the Angstrom computation.
struct Geo16 {
    double dx, dy, dz;
    float rinv;
    float p;
    int m;
    bool in;
};

// any_img: some lane of the wave holds a shifted image in this batch (a
// wave-uniform branch; most waves of a large box never take it)
__device__ inline Geo16 geo16(const uint4& ri, const uint4& rj, uint32_t e, bool any_img, const Rec16Frame& R,
                              double cut_sq_u, int nr)
{
    Geo16 g;
    unsigned sx = 0u, sy = 0u, sz = 0u;
    if (any_img) {
        // entry 62 of the table is the zero shift
        const uint4 s = R.shift[e >> kIndexBits];
        sx = s.x;
        sy = s.y;
        sz = s.z;
    }
    g.dx = (double)(int)(ri.x - rj.x - sx);
    g.dy = (double)(int)(ri.y - rj.y - sy);
    g.dz = (double)(int)(ri.z - rj.z - sz);
    const double rsq = g.dx * g.dx + g.dy * g.dy + g.dz * g.dz;
    g.in = rsq < cut_sq_u;
    // r to ~1e-14: FP32 reciprocal root, one FP64 Newton step for the root
    const float sf = (float)rsq;
    g.rinv = __builtin_amdgcn_rsqf(sf);
    const double r0 = (double)(sf * g.rinv);
    const double r = fma(fma(-r0, r0, rsq), (double)(0.5f * g.rinv), r0);
    g.m = knot(r, R.knot_u, nr, g.p);
    return g;
}

// 16-byte position records for the density pass: the coordinates on the
// record grid (32-bit integers modulo 2^32) and the element in the fourth
// word. Record differences are exact integers.
__global__ __launch_bounds__(kBlock) void k_make_records(int n, const double4* __restrict__ x, Rec16Frame R,
                                                         uint4* __restrict__ rec)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 xi = x[i];
    rec[i] = make_uint4((unsigned)__double2ll_rn(xi.x * R.inv_q), (unsigned)__double2ll_rn(xi.y * R.inv_q),
                        (unsigned)__double2ll_rn(xi.z * R.inv_q), (unsigned)(int)xi.w);
}

// pass 1 per atom: rho_i = sum_j rho_{j->i}(r_ij), an FP64 sum of FP32 spline
// values. Tab supplies the spline rows (global table or an LDS window).
template <int B, bool ONE, class Tab>
__device__ inline double rho_atom(int i, const uint4& ri, int cnt, const uint4* __restrict__ rec,
                                  const uint4* __restrict__ list4, int ldl, int maxb, const EamArgs& P,
                                  const Rec16Frame& R, const Tab& T)
{
    const int ti = ONE ? 0 : (int)ri.w;
    double rho = 0.0;
    for (int k0 = 0; k0 < cnt; k0 += B) {
        uint32_t e[B];
        uint4 rj[B];
        int row[B];
        float p[B];
        bool in[B];
        list_entries<B>(list4, ldl, maxb, k0, i, cnt, e);
#pragma unroll
        for (int b = 0; b < B; ++b) rj[b] = rec[e[b] & kIndexMask];
        const bool any_img = wave_has_image(e);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int tt = ONE ? 0 : ti * P.ne + (int)rj[b].w;
            const Geo16 g = geo16(ri, rj[b], e[b], any_img, R, R.cut_sq_u[tt], P.nr);
            in[b] = g.in && k0 + b < cnt;
            row[b] = in[b] ? tt * P.nr + g.m : 0;
            p[b] = g.p;
        }
        float4 c[B];
        T.fetch(row, in, c);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const float v = ((c[b].w * p[b] + c[b].z) * p[b] + c[b].y) * p[b] + c[b].x;
            rho += (double)(in[b] ? v : 0.f);
        }
    }
    return rho;
}

// density spline rows from the global table [ne*ne][nr]
struct RhoRowsGlobal {
    const float4* tab;
    template <int B>
    __device__ void fetch(const int (&row)[B], const bool (&)[B], float4 (&c)[B]) const
    {
#pragma unroll
        for (int b = 0; b < B; ++b) c[b] = *(const float4*)((const char*)tab + (size_t)((uint32_t)row[b] << 4));
    }
};

// density spline rows m_lo.. held in LDS (one element); rows below the window,
// rare in a condensed phase, come from the global table
struct RhoRowsLds {
    const float4* tab;
    const float4* win;
    int m_lo;
    template <int B>
    __device__ void fetch(const int (&row)[B], const bool (&in)[B], float4 (&c)[B]) const
    {
        bool low[B], any = false;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int r = row[b] - m_lo;
            low[b] = in[b] && r < 0;
            any |= low[b];
            c[b] = win[r < 0 ? 0 : r];
        }
        if (__ballot(any) != 0ull) {
#pragma unroll
            for (int b = 0; b < B; ++b)
                if (low[b]) c[b] = tab[row[b]];
        }
    }
};

// pass 2 per atom: pair + embedding forces on the owner (no scatter), and with
// EV the sums of the full-list pairs' energy (ev[0], phi) and virial
// (ev[1..6], in grid units: u dx dy) added to ev; without EV (the
// integrator's steps that read neither) the FP64 energy and virial arithmetic
// is left out. One element: neighbor records are the 16-byte pass-2 records
// (coordinates, F'); otherwise the density records (coordinates, element) and
// F' from dF.
template <int B, bool ONE, bool EV, class Tab>
__device__ inline void force_atom(int i, const uint4& ri, int ti, float dFi, int cnt, const uint4* __restrict__ rec,
                                  const uint4* __restrict__ xr, const float* __restrict__ dF,
                                  const uint4* __restrict__ list4, int ldl, int maxb, const EamArgs& P,
                                  const Rec16Frame& R, const Tab& T, float& fx, float& fy, float& fz,
                                  double (&ev)[7])
{
    // the list entries of the next batch are loaded one batch ahead (this
    // pass's batches are short, so one memory round trip less per batch pays)
    uint32_t en[B];
    list_entries<B>(list4, ldl, maxb, 0, i, cnt, en);
    for (int k0 = 0; k0 < cnt; k0 += B) {
        uint32_t e[B];
        uint4 rj[B];
        float dFj[B];
        Geo16 g[B];
        int row[B];
        bool in[B];
#pragma unroll
        for (int b = 0; b < B; ++b) e[b] = en[b];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const uint32_t j = e[b] & kIndexMask;
            if (ONE) {
                rj[b] = xr[j];
            } else {
                rj[b] = rec[j];
                dFj[b] = dF[j];
            }
        }
        list_entries<B>(list4, ldl, maxb, k0 + B, i, cnt, en);
        const bool any_img = wave_has_image(e);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int tt = ONE ? 0 : ti * P.ne + (int)rj[b].w;
            g[b] = geo16(ri, rj[b], e[b], any_img, R, R.cut_sq_u[tt], P.nr);
            in[b] = g[b].in && k0 + b < cnt;
            row[b] = in[b] ? tt * P.nr + g[b].m : 0;
            if (ONE) dFj[b] = __uint_as_float(rj[b].w);
        }
        // spline rows: (rho' coefficients at i, r*phi c0), (r*phi c1..c3),
        // and with several elements (rho' coefficients at j)
        float4 ca[B], cz[B], cb[B];
        T.fetch(row, in, ca, cz, cb);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const float p = g[b].p;
            const float d_in = (ca[b].z * p + ca[b].y) * p + ca[b].x;
            const float d_out = ONE ? d_in : (cb[b].z * p + cb[b].y) * p + cb[b].x;
            // r*phi and its derivative in p by synthetic division
            const float b1 = cz[b].z * p + cz[b].y;
            const float b0 = b1 * p + cz[b].x;
            const float z = b0 * p + ca[b].w;
            const float zp = ((cz[b].z * p + b1) * p + b0) * P.dr_invf;
            const float rinv = g[b].rinv * R.inv_qf;  // 1/r in 1/Angstrom
            const float phi = z * rinv;
            const float dphi = (zp - phi) * rinv;
            // u = f_pair * q, so that the force is u times the grid separation
            float u = -(dphi + dFi * d_in + dFj[b] * d_out) * g[b].rinv;
            u = in[b] ? u : 0.f;
            fx += u * (float)g[b].dx;
            fy += u * (float)g[b].dy;
            fz += u * (float)g[b].dz;
            if (EV) {
                ev[0] += (double)(in[b] ? phi : 0.f);
                const double ud = (double)u;
                const double ux = ud * g[b].dx, uy = ud * g[b].dy, uz = ud * g[b].dz;
                ev[1] += ux * g[b].dx;
                ev[2] += ux * g[b].dy;
                ev[3] += ux * g[b].dz;
                ev[4] += uy * g[b].dy;
                ev[5] += uy * g[b].dz;
                ev[6] += uz * g[b].dz;
            }
        }
    }
}

// force spline rows from the global tables: [nr][2] (one element) or
// [ne*ne][nr][3]
template <bool ONE>
struct ForceRowsGlobal {
    const float4* tab;
    template <int B>
    __device__ void fetch(const int (&row)[B], const bool (&)[B], float4 (&ca)[B], float4 (&cz)[B],
                          float4 (&cb)[B]) const
    {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (ONE) {
                const float4* c = (const float4*)((const char*)tab + (size_t)((uint32_t)row[b] << 5));
                ca[b] = c[0];
                cz[b] = c[1];
            } else {
                const float4* c = (const float4*)((const char*)tab + (size_t)((uint32_t)row[b] * 48u));
                ca[b] = c[0];
                cz[b] = c[1];
                cb[b] = c[2];
            }
        }
    }
};

// one-element force rows m_lo.. held in LDS as 28 bytes: (rho' d1..d3, r*phi
// c0) [nwin] float4, (c1, c2) [nwin] float2, c3 [nwin] float; rows below the
// window come from the global table
struct ForceRowsLds {
    const float4* tab;
    const float4* w4;
    const float2* w2;
    const float* w1;
    int m_lo;
    template <int B>
    __device__ void fetch(const int (&row)[B], const bool (&in)[B], float4 (&ca)[B], float4 (&cz)[B],
                          float4 (&)[B]) const
    {
        bool low[B], any = false;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            int r = row[b] - m_lo;
            low[b] = in[b] && r < 0;
            any |= low[b];
            r = r < 0 ? 0 : r;
            ca[b] = w4[r];
            const float2 t = w2[r];
            cz[b] = make_float4(t.x, t.y, w1[r], 0.f);
        }
        if (__ballot(any) != 0ull) {
#pragma unroll
            for (int b = 0; b < B; ++b)
                if (low[b]) {
                    ca[b] = tab[(size_t)row[b] * 2];
                    cz[b] = tab[(size_t)row[b] * 2 + 1];
                }
        }
    }
};

// second velocity-Verlet half kick of the owner of f_i, m v (x) v, and the sum
// of f / m that the next step's centre-of-mass shift needs
__device__ inline void kick_atom(int i, int ti, float fx, float fy, float fz, const KickArgs& K,
                                 double4* __restrict__ v, double (&mvv)[6], double (&fm)[3])
{
    double4 vi = v[i];
    const double m = K.mass[ti];
    vi.x += (double)fx * K.dt2 / m;
    vi.y += (double)fy * K.dt2 / m;
    vi.z += (double)fz * K.dt2 / m;
    v[i] = vi;
    mvv[0] += vi.x * vi.x * m;
    mvv[1] += vi.x * vi.y * m;
    mvv[2] += vi.x * vi.z * m;
    mvv[3] += vi.y * vi.y * m;
    mvv[4] += vi.y * vi.z * m;
    mvv[5] += vi.z * vi.z * m;
    fm[0] += (double)fx / m;
    fm[1] += (double)fy / m;
    fm[2] += (double)fz / m;
}

// pass 1 (one thread per atom, 256-thread workgroups): rho_i, F(rho_i) and
// F'(rho_i); embedding energy reduced per workgroup. With one element,
// F'(rho_i) goes into the fourth word of the pass-2 record (xr); otherwise
// into dF.
template <int B, bool ONE>
__global__ __launch_bounds__(kBlock) void k_eam_rho(int n, const uint4* __restrict__ rec, const int* __restrict__ nn,
                                                    const uint4* __restrict__ list4, int ldl, int maxb, EamArgs P,
                                                    Rec16Frame R, double* __restrict__ rho_out,
                                                    uint4* __restrict__ xr, float* __restrict__ dF_out,
                                                    double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        const uint4 ri = rec[i];
        const int ti = ONE ? 0 : (int)ri.w;
        const double rho = rho_atom<B, ONE>(i, ri, nn[i], rec, list4, ldl, maxb, P, R, RhoRowsGlobal{P.tabA});
        double F, dF;
        embed(P, ti, rho, F, dF);
        rho_out[i] = rho;
        if (ONE)
            xr[i] = make_uint4(ri.x, ri.y, ri.z, __float_as_uint((float)dF));
        else
            dF_out[i] = (float)dF;
        acc[0] = F;
    }
    block_sum_store<1>(acc, part + (size_t)blk * kPartW);
}

// pass 2 (one thread per atom, 256-thread workgroups) with the fused second
// half kick when K.on; energy and virial partials only with EV
template <int B, bool ONE, bool EV>
__global__ __launch_bounds__(kBlock) void k_eam_force(int n, const uint4* __restrict__ rec,
                                                      const uint4* __restrict__ xr, const float* __restrict__ dF,
                                                      Rec16Frame R, const int* __restrict__ nn,
                                                      const uint4* __restrict__ list4, int ldl, int maxb, EamArgs P,
                                                      float4* __restrict__ f_out, double* __restrict__ part,
                                                      KickArgs K, double4* __restrict__ v,
                                                      double* __restrict__ part_mvv, double* __restrict__ part_cv)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    double mvv[6] = {0, 0, 0, 0, 0, 0};
    double fm[3] = {0, 0, 0};
    if (i < n) {
        const uint4 ri = ONE ? xr[i] : rec[i];
        const int ti = ONE ? 0 : (int)ri.w;
        const float dFi = ONE ? __uint_as_float(ri.w) : dF[i];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        force_atom<B, ONE, EV>(i, ri, ti, dFi, nn[i], rec, xr, dF, list4, ldl, maxb, P, R,
                               ForceRowsGlobal<ONE>{ONE ? P.tabC : P.tabB}, fx, fy, fz, acc);
        f_out[i] = make_float4(fx, fy, fz, 0.f);
        if (K.on) kick_atom(i, ti, fx, fy, fz, K, v, mvv, fm);
    }
    if (EV) {
        // half of each full-list pair: energy phi / 2, virial (q/2) u dx dy
        // (powers of two: scaling the sums is exact)
        acc[0] *= 0.5;
        for (int k = 1; k < 7; ++k) acc[k] *= 0.5 * R.q;
        block_sum_store<7>(acc, part + (size_t)blk * kPartW);
    }
    if (K.on) {
        if (EV) __syncthreads();
        block_sum_store<6>(mvv, part_mvv + (size_t)blk * kPartW);
        __syncthreads();
        block_sum_store<3>(fm, part_cv + (size_t)blk * kPartW);
    }
}

// Persistent passes for one element with the spline rows in LDS: one
// 1024-thread workgroup per CU loads the row window once and walks chunks of
// kBigBlock atoms. The chunks form 8 contiguous ranges, one per XCD (where the
// workgroup runs); the workgroups of an XCD take consecutive chunks side by
// side and stride through the range together, so the chunks in flight on an
// XCD form one compact region whose neighbor halos are shared in its L2 (a
// contiguous run per workgroup measured 15-25% more HBM fetch).
constexpr int kBigBlock = 1024;

struct ChunkWalk {
    int c, end, step;
};
__device__ inline ChunkWalk chunk_walk(int n)
{
    const int nchunks = (n + kBigBlock - 1) / kBigBlock;
    const int per = (nchunks + kXcd - 1) / kXcd;
    const int x = blockIdx.x % kXcd;
    ChunkWalk w;
    w.c = x * per + (int)blockIdx.x / kXcd;
    w.end = min((x + 1) * per, nchunks);
    w.step = gridDim.x / kXcd;
    return w;
}

template <int B>
__global__ __launch_bounds__(kBigBlock) void k_eam_rho_lds(int n, const uint4* __restrict__ rec,
                                                           const int* __restrict__ nn,
                                                           const uint4* __restrict__ list4, int ldl, int maxb,
                                                           EamArgs P, Rec16Frame R, int m_lo, int nwin,
                                                           double* __restrict__ rho_out, uint4* __restrict__ xr,
                                                           double* __restrict__ part)
{
    extern __shared__ float4 lds_rows[];
    for (int r = threadIdx.x; r < nwin; r += kBigBlock) lds_rows[r] = P.tabA[m_lo + r];
    __syncthreads();
    const RhoRowsLds T{P.tabA, lds_rows, m_lo};
    double acc[1] = {0.0};
    for (ChunkWalk w = chunk_walk(n); w.c < w.end; w.c += w.step) {
        const int i = w.c * kBigBlock + threadIdx.x;
        if (i >= n) continue;
        const uint4 ri = rec[i];
        const double rho = rho_atom<B, true>(i, ri, nn[i], rec, list4, ldl, maxb, P, R, T);
        double F, dF;
        embed(P, 0, rho, F, dF);
        rho_out[i] = rho;
        xr[i] = make_uint4(ri.x, ri.y, ri.z, __float_as_uint((float)dF));
        acc[0] += F;
    }
    block_sum_store<1, kBigBlock>(acc, part + (size_t)blockIdx.x * kPartW);
}

template <int B, bool EV>
__global__ __launch_bounds__(kBigBlock) void k_eam_force_lds(int n, const uint4* __restrict__ xr, Rec16Frame R,
                                                             const int* __restrict__ nn,
                                                             const uint4* __restrict__ list4, int ldl, int maxb,
                                                             EamArgs P, int m_lo, int nwin,
                                                             float4* __restrict__ f_out, double* __restrict__ part,
                                                             KickArgs K, double4* __restrict__ v,
                                                             double* __restrict__ part_mvv,
                                                             double* __restrict__ part_cv)
{
    extern __shared__ float4 lds_rows[];
    float2* w2 = (float2*)(lds_rows + nwin);
    float* w1 = (float*)(w2 + nwin);
    for (int r = threadIdx.x; r < nwin; r += kBigBlock) {
        const float4 a = P.tabC[(size_t)(m_lo + r) * 2], c = P.tabC[(size_t)(m_lo + r) * 2 + 1];
        lds_rows[r] = a;
        w2[r] = make_float2(c.x, c.y);
        w1[r] = c.z;
    }
    __syncthreads();
    const ForceRowsLds T{P.tabC, lds_rows, w2, w1, m_lo};
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    double mvv[6] = {0, 0, 0, 0, 0, 0};
    double fm[3] = {0, 0, 0};
    for (ChunkWalk cw = chunk_walk(n); cw.c < cw.end; cw.c += cw.step) {
        const int i = cw.c * kBigBlock + threadIdx.x;
        if (i >= n) continue;
        const uint4 ri = xr[i];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        force_atom<B, true, EV>(i, ri, 0, __uint_as_float(ri.w), nn[i], nullptr, xr, nullptr, list4, ldl, maxb, P,
                                R, T, fx, fy, fz, acc);
        f_out[i] = make_float4(fx, fy, fz, 0.f);
        if (K.on) kick_atom(i, 0, fx, fy, fz, K, v, mvv, fm);
    }
    if (EV) {
        // half of each full-list pair: energy phi / 2, virial (q/2) u dx dy
        // (powers of two: scaling the sums is exact)
        acc[0] *= 0.5;
        for (int k = 1; k < 7; ++k) acc[k] *= 0.5 * R.q;
        block_sum_store<7, kBigBlock>(acc, part + (size_t)blockIdx.x * kPartW);
    }
    if (K.on) {
        if (EV) __syncthreads();
        block_sum_store<6, kBigBlock>(mvv, part_mvv + (size_t)blockIdx.x * kPartW);
        __syncthreads();
        block_sum_store<3, kBigBlock>(fm, part_cv + (size_t)blockIdx.x * kPartW);
    }
}

// ---------------- NVT ----------------

// nhc state layout (doubles)
// NH_SV: sum of v after the last kick, NH_SF: sum of f / m of the last forces
enum { NH_FAC = 0, NH_T = 1, NH_ETA = 2, NH_MVV = 5, NH_COM = 11, NH_SV = 14, NH_SF = 17, NH_SIZE = 20 };

struct NvtArgs {
    double dt, dt2, kB, T, ddt, ddt2, freq_sq;
    int ndof, natms;
    double mass[kMaxElems];
};

// remove the centre-of-mass drift of the step, then the skin check of
// DynamicMD::decide (|x - x0|^2 > skin^2 / 4 anywhere -> rebuild); with rec
// set, also the 16-byte position record of the shifted position (the force
// passes then skip their own record pass; a rebuild rewrites the records)
__global__ __launch_bounds__(kBlock) void k_shift_decide(int n, double4* __restrict__ x, const double4* __restrict__ x0,
                                                         const double* __restrict__ com, double inv_n, double skin_sq4,
                                                         int* __restrict__ flags, Rec16Frame R, uint4* __restrict__ rec)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    double4 xi = x[i];
    xi.x -= com[0] * inv_n;
    xi.y -= com[1] * inv_n;
    xi.z -= com[2] * inv_n;
    x[i] = xi;
    if (rec)
        rec[i] = make_uint4((unsigned)__double2ll_rn(xi.x * R.inv_q), (unsigned)__double2ll_rn(xi.y * R.inv_q),
                            (unsigned)__double2ll_rn(xi.z * R.inv_q), (unsigned)(int)xi.w);
    const double4 a = x0[i];
    if (rsq_exact(a.x, a.y, a.z, xi.x, xi.y, xi.z) > skin_sq4) atomicOr(&flags[1], 1);
}

// The first half of a velocity-Verlet step in one pass (update_x_d__x,
// src/md_nvt.cpp:70-102, then DynamicMD::decide): v <- v fac + f dt/2 / m,
// x <- x + v dt, then x <- x - c, where c = (sum of v dt) / natms is the
// centre-of-mass displacement of the step, which k_nhc computes before the
// pass from the sums of v and f / m the previous force pass left; then the
// skin check against x0 and the 16-byte position record. Partial sums of the
// new v go to part_sv.
__global__ __launch_bounds__(kBlock) void k_step(int n, double4* __restrict__ x, double4* __restrict__ v,
                                                 const float4* __restrict__ f, NvtArgs A,
                                                 const double* __restrict__ nh, const double4* __restrict__ x0,
                                                 double skin_sq4, int* __restrict__ flags, Rec16Frame R,
                                                 uint4* __restrict__ rec, double* __restrict__ part_sv)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double sv[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double fac = nh[NH_FAC];
        double4 xi = x[i];
        double4 vi = v[i];
        const float4 fi = f[i];
        const double m = A.mass[(int)xi.w];
        vi.x = vi.x * fac + (double)fi.x * A.dt2 / m;
        vi.y = vi.y * fac + (double)fi.y * A.dt2 / m;
        vi.z = vi.z * fac + (double)fi.z * A.dt2 / m;
        xi.x += vi.x * A.dt;
        xi.y += vi.y * A.dt;
        xi.z += vi.z * A.dt;
        xi.x -= nh[NH_COM];
        xi.y -= nh[NH_COM + 1];
        xi.z -= nh[NH_COM + 2];
        x[i] = xi;
        v[i] = vi;
        rec[i] = make_uint4((unsigned)__double2ll_rn(xi.x * R.inv_q), (unsigned)__double2ll_rn(xi.y * R.inv_q),
                            (unsigned)__double2ll_rn(xi.z * R.inv_q), (unsigned)(int)xi.w);
        const double4 a = x0[i];
        if (rsq_exact(a.x, a.y, a.z, xi.x, xi.y, xi.z) > skin_sq4) atomicOr(&flags[1], 1);
        sv[0] = vi.x;
        sv[1] = vi.y;
        sv[2] = vi.z;
    }
    block_sum_store<3>(sv, part_sv + (size_t)blk * kPartW);
}

// second half kick (when kick), the kinetic tensor sum m v (x) v, and the sums
// of v and f / m for the next step's centre-of-mass shift
__global__ __launch_bounds__(kBlock) void k_kick2(int n, const double4* __restrict__ x, double4* __restrict__ v,
                                                  const float4* __restrict__ f, NvtArgs A, int kick,
                                                  double* __restrict__ part, double* __restrict__ part_cv)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double cv[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double4 vi = v[i];
        const double m = A.mass[(int)x[i].w];
        const float4 fi = f[i];
        if (kick) {
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
        cv[0] = vi.x;
        cv[1] = vi.y;
        cv[2] = vi.z;
        cv[3] = (double)fi.x / m;
        cv[4] = (double)fi.y / m;
        cv[5] = (double)fi.z / m;
    }
    block_sum_store<6>(acc, part + (size_t)blk * kPartW);
    __syncthreads();
    block_sum_store<6>(cv, part_cv + (size_t)blk * kPartW);
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

// mode 0: initial T from mvv, fac = 1, the sums of v and f / m from cv[0..5];
// mode 1: thermostat at the start of a step (fac *= s), then the step's
// centre-of-mass displacement c = dt (fac sum v + dt/2 sum f / m) / natms;
// mode 2: T from the new mvv, the sums of v after the kick (cv[0..2], the
// step's v, plus dt/2 cv[3..5], the new f / m), thermostat after the step
// (fac = s)
__global__ void k_nhc(double* __restrict__ nh, const double* __restrict__ mvv, const double* __restrict__ cv,
                      NvtArgs A, int mode)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (mode == 0)
        for (int k = 0; k < 3; ++k) {
            nh[NH_SV + k] = cv[k];
            nh[NH_SF + k] = cv[3 + k];
        }
    if (mode == 2)
        for (int k = 0; k < 3; ++k) {
            nh[NH_SV + k] = cv[k] + A.dt2 * cv[3 + k];
            nh[NH_SF + k] = cv[3 + k];
        }
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
    if (mode == 1)
        for (int k = 0; k < 3; ++k)
            nh[NH_COM + k] = (nh[NH_FAC] * nh[NH_SV + k] + A.dt2 * nh[NH_SF + k]) * A.dt / A.natms;
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

// partial sums of the velocities (the centre-of-mass sums after an exchange
// has changed the atoms)
__global__ __launch_bounds__(kBlock) void k_sum_v(int n, const double4* __restrict__ v, double* __restrict__ part)
{
    const int blk = xcd_block();
    const int i = blk * kBlock + threadIdx.x;
    double sv[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double4 vi = v[i];
        sv[0] = vi.x;
        sv[1] = vi.y;
        sv[2] = vi.z;
    }
    block_sum_store<3>(sv, part + (size_t)blk * kPartW);
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
    // derivative coefficients of a density spline, d/dr in 1/Angstrom: the
    // factors 1, 2, 3 and 1/dr applied in FP64 before the rounding to FP32
    const double dri = 1.0 / t_.dr;
    auto dcoef = [dri](const double* c, float w) {
        return make_float4((float)(c[1] * dri), (float)(2.0 * c[2] * dri), (float)(3.0 * c[3] * dri), w);
    };
    for (int a = 0; a < ne; ++a)
        for (int b = 0; b < ne; ++b) {
            const size_t tt = (size_t)a * ne + b;
            for (int m = 0; m < nr; ++m) {
                const double* in = t_.rho_at(b, a) + (size_t)m * 4;   // density at a due to b
                const double* out = t_.rho_at(a, b) + (size_t)m * 4;  // density at b due to a
                const double* rp = t_.rphi_at(a, b) + (size_t)m * 4;
                A[tt * nr + m] = f4(in);
                B[(tt * nr + m) * 3 + 0] = dcoef(in, (float)rp[0]);
                B[(tt * nr + m) * 3 + 1] = make_float4((float)rp[1], (float)rp[2], (float)rp[3], 0.f);
                B[(tt * nr + m) * 3 + 2] = dcoef(out, 0.f);
            }
        }
    // one element: the density pair is symmetric, so a row needs only the
    // rho derivative coefficients and the four r*phi coefficients (32 B)
    std::vector<float4> Cc(ne == 1 ? (size_t)nr * 2 : 1);
    if (ne == 1)
        for (int m = 0; m < nr; ++m) {
            Cc[(size_t)m * 2] = B[(size_t)m * 3];
            Cc[(size_t)m * 2 + 1] = B[(size_t)m * 3 + 1];
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
    elem_host_.assign(elem, elem + natms);
    MAPP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    box_ = make_box(H);
    max_cut_ = ff_->tables().max_rc() + skin_;
    setup_grid();
    x_.reserve(n_); x_alt_.reserve(n_); s_.reserve(n_); x0_.reserve(n_);
    v_.reserve(n_); v_alt_.reserve(n_); id_.reserve(n_); id_alt_.reserve(n_); slot_of_.reserve(n_);
    f_.reserve(n_); rho_.reserve(n_); dF_.reserve(n_); xr_.reserve(n_); rec_.reserve(n_);
    keys_.reserve(n_); keys_alt_.reserve(n_); vals_.reserve(n_); vals_alt_.reserve(n_); s_tmp_.reserve(n_);
    nn_.reserve(n_);
    flags_.reserve(4);
    part_.reserve((size_t)5 * grid_for(n_) * kPartW + 64);  // rho, force, integrator, two centre-of-mass rows
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
    plan_lds();
    rebuild();
}

// A new box (and, when x is given, new positions in caller order): the box
// algebra, the cell grid, the record frame and image shifts, then a rebuild.
// A box the grid cannot hold leaves the system as it was.
void MDSystem::set_box(const double* H, const double* x)
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
    if (x) set_x(x);
    rebuild();
}

// A new set of atoms in the same box (the end of a grand-canonical exchange,
// DynamicMD::fin_xchng): positions, elements and velocities in the new caller
// order, then the cell sort and the list as at construction.
void MDSystem::reset_atoms(int natms, const double* x, const uint8_t* elem, const double* v)
{
    MAPP_REQUIRE(natms >= 1, ERR_ARG, "at least one atom is required");
    MAPP_REQUIRE((uint32_t)natms <= kIndexMask, ERR_CAPACITY, "at most 33554431 atoms per system");
    const int ne = ff_->tables().ne;
    for (int i = 0; i < natms; ++i) MAPP_REQUIRE(elem[i] < ne, ERR_ARG, "element index out of range");
    sync();
    n_ = natms;
    elem_host_.assign(elem, elem + natms);
    x_.reserve(n_); x_alt_.reserve(n_); s_.reserve(n_); x0_.reserve(n_);
    v_.reserve(n_); v_alt_.reserve(n_); id_.reserve(n_); id_alt_.reserve(n_); slot_of_.reserve(n_);
    f_.reserve(n_); rho_.reserve(n_); dF_.reserve(n_); xr_.reserve(n_); rec_.reserve(n_);
    keys_.reserve(n_); keys_alt_.reserve(n_); vals_.reserve(n_); vals_alt_.reserve(n_); s_tmp_.reserve(n_);
    nn_.reserve(n_);
    part_.reserve((size_t)5 * grid_for(n_) * kPartW + 64);
    std::vector<double4> xh(n_), vh(n_);
    std::vector<int> idh(n_);
    for (int i = 0; i < n_; ++i) {
        xh[i] = make_double4(x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], (double)elem[i]);
        vh[i] = v ? make_double4(v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2], 0.0)
                  : make_double4(0.0, 0.0, 0.0, 0.0);
        idh[i] = i;
    }
    MAPP_HIP(hipMemcpyAsync(x_.get(), xh.data(), sizeof(double4) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(v_.get(), vh.data(), sizeof(double4) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemcpyAsync(id_.get(), idh.data(), sizeof(int) * n_, hipMemcpyHostToDevice, stream_));
    MAPP_HIP(hipMemsetAsync(f_.get(), 0, sizeof(float4) * n_, stream_));
    setup_grid();
    rebuild();
    sync();
}

void MDSystem::get_elem(uint8_t* elem) const
{
    std::memcpy(elem, elem_host_.data(), elem_host_.size());
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

// LDS row windows of the persistent one-element passes: the rows of the
// largest r that fit the device's workgroup LDS (16 B per density row, 28 B
// per force row), leaving room for the static reduction arrays
void MDSystem::plan_lds()
{
    lds_ok_ = use_lds_ = false;
    const EamArgs P = ff_->args();
    if (P.ne != 1 || P.nr < 2) return;
    int dev = 0, lds_max = 0, ncu = 0;
    MAPP_HIP(hipGetDevice(&dev));
    MAPP_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    MAPP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long budget = (long)lds_max - 4096;
    if (budget < 16384 || ncu < 1) return;
    const int rows = P.nr - 1;  // spline intervals 0 .. nr-2
    rho_nwin_ = (int)std::min<long>(rows, budget / 16);
    rho_m_lo_ = rows - rho_nwin_;
    f_nwin_ = (int)std::min<long>(rows, budget / 28);
    f_m_lo_ = rows - f_nwin_;
    rho_lds_bytes_ = (size_t)rho_nwin_ * 16;
    f_lds_bytes_ = (size_t)f_nwin_ * 28;
    lds_grid_ = std::max(1, ncu / kXcd) * kXcd;
    const void* fns[] = {(const void*)k_eam_rho_lds<1>,         (const void*)k_eam_rho_lds<2>,
                         (const void*)k_eam_rho_lds<4>,         (const void*)k_eam_rho_lds<8>,
                         (const void*)k_eam_force_lds<1, true>, (const void*)k_eam_force_lds<2, true>,
                         (const void*)k_eam_force_lds<1, false>, (const void*)k_eam_force_lds<2, false>,
                         (const void*)k_eam_force_lds<4, false>};
    for (int k = 0; k < 9; ++k) {
        const size_t bytes = k < 4 ? rho_lds_bytes_ : f_lds_bytes_;
        if (hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
    }
    lds_ok_ = use_lds_ = true;
}

bool MDSystem::set_lds(bool on)
{
    use_lds_ = on && lds_ok_;
    return use_lds_;
}

void MDSystem::set_batch(int rho, int force)
{
    MAPP_REQUIRE(rho == 1 || rho == 2 || rho == 4 || rho == 8, ERR_ARG, "rho batch must be 1, 2, 4 or 8");
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
    // grid shift of every image code (n0+2) + 5 (n1+2) + 25 (n2+2)
    refresh_frame();
    // q is a power of two: the scaled cut-offs and knot spacing are exact
    {
        const EamArgs P = ff_->args();
        for (int k = 0; k < P.ne * P.ne; ++k) r16_.cut_sq_u[k] = P.cut_sq[k] * r16_.inv_q * r16_.inv_q;
        r16_.knot_u = r16_.q * P.dr_inv;
        r16_.inv_qf = (float)r16_.inv_q;
    }
    // the cell ranks and stencil order depend on the cell counts alone: a box
    // change that keeps them (the NST rebuilds) reuses them
    const int topo[6] = {grid_.nc[0], grid_.nc[1], grid_.nc[2], grid_.R[0], grid_.R[1], grid_.R[2]};
    if (!std::equal(topo, topo + 6, topo_)) {
        std::copy(topo, topo + 6, topo_);
        cell_topology();
    }
    // neighbors expected in the search sphere at this density, with margin;
    // the capacity is a whole number of 4-entry blocks
    const double dens = n_ / box_.vol;
    const double sphere = 4.0 / 3.0 * M_PI * max_cut_ * max_cut_ * max_cut_;
    maxn_ = std::max(maxn_, (int)(1.25 * dens * sphere) + 16);
    maxn_ = (maxn_ + 3) & ~3;
}

void MDSystem::cell_topology()
{
    const int ncells = grid_.ncells;
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
    // with a 3x3x3 stencil of distinct cells, each cell's stencil in Morton-rank order
    ordered_ = grid_.R[0] == 1 && grid_.R[1] == 1 && grid_.R[2] == 1 && grid_.nc[0] >= 3 && grid_.nc[1] >= 3 &&
               grid_.nc[2] >= 3 && (long long)ncells < (1ll << 27);
    if (ordered_) {
        GridArgs g{};
        for (int d = 0; d < 3; ++d) g.nc[d] = grid_.nc[d];
        cell_order_.reserve((size_t)2 * ncells);
        hipLaunchKernelGGL(k_cell_order, dim3((ncells + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, ncells, g,
                           cell_rank_.get(), cell_order_.get());
        MAPP_HIP(hipGetLastError());
    }
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
                       id_.get(), bd, r16_, x_alt_.get(), v_alt_.get(), id_alt_.get(), s_.get(), x0_.get(),
                       slot_of_.get(), rec_.get());
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
            // Record separations differ from the exact ones by at most 4 q
            // per coordinate (q/2 rounding of each position, q/2 per H row
            // for up to two box vectors in each of three rows), so at most
            // 4 sqrt(3) q < 7 q in length; the band takes 16 q either side.
            const double cu = c * r16_.inv_q;
            g.rec_in_sq[a * t.ne + b] = (cu - 16.0) * (cu - 16.0);
            g.rec_out_sq[a * t.ne + b] = (cu + 16.0) * (cu + 16.0);
        }
    BoxDev bd;
    std::memcpy(bd.h, box_.h, sizeof bd.h);
    std::memcpy(bd.b, box_.b, sizeof bd.b);
    ldl_ = n_;
    for (int attempt = 0; attempt < 2; ++attempt) {
        list_.reserve((size_t)maxn_ * ldl_);
        MAPP_HIP(hipMemsetAsync(flags_.get(), 0, sizeof(int), stream_));
        prof_.begin(K_BUILD_LIST, stream_);
        if (t.ne == 1)
            hipLaunchKernelGGL(k_build_list<true>, dim3(nb), dim3(kBlock), 0, stream_, n_, s_.get(), rec_.get(), r16_,
                               bd, g, cell_rank_.get(), cell_start_.get(), ordered_ ? cell_order_.get() : nullptr,
                               maxn_ / 4, ldl_, list4(), nn_.get(), flags_.get());
        else
            hipLaunchKernelGGL(k_build_list<false>, dim3(nb), dim3(kBlock), 0, stream_, n_, s_.get(), rec_.get(),
                               r16_, bd, g, cell_rank_.get(), cell_start_.get(),
                               ordered_ ? cell_order_.get() : nullptr, maxn_ / 4, ldl_, list4(), nn_.get(),
                               flags_.get());
        prof_.end(stream_);
        int over = 0;
        MAPP_HIP(hipMemcpyAsync(&over, flags_.get(), sizeof(int), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        if (over <= maxn_) return;
        maxn_ = (over + over / 8 + 8 + 3) & ~3;
    }
    throw Error(ERR_STATE, "neighbor list capacity did not converge");
}

void MDSystem::force_passes(const KickArgs* kick, bool records_fresh, float4* f_out, double* rho_out, bool ev)
{
    const int nb = grid_for(n_);
    float4* fo = f_out ? f_out : f_.get();
    double* ro = rho_out ? rho_out : rho_.get();
    const EamArgs P = ff_->args();
    prof_.begin(K_EAM_RHO, stream_);
    if (!records_fresh)
        hipLaunchKernelGGL(k_make_records, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), r16_, rec_.get());
    const bool one = P.ne == 1;
    const bool lds = one && use_lds_;
    KickArgs K{};
    if (kick) {
        K = *kick;
        K.on = 1;
    }
    double* p3 = part_.get() + (size_t)2 * nb * kPartW;
    double* p4 = part_.get() + (size_t)3 * nb * kPartW;
    const int maxb = maxn_ / 4;
    if (lds) {
        // at most one workgroup per CU, and never more rows than the partial-sum buffer has per pass
        const int G = std::min(lds_grid_, nb);
#define MAPP_RHO(B) hipLaunchKernelGGL((k_eam_rho_lds<B>), dim3(G), dim3(kBigBlock), rho_lds_bytes_, stream_, n_, \
                                       rec_.get(), nn_.get(), list4(), ldl_, maxb, P, r16_, rho_m_lo_, rho_nwin_, \
                                       ro, xr_.get(), part_.get())
        switch (batch_rho_) {
            case 1: MAPP_RHO(1); break;
            case 2: MAPP_RHO(2); break;
            case 4: MAPP_RHO(4); break;
            default: MAPP_RHO(8); break;
        }
#undef MAPP_RHO
        prof_.end(stream_);
        prof_.begin(K_EAM_FORCE, stream_);
        double* p2 = part_.get() + (size_t)G * kPartW;
#define MAPP_FORCE(B, EV) hipLaunchKernelGGL((k_eam_force_lds<B, EV>), dim3(G), dim3(kBigBlock), f_lds_bytes_, \
                                             stream_, n_, xr_.get(), r16_, nn_.get(), list4(), ldl_, maxb, P, f_m_lo_, \
                                             f_nwin_, fo, p2, K, v_.get(), p3, p4)
        // 1024-thread workgroups leave 128 VGPRs per lane: batches of 4 fit
        // without the FP64 energy and virial sums only
        if (batch_force_ == 1) {
            if (ev) MAPP_FORCE(1, true); else MAPP_FORCE(1, false);
        } else if (batch_force_ >= 4 && !ev) {
            MAPP_FORCE(4, false);
        } else {
            if (ev) MAPP_FORCE(2, true); else MAPP_FORCE(2, false);
        }
#undef MAPP_FORCE
        prof_.end(stream_);
        mvv_rows_ = G;
        if (ev) reduce_rows(part_.get(), 2 * G, 7, red_.get());
        return;
    }
#define MAPP_RHO(B, ONE) hipLaunchKernelGGL((k_eam_rho<B, ONE>), dim3(nb), dim3(kBlock), 0, stream_, n_, rec_.get(), \
                                            nn_.get(), list4(), ldl_, maxb, P, r16_, ro, xr_.get(), \
                                            dF_.get(), part_.get())
    switch (batch_rho_) {
        case 1: if (one) MAPP_RHO(1, true); else MAPP_RHO(1, false); break;
        case 2: if (one) MAPP_RHO(2, true); else MAPP_RHO(2, false); break;
        case 4: if (one) MAPP_RHO(4, true); else MAPP_RHO(4, false); break;
        default: if (one) MAPP_RHO(8, true); else MAPP_RHO(8, false); break;
    }
#undef MAPP_RHO
    prof_.end(stream_);
    prof_.begin(K_EAM_FORCE, stream_);
    double* p2 = part_.get() + (size_t)nb * kPartW;
#define MAPP_FORCE(B, ONE, EV) hipLaunchKernelGGL((k_eam_force<B, ONE, EV>), dim3(nb), dim3(kBlock), 0, stream_, n_, \
                                                  rec_.get(), xr_.get(), dF_.get(), r16_, nn_.get(), list4(), ldl_, \
                                                  maxb, P, fo, p2, K, v_.get(), p3, p4)
#define MAPP_FORCE_EV(B, ONE) \
    do {                      \
        if (ev)               \
            MAPP_FORCE(B, ONE, true); \
        else                  \
            MAPP_FORCE(B, ONE, false); \
    } while (0)
    switch (batch_force_) {
        case 1: if (one) MAPP_FORCE_EV(1, true); else MAPP_FORCE_EV(1, false); break;
        case 2: if (one) MAPP_FORCE_EV(2, true); else MAPP_FORCE_EV(2, false); break;
        case 8: if (one) MAPP_FORCE_EV(8, true); else MAPP_FORCE_EV(8, false); break;
        default: if (one) MAPP_FORCE_EV(4, true); else MAPP_FORCE_EV(4, false); break;
    }
#undef MAPP_FORCE_EV
#undef MAPP_FORCE
    prof_.end(stream_);
    mvv_rows_ = nb;
    if (ev) reduce_rows(part_.get(), 2 * nb, 7, red_.get());
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
    // the energy of the two force passes (the EAM energy the reference's
    // __energy_calc sums); their forces and densities go to spare buffers so
    // that the last force_calc's stay readable
    val_f_.reserve(n_);
    val_rho_.reserve(n_);
    force_passes(nullptr, false, val_f_.get(), val_rho_.get());
    double e = 0.0;
    MAPP_HIP(hipMemcpyAsync(&e, red_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    return e;
}

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxScratchDevices];

// this GPU's address of the scratch array, looked up once per device with the
// device current
unsigned char* scratch_rows()
{
    int dev = 0;
    MAPP_HIP(hipGetDevice(&dev));
    MAPP_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, ERR_ARG, "device index out of range of the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        MAPP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MAPP_REQUIRE(p != nullptr, ERR_HIP, "the scratch array has no device address");
        g_scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_dev[dev];
}
}  // namespace

double MDSystem::value_noalloc()
{
    MAPP_REQUIRE(n_ <= kNoAllocMaxAtoms, ERR_CAPACITY,
                 "value without allocation holds at most " + std::to_string(kNoAllocMaxAtoms) + " atoms, the system has " +
                     std::to_string(n_));
    unsigned char* rows = scratch_rows();
    // the passes overwrite every row they use: nothing needs clearing
    force_passes(nullptr, false, reinterpret_cast<float4*>(rows),
                 reinterpret_cast<double*>(rows + (size_t)kNoAllocMaxAtoms * 16));
    double e = 0.0;
    MAPP_HIP(hipMemcpyAsync(&e, red_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    prof_.flush();
    return e;
}

bool MDSystem::decide()
{
    const int nb = grid_for(n_);
    MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
    // no centre-of-mass shift outside the integrator: com sums are zero
    MAPP_HIP(hipMemsetAsync(red_.get() + 16, 0, sizeof(double) * 3, stream_));
    hipLaunchKernelGGL(k_shift_decide, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), x0_.get(), red_.get() + 16,
                       1.0 / n_, 0.25 * skin_ * skin_, flags_.get(), r16_, (uint4*)nullptr);
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

void MDSystem::export_cfg(const char* path, int nelems, const double* masses, const std::vector<std::string>& names,
                          bool with_vel)
{
    MAPP_REQUIRE(nelems >= ff_->tables().ne, ERR_ARG, "one name and mass per element of the force field");
    std::vector<double> x((size_t)3 * n_), v;
    get_x(x.data());
    if (with_vel) {
        v.resize((size_t)3 * n_);
        get_v(v.data());
    }
    double H[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H[i * 3 + j] = box_.H[i][j];
    write_cfg_md(path, n_, H, x.data(), elem_host_.data(), with_vel ? v.data() : nullptr, 0, nullptr, {}, nelems,
                 masses, names);
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
            const uint32_t e = h.list[((size_t)(k >> 2) * ldl_ + s) * 4 + (k & 3)];
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
            const uint32_t e = h.list[((size_t)(k >> 2) * ldl_ + s) * 4 + (k & 3)];
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

int MDSystem::nvt_run(const double* masses, double kB, double T, double dt, int nsteps, MDReport* rep, double* T_part,
                      const ExchangeHook* hook, double* hist_T)
{
    MAPP_REQUIRE(nsteps >= 0 && dt > 0.0 && T > 0.0 && kB > 0.0, ERR_ARG, "invalid NVT parameters");
    int nb = grid_for(n_);
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
    double* cv = red_.get() + 24;  // sums of v and f / m for the centre-of-mass shift
    double* mvv = red_.get() + 8;
    double* rows_mvv = part_.get() + (size_t)2 * nb * kPartW;
    double* rows_cv = part_.get() + (size_t)3 * nb * kPartW;
    double* rows_sv = part_.get() + (size_t)4 * nb * kPartW;
    const int nevery = hook && hook->run ? hook->nevery : 0;
    MAPP_HIP(hipMemsetAsync(nh, 0, sizeof(double) * NH_SIZE, stream_));
    const int rebuilds0 = nrebuild_;
    force_passes(nullptr);
    hipLaunchKernelGGL(k_kick2, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), A, 0, rows_mvv,
                       rows_cv);
    reduce_rows(rows_mvv, nb, 6, mvv);
    reduce_rows(rows_cv, nb, 6, cv);
    hipLaunchKernelGGL(k_nhc, dim3(1), dim3(64), 0, stream_, nh, mvv, cv, A, 0);
    const double skin_sq4 = 0.25 * skin_ * skin_;
    KickArgs kick{};
    kick.dt2 = A.dt2;
    for (int a = 0; a < t.ne; ++a) kick.mass[a] = A.mass[a];
    int flag = 0;
    for (int step = 0; step < nsteps; ++step) {
        prof_.begin(K_INTEGRATE, stream_);
        hipLaunchKernelGGL(k_nhc, dim3(1), dim3(64), 0, stream_, nh, mvv, cv, A, 1);
        MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
        hipLaunchKernelGGL(k_step, dim3(nb), dim3(kBlock), 0, stream_, n_, x_.get(), v_.get(), f_.get(), A, nh,
                           x0_.get(), skin_sq4, flags_.get(), r16_, rec_.get(), rows_sv);
        reduce_rows(rows_sv, nb, 3, cv);
        prof_.end(stream_);
        MAPP_HIP(hipMemcpyAsync(&flag, flags_.get() + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        if (nevery > 0 && (step + 1) % nevery == 0) {
            // the exchange takes the place of the skin check (MDMuVT::run,
            // src/md_muvt.cpp:172-195); it ends with a rebuild, and the atom
            // count, ndof_part = 3 natms and the sums of v follow it
            hook->run();
            nb = grid_for(n_);
            rows_mvv = part_.get() + (size_t)2 * nb * kPartW;
            rows_cv = part_.get() + (size_t)3 * nb * kPartW;
            rows_sv = part_.get() + (size_t)4 * nb * kPartW;
            A.ndof = 3 * n_;
            A.natms = n_;
            hipLaunchKernelGGL(k_sum_v, dim3(nb), dim3(kBlock), 0, stream_, n_, v_.get(), rows_sv);
            reduce_rows(rows_sv, nb, 3, cv);
        } else if (flag) {
            rebuild();  // the permute rewrites the records
        }
        // energy and virial only where they are read: after the last step
        force_passes(&kick, true, nullptr, nullptr, step == nsteps - 1);
        prof_.begin(K_INTEGRATE, stream_);
        reduce_rows(rows_mvv, mvv_rows_, 6, mvv);
        reduce_rows(rows_cv, mvv_rows_, 3, cv + 3);
        hipLaunchKernelGGL(k_nhc, dim3(1), dim3(64), 0, stream_, nh, mvv, cv, A, 2);
        prof_.end(stream_);
        if (hist_T) MAPP_HIP(hipMemcpyAsync(hist_T + step, nh + NH_T, sizeof(double), hipMemcpyDeviceToHost, stream_));
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
