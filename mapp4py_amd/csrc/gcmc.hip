// gcmc.hip -- batched grand-canonical exchange (SGCMC::attmpt with the EAM
// exchange energies) over thousands of independent periodic cells.
//
// Reference behaviour restated here:
//   SGCMC::attmpt draw order and acceptance        src/sgcmc.cpp:255-327
//   prep_s_x_buff (periodic images of the trial)   src/sgcmc.cpp:448-497
//   next_jatm_reg / next_jatm_self                 src/sgcmc.cpp:560-640
//   ins_succ / del_succ (append / move last)       src/sgcmc.cpp:331-444
//   ForceFieldEAM::pre_/xchng_/post_xchng_energy   src/ff_eam.cpp:323-483
//   ForceFieldEAM::init_xchng                      src/ff_eam.cpp:509-578
//   GCMC::init / box_setup                         src/gcmc.cpp:86-150
//   Random (Park-Miller, polar Gaussian)           src/random.cpp:7-68
//
// Design: a cell's chain is sequential, so the parallelism is across cells
// and, inside an attempt, across the cell's atoms. One wave (or all four
// waves of a workgroup, for large cells) owns a cell for the whole launch:
// every lane runs the same random stream in registers, so the decision is
// uniform without any broadcast; the lanes split the atoms, each lane keeps
// the density change of its own atoms and a bitmask of those the trial
// touched, and the few cross-lane values (energy sums, trial density) are
// reduced with shuffles. The reference walks a 27-cell stencil around each
// trial image; every pair within the cutoff lies inside that stencil, so
// testing all atoms of the cell gives the same pair set while keeping atom
// order (which decides the deleted atom) exactly the reference's. All
// arithmetic is FP64 with contraction off so the exchange energies are the
// reference's to rounding of the sums, and accept/reject decisions match.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "box.hpp"
#include "gcmc.hpp"
#include "kern.hpp"

#pragma clang fp contract(off)

namespace mapp {

int gcmc_class_of(int n)
{
    if (n < 256) return 0;
    if (n < 512) return 1;
    if (n < 1024) return 2;
    if (n < 2048) return 3;
    return 4;
}

namespace {

enum { MODE_INS = 0, MODE_DEL = 1 };

__device__ inline double hval(const double4 c, double p) { return ((c.w * p + c.z) * p + c.y) * p + c.x; }

__device__ inline void spline_at(double v, double inv, int n, int& m, double& p)
{
    p = v * inv;
    long long mm = (long long)p;  // truncation toward zero, as the reference's size_t cast for p > -1
    if (mm > n - 2) mm = n - 2;
    if (mm < 0) mm = 0;
    p -= (double)mm;
    if (p > 1.0) p = 1.0;
    m = (int)mm;
}

// F(rho), with the linear extension past rho_max when asked (src/ff_eam.cpp:395-403)
__device__ inline double f_embed(const GcmcArgs& A, int e, double rho, bool extrapolate)
{
    int m;
    double p;
    spline_at(rho, A.drho_inv, A.nrho, m, p);
    const double4 c = A.F[(size_t)e * A.nrho + m];
    double out = hval(c, p);
    if (extrapolate && rho > A.rho_max) out += ((3.0 * c.w * p + 2.0 * c.z) * p + c.y) * A.drho_inv * (rho - A.rho_max);
    return out;
}

// dx0^2 + (dx1^2 + dx2^2) of a - b (src/xmath.h:2464-2477)
__device__ inline double rsq_of(double a0, double a1, double a2, double b0, double b1, double b2)
{
    const double d0 = a0 - b0, d1 = a1 - b1, d2 = a2 - b2;
    return d0 * d0 + (d1 * d1 + d2 * d2);
}

struct Stream {
    int seed, chk;
    double keep;
    __device__ double uniform()
    {
        const int q = seed / 127773, r = seed % 127773;
        const int t = 16807 * r - 2836 * q;
        seed = t > 0 ? t : t + 2147483647;
        return seed * 1.0 / 2147483647;
    }
    __device__ double gaussian()
    {
        if (chk) {
            chk = 0;
            return keep;
        }
        chk = 1;
        double u = 0.0, v = 0.0, s = 1.0;
        while (s >= 1.0) {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            s = u * u + v * v;
        }
        const double f = sqrt(-2.0 * log(s) / s);
        keep = v * f;
        return u * f;
    }
};

// Images of s inside [lo, hi) per dimension in the reference's order: s, s+1,
// ... then s-1, s-2, ... (each by repeated +-1.0). s must lie in [0, 1).
struct Images {
    int up[3], cnt[3];
    __device__ int total() const { return cnt[0] * cnt[1] * cnt[2]; }
};

__device__ inline Images count_images(const double* s, const double* lo, const double* hi)
{
    Images im;
    for (int d = 0; d < 3; ++d) {
        int u = 0, dn = 0;
        for (double t = s[d]; t < hi[d]; t = t + 1.0) ++u;
        for (double t = s[d] - 1.0; lo[d] <= t; t = t - 1.0) ++dn;
        im.up[d] = u;
        im.cnt[d] = u + dn;
    }
    return im;
}

__device__ inline double image_coord(double s, int up, int i)
{
    double t;
    if (i < up) {
        t = s;
        for (int k = 0; k < i; ++k) t = t + 1.0;
    } else {
        t = s - 1.0;
        for (int k = up; k < i; ++k) t = t - 1.0;
    }
    return t;
}

// image k (first dimension fastest) of s, as Cartesian coordinates
__device__ inline void image_x(const double* h, const double* s, const Images& im, int k, double* x)
{
    const int i0 = k % im.cnt[0], i1 = (k / im.cnt[0]) % im.cnt[1], i2 = k / (im.cnt[0] * im.cnt[1]);
    const double b0 = image_coord(s[0], im.up[0], i0);
    const double b1 = image_coord(s[1], im.up[1], i1);
    const double b2 = image_coord(s[2], im.up[2], i2);
    double x0, x1, x2;
    s_to_x(h, b0, b1, b2, x0, x1, x2);
    x[0] = x0;
    x[1] = x1;
    x[2] = x2;
}

template <int W>
__device__ inline void group_sync()
{
    if constexpr (W == 1) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// Values that are the same in every lane, marked so for the compiler (read
// from the first lane): branches on them stay scalar.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return uni(v);  // one lane's value, so every lane holds the same bits
}

struct Arrays {
    double4* xe;
    double4* s;
    double4* v;
    double* rho;
    double* F;
    double* rho_x;
    double* F_x;
    uint8_t* el;
};

// One cell, nattempts attempts, on W waves (w = this wave's index in the
// group, wave-uniform). Control flow is wave-uniform throughout: loop bounds
// and branch conditions are the same in every lane (the cell index comes from
// readfirstlane, so the cell's header and everything derived from it are
// scalar), per-lane work sits in branch-free predicated blocks, and values
// that one lane would store are stored by all lanes alike. No lane ever
// leaves a loop that its wave stays in, so every cross-lane operation runs
// with the whole wave.
// gas atoms among el[lo, hi): ballots over 64 atoms, the loads of four of
// them in flight at a time (the count is the same in every lane)
__device__ inline int count_gas(const uint8_t* __restrict__ el, int lo, int hi, int gas, int lane)
{
    int cnt = 0;
    for (int base = lo; base < hi; base += 256) {
        uint8_t e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = base + 64 * q + lane;
            e[q] = el[j < hi ? j : hi - 1];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) cnt += __popcll(__ballot((base + 64 * q + lane < hi) & (e[q] == gas)));
    }
    return cnt;
}

// index of the k-th gas atom (from 0) among el[lo, hi), or -1: the same
// ballots, stopping at the chunk that holds it
__device__ inline int nth_gas(const uint8_t* __restrict__ el, int lo, int hi, int gas, int k, int lane)
{
    int seen = 0, idx = -1;
    for (int base = lo; base < hi && idx < 0; base += 256) {
        uint8_t e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = base + 64 * q + lane;
            e[q] = el[j < hi ? j : hi - 1];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const unsigned long long m = __ballot((base + 64 * q + lane < hi) & (e[q] == gas));
            const int cnt = __popcll(m);
            if (idx < 0) {
                if (seen + cnt > k) {
                    unsigned long long mm = m;
                    for (int t = 0; t < k - seen; ++t) mm &= mm - 1;
                    idx = base + 64 * q + __ffsll((long long)mm) - 1;
                } else {
                    seen += cnt;
                }
            }
        }
    }
    return idx;
}

template <int W>
__device__ void run_cell(const GcmcArgs& A, GcmcCell* __restrict__ cells, int c, int nattempts, int8_t* dec,
                         double* dU, const Arrays& R, double* trial, double* red, int w, int lane)
{
    GcmcCell hd = cells[c];
    for (int d = 0; d < 3; ++d) {
        hd.lo[d] = uni(hd.lo[d]);
        hd.hi[d] = uni(hd.hi[d]);
    }
    for (int k = 0; k < 6; ++k) hd.h[k] = uni(hd.h[k]);
    hd.zv = uni(hd.zv);
    const int off = uni(hd.off), cap = uni(hd.cap), gas = A.gas, ne = A.ne, nr = A.nr;
    int n = uni(hd.n), ngas = uni(hd.ngas), status = 0;
    Stream rng{uni(hd.seed), uni(hd.gchk), uni(hd.gval)};
    const double cut_gg = A.cut_sq[gas * ne + gas];
    for (int a = 0; a < nattempts; ++a) {
        int mode, del_idx = -1;
        double sb[3];
        if (rng.uniform() < 0.5) {
            mode = MODE_INS;
            for (int d = 0; d < 3; ++d) sb[d] = rng.uniform();
        } else {
            if (ngas == 0) {
                if (w == 0 && dec) {
                    dec[(size_t)c * nattempts + a] = 0;
                    dU[(size_t)c * nattempts + a] = 0.0;
                }
                continue;
            }
            mode = MODE_DEL;
            const int igas = (int)(ngas * rng.uniform());
            // igas-th gas atom in index order
            if constexpr (W < 16) {
                // every wave finds it alike (cells of at most 4,000 atoms in
                // a batch: three workgroup barriers cost more than the scan)
                del_idx = nth_gas(R.el + off, 0, n, gas, igas, lane);
            } else {
                // one large cell: each wave counts the gas atoms of its slice
                // of 64-atom chunks; the wave whose slice holds the igas-th
                // finds it
                int* ish = reinterpret_cast<int*>(red);
                const int slice = ((n + W - 1) / W + 63) & ~63;
                const int lo = min(n, w * slice), hi = min(n, lo + slice);
                const int cnt = count_gas(R.el + off, lo, hi, gas, lane);
                if (w == 0) ish[W] = -1;
                ish[w] = cnt;  // every lane of the wave stores the same count
                __syncthreads();
                int before = 0;
                for (int q = 0; q < w; ++q) before += ish[q];
                __syncthreads();
                if (before <= igas && igas < before + cnt) ish[W] = nth_gas(R.el + off, lo, hi, gas, igas - before, lane);
                __syncthreads();
                del_idx = __builtin_amdgcn_readfirstlane(ish[W]);
            }
            if (del_idx < 0) {  // gas count and element array disagree: stop this chain
                status = 2;
                break;
            }
            const double4 s4 = R.s[off + del_idx];
            sb[0] = uni(s4.x);
            sb[1] = uni(s4.y);
            sb[2] = uni(s4.z);
        }
        const Images im = count_images(sb, hd.lo, hd.hi);
        const int ntrial = im.total();
        if (ntrial > kGcmcMaxTrials) {  // the host bound rules this out; stop the chain rather than overrun
            status = 3;
            break;
        }
        if (w == 0 && lane < ntrial) image_x(hd.h, sb, im, lane, trial + 3 * lane);
        group_sync<W>();

        // ---- exchange energy (pre_xchng_energy + xchng_energy) ----
        const double c0 = mode == MODE_DEL ? -1.0 : 1.0;
        double en = 0.0, en0 = 0.0, rhoi = 0.0;
        unsigned long long mask = 0ull;
        const int nsweep = (n - w * 64 + 64 * W - 1) / (64 * W);  // atoms per lane, rounded up (uniform)
        for (int it = 0; it < nsweep; ++it) {
            const int j = w * 64 + lane + it * 64 * W;
            if (j < n) {
                const double4 xj = R.xe[off + j];
                const int ej = (int)xj.w;
                const double cut = A.cut_sq[gas * ne + ej];
                const double4* tp = A.phi + (size_t)(gas * ne + ej) * nr;
                const double4* tr_ij = A.rho + (size_t)(gas * ne + ej) * nr;
                const double4* tr_ji = A.rho + (size_t)(ej * ne + gas) * nr;
                bool hit = false;
                double rx = 0.0;
                for (int k = 0; k < ntrial; ++k) {
                    const double rsq = rsq_of(trial[3 * k], trial[3 * k + 1], trial[3 * k + 2], xj.x, xj.y, xj.z);
                    if (rsq < cut && !(k == 0 && j == del_idx)) {
                        const double r = sqrt(rsq);
                        int m;
                        double p;
                        spline_at(r, A.dr_inv, nr, m, p);
                        en += hval(tp[m], p) / r;
                        if (!hit) {
                            rx = R.rho[off + j];
                            hit = true;
                        }
                        rx += c0 * hval(tr_ij[m], p);
                        rhoi += hval(tr_ji[m], p);
                    }
                }
                if (hit) {
                    const double fx = f_embed(A, ej, rx, true);
                    en0 += fx - R.F[off + j];
                    R.rho_x[off + j] = rx;
                    R.F_x[off + j] = fx;
                    mask |= 1ull << it;
                }
            }
        }
        // the trial point's own images, half of each pair (next_jatm_self)
        if (w == 0 && lane >= 1 && lane < ntrial) {
            const double rsq =
                rsq_of(trial[0], trial[1], trial[2], trial[3 * lane], trial[3 * lane + 1], trial[3 * lane + 2]);
            if (rsq < cut_gg) {
                const double r = sqrt(rsq);
                int m;
                double p;
                spline_at(r, A.dr_inv, nr, m, p);
                en += 0.5 * hval(A.phi[(size_t)(gas * ne + gas) * nr + m], p) / r;
                rhoi += hval(A.rho[(size_t)(gas * ne + gas) * nr + m], p);
            }
        }
        en = wave_sum(en);
        en0 = wave_sum(en0);
        rhoi = wave_sum(rhoi);
        if constexpr (W > 1) {
            red[4 * w + 0] = en;  // every lane of the wave stores the same value
            red[4 * w + 1] = en0;
            red[4 * w + 2] = rhoi;
            __syncthreads();
            en = en0 = rhoi = 0.0;
            for (int q = 0; q < W; ++q) {
                en += red[4 * q + 0];
                en0 += red[4 * q + 1];
                rhoi += red[4 * q + 2];
            }
            en = uni(en);
            en0 = uni(en0);
            rhoi = uni(rhoi);
        }
        double du = en + en0 * c0;
        du = du + f_embed(A, gas, rhoi, true);

        // ---- acceptance ----
        bool acc;
        if (mode == MODE_INS)
            acc = rng.uniform() < hd.zv / (((double)ngas + 1.0) * exp(A.beta * du));
        else
            acc = rng.uniform() < (double)ngas * exp(A.beta * du) / hd.zv;
        if (w == 0 && dec) {
            dec[(size_t)c * nattempts + a] = (int8_t)(mode == MODE_INS ? (acc ? 1 : -1) : (acc ? 2 : -2));
            dU[(size_t)c * nattempts + a] = du;
        }
        if (acc) {
            if (mode == MODE_INS && n == cap) {
                status = 1;
                break;
            }
            // post_xchng_energy: every touched atom takes its new rho and F
            for (int it = 0; it < nsweep; ++it) {
                if ((mask >> it) & 1ull) {
                    const int j = w * 64 + lane + it * 64 * W;
                    R.rho[off + j] = R.rho_x[off + j];
                    R.F[off + j] = R.F_x[off + j];
                }
            }
            if (mode == MODE_INS) {
                double vel[3];
                for (int d = 0; d < 3; ++d) vel[d] = rng.gaussian() * A.sigma;
                if (w == 0) {  // all lanes store the same new atom
                    const int k = off + n;
                    R.xe[k] = make_double4(trial[0], trial[1], trial[2], (double)gas);
                    R.s[k] = make_double4(sb[0], sb[1], sb[2], 0.0);
                    R.v[k] = make_double4(vel[0], vel[1], vel[2], 0.0);
                    R.rho[k] = rhoi;
                    R.F[k] = f_embed(A, gas, rhoi, true);
                    R.el[k] = (uint8_t)gas;
                }
                ++n;
                ++ngas;
            } else {
                group_sync<W>();  // the last atom may have just been updated by another lane
                if (w == 0 && del_idx != n - 1) {  // all lanes copy the same last atom into the hole
                    const int to = off + del_idx, from = off + n - 1;
                    const double4 xe = R.xe[from], s4 = R.s[from], v4 = R.v[from];
                    const double rho = R.rho[from], F = R.F[from];
                    const uint8_t el = R.el[from];
                    group_sync<1>();
                    R.xe[to] = xe;
                    R.s[to] = s4;
                    R.v[to] = v4;
                    R.rho[to] = rho;
                    R.F[to] = F;
                    R.el[to] = el;
                }
                --n;
                --ngas;
            }
        }
        group_sync<W>();
    }
    if (w == 0) {
        GcmcCell& o = cells[c];
        o.n = n;
        o.ngas = ngas;
        o.seed = rng.seed;
        o.gchk = rng.chk;
        o.gval = rng.keep;
        o.status = status;
    }
}

// Persistent launch: workgroups first take the four-wave cells (largest
// first) from queue[0]; when those are gone each wave takes one-wave cells
// from queue[1] on its own until that queue is empty. Every lane of a wave
// joins the queue atomic (lane 0 adds 1, the others 0) and the ticket is
// read from the first lane, so the cell index is wave-uniform.
__global__ __launch_bounds__(256) void k_gcmc(GcmcArgs A, GcmcCell* cells, const int* __restrict__ q4, int n4,
                                              const int* __restrict__ q1, int n1, int* queue, int nattempts,
                                              int8_t* dec, double* dU, Arrays R)
{
    __shared__ double trial[4][3 * kGcmcMaxTrials];
    __shared__ double red[16];
    __shared__ int next_sh;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    while (true) {
        if (w == 0) {
            const int t = __builtin_amdgcn_readfirstlane(atomicAdd(&queue[0], lane == 0 ? 1 : 0));
            next_sh = t;  // every lane of wave 0 stores the same ticket
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(next_sh);
        __syncthreads();
        if (t >= n4) break;
        run_cell<4>(A, cells, q4[t], nattempts, dec, dU, R, trial[0], red, w, lane);
        __syncthreads();
    }
    while (true) {
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(&queue[1], lane == 0 ? 1 : 0));
        if (t >= n1) break;
        run_cell<1>(A, cells, q1[t], nattempts, dec, dU, R, trial[w], nullptr, 0, lane);
    }
}

// A batch of one cell (the grand-canonical exchange of an MD system, up to
// 16,384 atoms) over one workgroup of 16 waves: each attempt's sweep over the
// atoms takes a quarter of the steps of four waves.
constexpr int kGcmcOneWaves = 16;
__global__ __launch_bounds__(64 * kGcmcOneWaves) void k_gcmc_one(GcmcArgs A, GcmcCell* cells, int nattempts,
                                                                 int8_t* dec, double* dU, Arrays R)
{
    __shared__ double trial[3 * kGcmcMaxTrials];
    __shared__ double red[4 * kGcmcOneWaves];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    run_cell<kGcmcOneWaves>(A, cells, 0, nattempts, dec, dU, R, trial, red, w, lane);
}

__global__ void k_gcmc_reset(int* queue)
{
    if (threadIdx.x < 2) queue[threadIdx.x] = 0;
}

// init_xchng: rho_i over every periodic image of every atom within the
// cutoff, then F(rho_i) without the extension past rho_max. One workgroup
// per cell.
__device__ inline void init_atom(const GcmcArgs& A, const GcmcCell& hd, const Arrays& R, int i)
{
    const int off = hd.off, n = hd.n, ne = A.ne, nr = A.nr;
    const double4 si4 = R.s[off + i];
    const double si[3] = {si4.x, si4.y, si4.z};
    const int ei = (int)R.xe[off + i].w;
    const Images im = count_images(si, hd.lo2, hd.hi2);
    double rho = 0.0;
    for (int k = 0; k < im.total(); ++k) {
        double xi[3];
        image_x(hd.h, si, im, k, xi);
        for (int j = 0; j < n; ++j) {
            if (k == 0 && j == i) continue;
            const double4 xj = R.xe[off + j];
            const int ej = (int)xj.w;
            const double rsq = rsq_of(xi[0], xi[1], xi[2], xj.x, xj.y, xj.z);
            if (rsq >= A.cut_sq[ei * ne + ej]) continue;
            int m;
            double p;
            spline_at(sqrt(rsq), A.dr_inv, nr, m, p);
            rho += hval(A.rho[(size_t)(ej * ne + ei) * nr + m], p);
        }
    }
    R.rho[off + i] = rho;
    R.F[off + i] = f_embed(A, ei, rho, false);
}

__global__ __launch_bounds__(256) void k_gcmc_init(GcmcArgs A, const GcmcCell* __restrict__ cells, Arrays R)
{
    const GcmcCell hd = cells[blockIdx.x];
    for (int i = threadIdx.x; i < hd.n; i += blockDim.x) init_atom(A, hd, R, i);
}

// init_atom with the 64 lanes of a wave splitting the other atoms (lane l
// takes j = l, l + 64, ... for every image; the lanes' sums meet in a fixed
// butterfly, so the result is deterministic)
__device__ inline void init_atom_wave(const GcmcArgs& A, const GcmcCell& hd, const Arrays& R, int i, int lane)
{
    const int off = hd.off, n = hd.n, ne = A.ne, nr = A.nr;
    const double4 si4 = R.s[off + i];
    const double si[3] = {si4.x, si4.y, si4.z};
    const int ei = (int)R.xe[off + i].w;
    const Images im = count_images(si, hd.lo2, hd.hi2);
    double rho = 0.0;
    for (int k = 0; k < im.total(); ++k) {
        double xi[3];
        image_x(hd.h, si, im, k, xi);
        for (int j = lane; j < n; j += 64) {
            if (k == 0 && j == i) continue;
            const double4 xj = R.xe[off + j];
            const int ej = (int)xj.w;
            const double rsq = rsq_of(xi[0], xi[1], xi[2], xj.x, xj.y, xj.z);
            if (rsq >= A.cut_sq[ei * ne + ej]) continue;
            int m;
            double p;
            spline_at(sqrt(rsq), A.dr_inv, nr, m, p);
            rho += hval(A.rho[(size_t)(ej * ne + ei) * nr + m], p);
        }
    }
    for (int o = 32; o > 0; o >>= 1) rho += __shfl_xor(rho, o, 64);
    if (lane == 0) {
        R.rho[off + i] = rho;
        R.F[off + i] = f_embed(A, ei, rho, false);
    }
}

// the same for one cell c over many workgroups, one wave per atom (a large
// cell: the grand-canonical exchange of an MD system, up to 16,384 atoms,
// where one thread per atom would leave most CUs idle)
__global__ __launch_bounds__(256) void k_gcmc_init_cell(GcmcArgs A, const GcmcCell* __restrict__ cells, int c,
                                                        Arrays R)
{
    const GcmcCell hd = cells[c];
    const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);  // uniform over the wave
    if (i < hd.n) init_atom_wave(A, hd, R, i, (int)(threadIdx.x & 63));
}

// ---- single-trial hooks (one cell, one trial, driven by the caller) ----
// k_gcmc_xtrial: the exchange energy of one trial in cell c -- an insertion
// at s (mode MODE_INS) or the deletion of the igas-th gas atom in index order
// (MODE_DEL) -- computed as run_cell computes it with four waves (each thread
// takes atoms t, t + 256, ..., the same sums in the same order). rho_x / F_x
// of every atom of the cell take the trial's values (untouched atoms keep
// theirs). out: du, trial density, deletion index, status (0 ok, 2 no such
// gas atom, 3 too many images), trial position x (image 0), s.
__global__ __launch_bounds__(256) void k_gcmc_xtrial(GcmcArgs A, const GcmcCell* __restrict__ cells, int c, int mode,
                                                     double s0, double s1, double s2, int igas, Arrays R,
                                                     double* __restrict__ out)
{
    __shared__ double trial[3 * kGcmcMaxTrials];
    __shared__ double red[12];
    __shared__ double sb[3];
    __shared__ int sh_del;
    const GcmcCell hd = cells[c];
    const int off = hd.off, n = hd.n, gas = A.gas, ne = A.ne, nr = A.nr;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) {
        int del = -1;
        if (mode == MODE_DEL)
            for (int j = 0, seen = 0; j < n; ++j)
                if (R.el[off + j] == gas) {
                    if (seen == igas) {
                        del = j;
                        break;
                    }
                    ++seen;
                }
        sh_del = del;
        if (mode == MODE_DEL && del >= 0) {
            const double4 s4 = R.s[off + del];
            sb[0] = s4.x;
            sb[1] = s4.y;
            sb[2] = s4.z;
        } else {
            sb[0] = s0;
            sb[1] = s1;
            sb[2] = s2;
        }
    }
    __syncthreads();
    const int del_idx = sh_del;
    const double sv[3] = {sb[0], sb[1], sb[2]};
    const Images im = count_images(sv, hd.lo, hd.hi);
    const int ntrial = im.total();
    // uniform exits: every thread reads the same shared values
    if ((mode == MODE_DEL && del_idx < 0) || ntrial > kGcmcMaxTrials) {
        if (t == 0) {
            out[0] = 0.0;
            out[3] = mode == MODE_DEL && del_idx < 0 ? 2.0 : 3.0;
        }
        return;
    }
    if (t < ntrial) image_x(hd.h, sv, im, t, trial + 3 * t);
    __syncthreads();
    const double c0 = mode == MODE_DEL ? -1.0 : 1.0;
    double en = 0.0, en0 = 0.0, rhoi = 0.0;
    for (int j = t; j < n; j += 256) {
        const double4 xj = R.xe[off + j];
        const int ej = (int)xj.w;
        const double cut = A.cut_sq[gas * ne + ej];
        const double4* tp = A.phi + (size_t)(gas * ne + ej) * nr;
        const double4* tr_ij = A.rho + (size_t)(gas * ne + ej) * nr;
        const double4* tr_ji = A.rho + (size_t)(ej * ne + gas) * nr;
        bool hit = false;
        double rx = 0.0;
        for (int k = 0; k < ntrial; ++k) {
            const double rsq = rsq_of(trial[3 * k], trial[3 * k + 1], trial[3 * k + 2], xj.x, xj.y, xj.z);
            if (rsq < cut && !(k == 0 && j == del_idx)) {
                const double r = sqrt(rsq);
                int m;
                double p;
                spline_at(r, A.dr_inv, nr, m, p);
                en += hval(tp[m], p) / r;
                if (!hit) {
                    rx = R.rho[off + j];
                    hit = true;
                }
                rx += c0 * hval(tr_ij[m], p);
                rhoi += hval(tr_ji[m], p);
            }
        }
        if (hit) {
            const double fx = f_embed(A, ej, rx, true);
            en0 += fx - R.F[off + j];
            R.rho_x[off + j] = rx;
            R.F_x[off + j] = fx;
        } else {
            R.rho_x[off + j] = R.rho[off + j];
            R.F_x[off + j] = R.F[off + j];
        }
    }
    // the trial point's own images, half of each pair (wave 0, as run_cell)
    if (w == 0 && lane >= 1 && lane < ntrial) {
        const double rsq = rsq_of(trial[0], trial[1], trial[2], trial[3 * lane], trial[3 * lane + 1], trial[3 * lane + 2]);
        if (rsq < A.cut_sq[gas * ne + gas]) {
            const double r = sqrt(rsq);
            int m;
            double p;
            spline_at(r, A.dr_inv, nr, m, p);
            en += 0.5 * hval(A.phi[(size_t)(gas * ne + gas) * nr + m], p) / r;
            rhoi += hval(A.rho[(size_t)(gas * ne + gas) * nr + m], p);
        }
    }
    en = wave_sum(en);
    en0 = wave_sum(en0);
    rhoi = wave_sum(rhoi);
    if (lane == 0) {
        red[3 * w] = en;
        red[3 * w + 1] = en0;
        red[3 * w + 2] = rhoi;
    }
    __syncthreads();
    if (t == 0) {
        en = en0 = rhoi = 0.0;
        for (int q = 0; q < 4; ++q) {
            en += red[3 * q];
            en0 += red[3 * q + 1];
            rhoi += red[3 * q + 2];
        }
        out[0] = en + en0 * c0 + f_embed(A, gas, rhoi, true);
        out[1] = rhoi;
        out[2] = (double)del_idx;
        out[3] = 0.0;
        for (int d = 0; d < 3; ++d) {
            out[4 + d] = trial[d];
            out[7 + d] = sv[d];
        }
    }
}

// k_gcmc_xcommit: an accepted trial of cell c (post_xchng_energy, then
// ins_succ / del_succ): every atom takes its trial rho / F; an insertion
// appends the gas atom at x (fractional s, velocity v) with the trial
// density; a deletion moves the last atom into the hole.
__global__ __launch_bounds__(256) void k_gcmc_xcommit(GcmcArgs A, GcmcCell* __restrict__ cells, int c, int mode,
                                                      int del_idx, double rhoi, double x0, double x1, double x2,
                                                      double s0, double s1, double s2, double v0, double v1,
                                                      double v2, Arrays R)
{
    const int off = cells[c].off, n = cells[c].n, gas = A.gas;
    for (int j = threadIdx.x; j < n; j += 256) {
        R.rho[off + j] = R.rho_x[off + j];
        R.F[off + j] = R.F_x[off + j];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (mode == MODE_INS) {
        const int k = off + n;
        R.xe[k] = make_double4(x0, x1, x2, (double)gas);
        R.s[k] = make_double4(s0, s1, s2, 0.0);
        R.v[k] = make_double4(v0, v1, v2, 0.0);
        R.rho[k] = rhoi;
        R.F[k] = f_embed(A, gas, rhoi, true);
        R.el[k] = (uint8_t)gas;
        cells[c].n = n + 1;
        cells[c].ngas += 1;
    } else {
        if (del_idx != n - 1) {
            const int to = off + del_idx, from = off + n - 1;
            R.xe[to] = R.xe[from];
            R.s[to] = R.s[from];
            R.v[to] = R.v[from];
            R.rho[to] = R.rho[from];
            R.F[to] = R.F[from];
            R.el[to] = R.el[from];
        }
        cells[c].n = n - 1;
        cells[c].ngas -= 1;
    }
}

Arrays arrays_of(GcmcState& st)
{
    return Arrays{st.xe.get(), st.s.get(), st.v.get(), st.rho.get(), st.F.get(), st.rho_x.get(), st.F_x.get(),
                  st.el.get()};
}

void alloc_state(GcmcState& st, int nc, size_t cap)
{
    st.cells.reserve(nc);
    st.xe.reserve(cap);
    st.s.reserve(cap);
    st.v.reserve(cap);
    st.rho.reserve(cap);
    st.F.reserve(cap);
    st.rho_x.reserve(cap);
    st.F_x.reserve(cap);
    st.el.reserve(cap);
}

}  // namespace

GcmcBatch::GcmcBatch(std::shared_ptr<EamForceField> ff, int ncells, const int* natms, const double* H,
                     const double* x, const uint8_t* elem, const int* cap, int gas, double mu, double T, double kB,
                     double hP, const double* masses, const int* seeds)
    : ff_(std::move(ff)), nc_(ncells)
{
    const EamTables& tb = ff_->tables();
    MAPP_REQUIRE(!tb.dmd, ERR_ARG, "grand-canonical exchange needs a plain (non-DMD) EAM force field");
    MAPP_REQUIRE(ncells > 0, ERR_ARG, "ncells must be positive");
    MAPP_REQUIRE(gas >= 0 && gas < tb.ne, ERR_ARG, "gas element index out of range");
    MAPP_REQUIRE(T > 0.0 && kB > 0.0 && hP > 0.0 && masses[gas] > 0.0, ERR_ARG,
                 "T, kB, hP and the gas mass must be positive");
    const int ne = tb.ne;
    A_.ne = ne;
    A_.nr = tb.nr;
    A_.nrho = tb.nrho;
    A_.gas = gas;
    A_.dr_inv = 1.0 / tb.dr;
    A_.drho_inv = 1.0 / tb.drho;
    A_.rho_max = (double)tb.nrho * tb.drho;
    for (int i = 0; i < ne * ne; ++i) A_.cut_sq[i] = tb.rc[i] * tb.rc[i];
    // GCMC::init (src/gcmc.cpp:86-116), in the reference's operation order
    double cut = tb.rc[(size_t)gas * ne + 0];
    for (int i = 1; i < ne; ++i) cut = std::max(cut, tb.rc[(size_t)gas * ne + i]);
    const double kbT = kB * T;
    A_.beta = 1.0 / kbT;
    const double lambda = hP / std::sqrt(2.0 * M_PI * kbT * masses[gas]);
    A_.sigma = std::sqrt(kbT / masses[gas]);
    double z_fac = 1.0;
    for (int i = 0; i < 3; ++i) z_fac /= lambda;
    z_fac *= std::exp(A_.beta * mu);
    const double max_cut = tb.max_rc();

    // tables in FP64
    {
        std::vector<double4> ph((size_t)ne * ne * tb.nr), rh((size_t)ne * ne * tb.nr), fh((size_t)ne * tb.nrho);
        for (int a = 0; a < ne; ++a)
            for (int b = 0; b < ne; ++b)
                for (int k = 0; k < tb.nr; ++k) {
                    const double* p = tb.rphi_at(a, b) + 4 * k;
                    const double* r = tb.rho_at(a, b) + 4 * k;
                    ph[((size_t)a * ne + b) * tb.nr + k] = make_double4(p[0], p[1], p[2], p[3]);
                    rh[((size_t)a * ne + b) * tb.nr + k] = make_double4(r[0], r[1], r[2], r[3]);
                }
        for (int a = 0; a < ne; ++a)
            for (int k = 0; k < tb.nrho; ++k) {
                const double* f = tb.F_at(a) + (size_t)tb.nf() * k;
                fh[(size_t)a * tb.nrho + k] = make_double4(f[0], f[1], f[2], f[3]);
            }
        phi_.reserve(ph.size());
        rho4_.reserve(rh.size());
        F4_.reserve(fh.size());
        MAPP_HIP(hipMemcpy(phi_.get(), ph.data(), ph.size() * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(rho4_.get(), rh.data(), rh.size() * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(F4_.get(), fh.data(), fh.size() * sizeof(double4), hipMemcpyHostToDevice));
        A_.phi = phi_.get();
        A_.rho = rho4_.get();
        A_.F = F4_.get();
    }

    // per-cell constants, coordinates wrapped into the box (x2s_lcl, s2x)
    host_cells_.resize(ncells);
    n0_.resize(ncells);
    size_t off = 0, natot = 0;
    for (int c = 0; c < ncells; ++c) {
        MAPP_REQUIRE(natms[c] >= 0, ERR_ARG, "negative atom count");
        const int cp = cap ? cap[c] : natms[c] + std::max(32, natms[c] / 8);
        MAPP_REQUIRE(cp >= natms[c], ERR_ARG, "cell capacity below its atom count");
        MAPP_REQUIRE(cp <= 4 * kGcmcWaveCap, ERR_CAPACITY,
                     "a cell may hold at most " + std::to_string(4 * kGcmcWaveCap) + " atoms");
        GcmcCell& hc = host_cells_[c];
        std::memset(&hc, 0, sizeof(hc));
        const Box bx = make_box(H + 9 * (size_t)c);
        MAPP_REQUIRE(bx.H[0][0] > 0.0 && bx.H[1][1] > 0.0 && bx.H[2][2] > 0.0, ERR_ARG,
                     "box diagonal must be positive");
        hc.off = (int)off;
        hc.cap = cp;
        hc.n = natms[c];
        hc.seed = seeds[c];
        hc.cls = gcmc_class_of(natms[c]);
        for (int k = 0; k < 6; ++k) hc.h[k] = bx.h[k];
        double vol = 1.0;
        for (int i = 0; i < 3; ++i) vol *= bx.H[i][i];
        hc.zv = z_fac * vol;
        int ntr = 1;
        for (int i = 0; i < 3; ++i) {
            double t = 0.0;
            for (int j = i; j < 3; ++j) t += bx.B[j][i] * bx.B[j][i];
            const double cs = std::sqrt(t) * cut, cs2 = std::sqrt(t) * max_cut;
            hc.lo[i] = 0.0 - cs;
            hc.hi[i] = 1.0 + cs;
            hc.lo2[i] = 0.0 - cs2;
            hc.hi2[i] = 1.0 + cs2;
            ntr *= 1 + (int)std::floor(hc.hi[i]) - (int)std::floor(hc.lo[i]);
            MAPP_REQUIRE(1 + (int)std::floor(hc.hi2[i]) - (int)std::floor(hc.lo2[i]) <= 16, ERR_CAPACITY,
                         "box too small for its cutoff (more than 16 images along one direction)");
        }
        MAPP_REQUIRE(ntr <= kGcmcMaxTrials, ERR_CAPACITY,
                     "box too small for the gas cutoff: more than " + std::to_string(kGcmcMaxTrials) +
                         " periodic images of a trial point");
        n0_[c] = natms[c];
        off += cp;
        natot += natms[c];
    }
    total_cap_ = off;
    alloc_state(st_, ncells, total_cap_);
    {
        std::vector<double4> xe(total_cap_), sv(total_cap_), vv(total_cap_, make_double4(0, 0, 0, 0));
        std::vector<uint8_t> el(total_cap_, 0);
        size_t src = 0;
        for (int c = 0; c < ncells; ++c) {
            GcmcCell& hc = host_cells_[c];
            const Box bx = make_box(H + 9 * (size_t)c);
            for (int i = 0; i < hc.n; ++i, ++src) {
                const uint8_t e = elem[src];
                MAPP_REQUIRE(e < ne, ERR_ARG, "element index out of range");
                double s0, s1, s2, x0, x1, x2;
                x_to_s(bx.b, x[3 * src], x[3 * src + 1], x[3 * src + 2], s0, s1, s2);
                s_to_x(bx.h, s0, s1, s2, x0, x1, x2);
                xe[hc.off + i] = make_double4(x0, x1, x2, (double)e);
                sv[hc.off + i] = make_double4(s0, s1, s2, 0.0);
                el[hc.off + i] = e;
                if (e == gas) ++hc.ngas;
            }
        }
        (void)natot;
        MAPP_HIP(hipMemcpy(st_.xe.get(), xe.data(), total_cap_ * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.s.get(), sv.data(), total_cap_ * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.v.get(), vv.data(), total_cap_ * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.el.get(), el.data(), total_cap_, hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.cells.get(), host_cells_.data(), ncells * sizeof(GcmcCell), hipMemcpyHostToDevice));
    }
    MAPP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    queue_.reserve(2);
    if (ncells == 1) {
        if (host_cells_[0].n)
            k_gcmc_init_cell<<<(host_cells_[0].n + 3) / 4, 256, 0, stream_>>>(A_, st_.cells.get(), 0,
                                                                             arrays_of(st_));
    } else {
        k_gcmc_init<<<ncells, 256, 0, stream_>>>(A_, st_.cells.get(), arrays_of(st_));
    }
    MAPP_HIP(hipGetLastError());
    int dev = 0, ncu = 0, per = 0;
    MAPP_HIP(hipGetDevice(&dev));
    MAPP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MAPP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_gcmc, 256, 0));
    grid_ = std::max(1, ncu * std::max(1, per));
    MAPP_HIP(hipStreamSynchronize(stream_));
}

GcmcBatch::~GcmcBatch()
{
    if (graph_) (void)hipGraphExecDestroy(graph_);
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void GcmcBatch::set_waves(const int* waves)
{
    for (int k = 0; k < kGcmcClasses; ++k)
        MAPP_REQUIRE(waves[k] == 1 || waves[k] == 4, ERR_ARG, "waves per cell must be 1 or 4");
    for (int k = 0; k < kGcmcClasses; ++k) waves_[k] = waves[k];
    queues_dirty_ = true;
}

// Cells of each shape, largest first (longest chains start first).
void GcmcBatch::split_cells(const int* waves, std::vector<int>& q4, std::vector<int>& q1) const
{
    std::vector<int> order(nc_);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n0_[a] > n0_[b]; });
    q4.clear();
    q1.clear();
    for (int c : order) {
        const int wv = host_cells_[c].cap > kGcmcWaveCap ? 4 : waves[host_cells_[c].cls];
        (wv == 4 ? q4 : q1).push_back(c);
    }
}

void GcmcBatch::build_queues()
{
    std::vector<int> q4, q1;
    split_cells(waves_, q4, q1);
    n4_ = (int)q4.size();
    n1_ = (int)q1.size();
    q4_.reserve(std::max(1, n4_));
    q1_.reserve(std::max(1, n1_));
    if (n4_) MAPP_HIP(hipMemcpy(q4_.get(), q4.data(), n4_ * sizeof(int), hipMemcpyHostToDevice));
    if (n1_) MAPP_HIP(hipMemcpy(q1_.get(), q1.data(), n1_ * sizeof(int), hipMemcpyHostToDevice));
    queues_dirty_ = false;
    if (graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
        graph_k_ = -1;
    }
}

void GcmcBatch::launch(GcmcState& st, int nattempts, int8_t* dec, double* du, const int* q4, int n4, const int* q1,
                       int n1)
{
    if (nc_ == 1) {
        k_gcmc_one<<<1, 64 * kGcmcOneWaves, 0, stream_>>>(A_, st.cells.get(), nattempts, dec, du, arrays_of(st));
        MAPP_HIP(hipGetLastError());
        return;
    }
    k_gcmc_reset<<<1, 64, 0, stream_>>>(queue_.get());
    const int blocks = std::max(1, std::min(grid_, std::max(n4, (n1 + 3) / 4)));
    k_gcmc<<<blocks, 256, 0, stream_>>>(A_, st.cells.get(), q4, n4, q1, n1, queue_.get(), nattempts, dec, du,
                                        arrays_of(st));
    MAPP_HIP(hipGetLastError());
}

void GcmcBatch::use_graph(bool on)
{
    graph_on_ = on;
    if (!on && graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
        graph_k_ = -1;
    }
}

void GcmcBatch::check_status(GcmcState& st)
{
    std::vector<GcmcCell> hc(nc_);
    MAPP_HIP(hipMemcpy(hc.data(), st.cells.get(), nc_ * sizeof(GcmcCell), hipMemcpyDeviceToHost));
    for (int c = 0; c < nc_; ++c) {
        MAPP_REQUIRE(hc[c].status != 1, ERR_CAPACITY,
                     "cell " + std::to_string(c) + " is full (" + std::to_string(hc[c].cap) +
                         " atoms): an accepted insertion had no room; its chain stopped there");
        MAPP_REQUIRE(hc[c].status == 0, ERR_STATE,
                     "cell " + std::to_string(c) + " chain stopped: internal consistency check " +
                         std::to_string(hc[c].status) + " failed");
    }
}

void GcmcBatch::attempts(int nattempts, int8_t* decisions, double* dU)
{
    MAPP_REQUIRE(nattempts >= 0, ERR_ARG, "nattempts must be non-negative");
    if (nattempts == 0) return;
    ++ver_;
    if (queues_dirty_) build_queues();
    const size_t m = (size_t)nc_ * nattempts;
    dec_.reserve(m);
    du_.reserve(m);
    prof_.begin(K_GCMC, stream_);
    if (graph_on_) {
        if (!graph_ || graph_k_ != nattempts) {
            if (graph_) (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
            hipGraph_t g;
            MAPP_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            launch(st_, nattempts, dec_.get(), du_.get(), q4_.get(), n4_, q1_.get(), n1_);
            MAPP_HIP(hipStreamEndCapture(stream_, &g));
            MAPP_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            MAPP_HIP(hipGraphDestroy(g));
            graph_k_ = nattempts;
        }
        MAPP_HIP(hipGraphLaunch(graph_, stream_));
    } else {
        launch(st_, nattempts, dec_.get(), du_.get(), q4_.get(), n4_, q1_.get(), n1_);
    }
    prof_.end(stream_);
    MAPP_HIP(hipStreamSynchronize(stream_));
    if (decisions) MAPP_HIP(hipMemcpy(decisions, dec_.get(), m, hipMemcpyDeviceToHost));
    if (dU) MAPP_HIP(hipMemcpy(dU, du_.get(), m * sizeof(double), hipMemcpyDeviceToHost));
    check_status(st_);
}

void GcmcBatch::copy_state(GcmcState& dst, const GcmcState& src)
{
    alloc_state(dst, nc_, total_cap_);
    MAPP_HIP(hipMemcpyAsync(dst.cells.get(), src.cells.get(), nc_ * sizeof(GcmcCell), hipMemcpyDeviceToDevice,
                            stream_));
    MAPP_HIP(hipMemcpyAsync(dst.xe.get(), src.xe.get(), total_cap_ * sizeof(double4), hipMemcpyDeviceToDevice,
                            stream_));
    MAPP_HIP(hipMemcpyAsync(dst.s.get(), src.s.get(), total_cap_ * sizeof(double4), hipMemcpyDeviceToDevice,
                            stream_));
    MAPP_HIP(hipMemcpyAsync(dst.v.get(), src.v.get(), total_cap_ * sizeof(double4), hipMemcpyDeviceToDevice,
                            stream_));
    MAPP_HIP(hipMemcpyAsync(dst.rho.get(), src.rho.get(), total_cap_ * sizeof(double), hipMemcpyDeviceToDevice,
                            stream_));
    MAPP_HIP(hipMemcpyAsync(dst.F.get(), src.F.get(), total_cap_ * sizeof(double), hipMemcpyDeviceToDevice,
                            stream_));
    MAPP_HIP(hipMemcpyAsync(dst.el.get(), src.el.get(), total_cap_, hipMemcpyDeviceToDevice, stream_));
}

// Tile tuner: the launch shape is one wave per cell for the small size
// classes and four waves per cell from a threshold class up. Every threshold
// is timed on the whole batch (a scratch copy, so the production chains and
// their random streams are not advanced) and the fastest is kept.
GcmcTuneReport GcmcBatch::tune(int nattempts)
{
    MAPP_REQUIRE(nattempts > 0, ERR_ARG, "nattempts must be positive");
    GcmcTuneReport rep{};
    hipEvent_t e0, e1;
    MAPP_HIP(hipEventCreate(&e0));
    MAPP_HIP(hipEventCreate(&e1));
    DevBuf<int> d4, d1;
    d4.reserve(nc_);
    d1.reserve(nc_);
    double best = 1e300;
    int best_t = kGcmcClasses;
    std::vector<int> q4, q1;
    // one untimed launch first (code objects, caches), then the best of two
    // timed launches per candidate: a single timing let first-launch costs
    // decide the split
    for (int th = -1; th <= kGcmcClasses; ++th) {
        int wv[kGcmcClasses];
        for (int k = 0; k < kGcmcClasses; ++k) wv[k] = k < th ? 1 : 4;
        split_cells(wv, q4, q1);
        if (!q4.empty()) MAPP_HIP(hipMemcpy(d4.get(), q4.data(), q4.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!q1.empty()) MAPP_HIP(hipMemcpy(d1.get(), q1.data(), q1.size() * sizeof(int), hipMemcpyHostToDevice));
        float ms = 1e30f;
        for (int rep_i = 0; rep_i < (th < 0 ? 1 : 2); ++rep_i) {
            copy_state(scratch_, st_);
            MAPP_HIP(hipEventRecord(e0, stream_));
            launch(scratch_, nattempts, nullptr, nullptr, d4.get(), (int)q4.size(), d1.get(), (int)q1.size());
            MAPP_HIP(hipEventRecord(e1, stream_));
            MAPP_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            MAPP_HIP(hipEventElapsedTime(&t, e0, e1));
            ms = std::min(ms, t);
        }
        if (th < 0) continue;
        rep.ms[th] = ms;
        if (ms < best) {
            best = ms;
            best_t = th;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < kGcmcClasses; ++k) rep.waves[k] = k < best_t ? 1 : 4;
    set_waves(rep.waves);
    return rep;
}

// ForceFieldEAM::pre_xchng_energy + xchng_energy (src/ff_eam.cpp:323-446) for
// one trial of one cell, as SGCMC::attmpt (src/sgcmc.cpp:255-327) calls them:
// mode 1 inserts the gas element at fractional s, mode 2 deletes the igas-th
// gas atom. The trial is kept for xchng_commit.
double GcmcBatch::xchng_energy(int cell, int mode, const double* s, int igas)
{
    MAPP_REQUIRE(cell >= 0 && cell < nc_, ERR_ARG, "cell index out of range");
    MAPP_REQUIRE(mode == 1 || mode == 2, ERR_ARG, "mode must be 1 (insertion) or 2 (deletion)");
    if (mode == 1) {
        MAPP_REQUIRE(s != nullptr, ERR_ARG, "an insertion needs its fractional position s");
        for (int d = 0; d < 3; ++d)
            MAPP_REQUIRE(s[d] >= 0.0 && s[d] < 1.0, ERR_ARG, "s must lie in [0, 1) in every dimension");
    } else {
        MAPP_REQUIRE(igas >= 0, ERR_ARG, "igas must be non-negative");
    }
    pending_.cell = -1;
    xout_.reserve(10);
    const double s0 = mode == 1 ? s[0] : 0.0, s1 = mode == 1 ? s[1] : 0.0, s2 = mode == 1 ? s[2] : 0.0;
    k_gcmc_xtrial<<<1, 256, 0, stream_>>>(A_, st_.cells.get(), cell, mode == 1 ? MODE_INS : MODE_DEL, s0, s1, s2,
                                          igas, arrays_of(st_), xout_.get());
    double o[10];
    MAPP_HIP(hipMemcpyAsync(o, xout_.get(), sizeof o, hipMemcpyDeviceToHost, stream_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    MAPP_REQUIRE(o[3] != 2.0, ERR_ARG, "igas is not below the cell's gas count");
    MAPP_REQUIRE(o[3] == 0.0, ERR_CAPACITY, "the trial point has more periodic images than the kernel holds");
    pending_ = PendingTrial{cell, mode, (int)o[2], o[1], {o[4], o[5], o[6]}, {o[7], o[8], o[9]}, ver_};
    return o[0];
}

// ForceFieldEAM::post_xchng_energy (src/ff_eam.cpp:451-483) after
// SGCMC::ins_succ / del_succ (src/sgcmc.cpp:331-444) for the trial of the last
// xchng_energy call, which the caller accepted; v is the inserted atom's
// velocity (the reference draws it as gaussian() * sigma per component).
void GcmcBatch::xchng_commit(const double* v)
{
    MAPP_REQUIRE(pending_.cell >= 0 && pending_.ver == ver_, ERR_STATE,
                 "no trial to commit: call xchng_energy first (the cells changed since, or it was committed)");
    const int c = pending_.cell;
    GcmcCell hc;
    MAPP_HIP(hipMemcpy(&hc, st_.cells.get() + c, sizeof hc, hipMemcpyDeviceToHost));
    if (pending_.mode == 1) {
        MAPP_REQUIRE(v != nullptr, ERR_ARG, "an insertion needs the new atom's velocity");
        MAPP_REQUIRE(hc.n < hc.cap, ERR_CAPACITY, "the cell is full: the insertion cannot be committed");
    }
    const double v0 = v ? v[0] : 0.0, v1 = v ? v[1] : 0.0, v2 = v ? v[2] : 0.0;
    k_gcmc_xcommit<<<1, 256, 0, stream_>>>(A_, st_.cells.get(), c, pending_.mode == 1 ? MODE_INS : MODE_DEL,
                                           pending_.del_idx, pending_.rhoi, pending_.x[0], pending_.x[1],
                                           pending_.x[2], pending_.s[0], pending_.s[1], pending_.s[2], v0, v1, v2,
                                           arrays_of(st_));
    MAPP_HIP(hipStreamSynchronize(stream_));
    pending_.cell = -1;
    ++ver_;
}

void GcmcBatch::counts(int* natms, int* ngas)
{
    std::vector<GcmcCell> hc(nc_);
    MAPP_HIP(hipMemcpy(hc.data(), st_.cells.get(), nc_ * sizeof(GcmcCell), hipMemcpyDeviceToHost));
    for (int c = 0; c < nc_; ++c) {
        if (natms) natms[c] = hc[c].n;
        if (ngas) ngas[c] = hc[c].ngas;
    }
}

void GcmcBatch::load_cell(int c, int n, const double* x, const uint8_t* elem, const double* v)
{
    MAPP_REQUIRE(c >= 0 && c < nc_, ERR_ARG, "cell index out of range");
    GcmcCell hc;
    MAPP_HIP(hipStreamSynchronize(stream_));
    MAPP_HIP(hipMemcpy(&hc, st_.cells.get() + c, sizeof(GcmcCell), hipMemcpyDeviceToHost));
    MAPP_REQUIRE(n >= 0 && n <= hc.cap, ERR_CAPACITY,
                 "cell " + std::to_string(c) + " holds at most " + std::to_string(hc.cap) + " atoms");
    const int ne = A_.ne;
    double b6[6];
    {
        double H9[9] = {hc.h[0], 0.0, 0.0, hc.h[1], hc.h[3], 0.0, hc.h[2], hc.h[4], hc.h[5]};
        const Box bx = make_box(H9);
        for (int k = 0; k < 6; ++k) b6[k] = bx.b[k];
    }
    std::vector<double4> xe(std::max(n, 1)), sv(std::max(n, 1)), vv(std::max(n, 1));
    std::vector<uint8_t> el(std::max(n, 1));
    int ngas = 0;
    for (int i = 0; i < n; ++i) {
        const uint8_t e = elem[i];
        MAPP_REQUIRE(e < ne, ERR_ARG, "element index out of range");
        double s0, s1, s2, x0, x1, x2;
        x_to_s(b6, x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], s0, s1, s2);
        s_to_x(hc.h, s0, s1, s2, x0, x1, x2);
        xe[i] = make_double4(x0, x1, x2, (double)e);
        sv[i] = make_double4(s0, s1, s2, 0.0);
        vv[i] = v ? make_double4(v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2], 0.0)
                  : make_double4(0.0, 0.0, 0.0, 0.0);
        el[i] = e;
        if (e == A_.gas) ++ngas;
    }
    hc.n = n;
    hc.ngas = ngas;
    hc.status = 0;
    if (n) {
        MAPP_HIP(hipMemcpy(st_.xe.get() + hc.off, xe.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.s.get() + hc.off, sv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.v.get() + hc.off, vv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        MAPP_HIP(hipMemcpy(st_.el.get() + hc.off, el.data(), n, hipMemcpyHostToDevice));
    }
    MAPP_HIP(hipMemcpy(st_.cells.get() + c, &hc, sizeof(GcmcCell), hipMemcpyHostToDevice));
    host_cells_[c].n = n;
    host_cells_[c].ngas = ngas;
    if (n) {
        k_gcmc_init_cell<<<(n + 3) / 4, 256, 0, stream_>>>(A_, st_.cells.get(), c, arrays_of(st_));
        MAPP_HIP(hipGetLastError());
    }
    MAPP_HIP(hipStreamSynchronize(stream_));
    ++ver_;
    pending_.cell = -1;
}

void GcmcBatch::get_cell(int c, double* x, uint8_t* elem, double* rho, double* F, double* v, int* seed)
{
    MAPP_REQUIRE(c >= 0 && c < nc_, ERR_ARG, "cell index out of range");
    GcmcCell hc;
    MAPP_HIP(hipMemcpy(&hc, st_.cells.get() + c, sizeof(GcmcCell), hipMemcpyDeviceToHost));
    const int n = hc.n;
    if (seed) *seed = hc.seed;
    if (n == 0) return;
    std::vector<double4> b(n);
    if (x || elem) {
        MAPP_HIP(hipMemcpy(b.data(), st_.xe.get() + hc.off, n * sizeof(double4), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            if (x) {
                x[3 * i] = b[i].x;
                x[3 * i + 1] = b[i].y;
                x[3 * i + 2] = b[i].z;
            }
            if (elem) elem[i] = (uint8_t)b[i].w;
        }
    }
    if (v) {
        MAPP_HIP(hipMemcpy(b.data(), st_.v.get() + hc.off, n * sizeof(double4), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            v[3 * i] = b[i].x;
            v[3 * i + 1] = b[i].y;
            v[3 * i + 2] = b[i].z;
        }
    }
    if (rho) MAPP_HIP(hipMemcpy(rho, st_.rho.get() + hc.off, n * sizeof(double), hipMemcpyDeviceToHost));
    if (F) MAPP_HIP(hipMemcpy(F, st_.F.get() + hc.off, n * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace mapp
