// md_min.hip -- energy minimisation of MD systems on the GPU: conjugate
// gradients (MinCG::__run, src/min_cg.h:453-533) and L-BFGS
// (MinLBFGS::__run, src/min_l-bfgs.h:114-217) with the back-tracking line
// search (LineSearchBackTrack::min, src/ls.h:711-769) over the device EAM
// passes. The host keeps the scalar control (step lengths, two-loop
// coefficients); each vector operation over the atoms is one streaming
// kernel over slot-ordered double4 vectors that also writes the workgroup
// partial sums of the inner products the next host decision needs, so an
// L-BFGS direction with k stored pairs costs 2k + 1 passes over the vectors.
#include <algorithm>
#include <cmath>
#include <limits>

#include "md.hpp"
#include "kern.hpp"
#include "linesearch.hpp"

namespace mapp {

namespace {

constexpr int kMinRows = 2048;  // grid-stride vector kernels: at most this many workgroups


__device__ inline double4 fdbl(float4 f) { return make_double4((double)f.x, (double)f.y, (double)f.z, 0.0); }
__device__ inline double dot3(double4 a, double4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double amax3(double4 a) { return fmax(fabs(a.x), fmax(fabs(a.y), fabs(a.z))); }

// K sums and one maximum (of non-negative values) over a workgroup, written
// as one row of kPartW: the sums in columns 0..K-1, the maximum in column K
template <int K>
__device__ inline void block_sums_max_store(double (&v)[K], double m, double* out)
{
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, 64);
        m = fmax(m, __shfl_xor(m, off, 64));
    }
    __shared__ double sh[kBlock / 64][K + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        for (int k = 0; k < K; ++k) sh[w][k] = v[k];
        sh[w][K] = m;
    }
    __syncthreads();
    if (threadIdx.x < kPartW) {
        double s = 0.0;
        if ((int)threadIdx.x < K)
            for (int q = 0; q < kBlock / 64; ++q) s += sh[q][threadIdx.x];
        else if ((int)threadIdx.x == K)
            for (int q = 0; q < kBlock / 64; ++q) s = fmax(s, sh[q][K]);
        out[threadIdx.x] = s;
    }
}

// one workgroup: columns below max_col summed, column max_col (if < kPartW)
// maximised, in a fixed order
__global__ __launch_bounds__(kBlock) void k_min_rows(int nrows, int max_col, const double* __restrict__ part,
                                                     double* __restrict__ out)
{
    __shared__ double sh[kBlock][kPartW];
    double acc[kPartW];
    for (int k = 0; k < kPartW; ++k) acc[k] = 0.0;
    for (int r = threadIdx.x; r < nrows; r += kBlock)
        for (int k = 0; k < kPartW; ++k) {
            const double u = part[(size_t)r * kPartW + k];
            acc[k] = k == max_col ? fmax(acc[k], u) : acc[k] + u;
        }
    for (int k = 0; k < kPartW; ++k) sh[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < kPartW; ++k)
                sh[threadIdx.x][k] = k == max_col ? fmax(sh[threadIdx.x][k], sh[threadIdx.x + s][k])
                                                  : sh[threadIdx.x][k] + sh[threadIdx.x + s][k];
        __syncthreads();
    }
    if ((int)threadIdx.x < kPartW) out[threadIdx.x] = sh[0][threadIdx.x];
}

// start of an iteration: h = f (cg_ratio unused) or h = ratio h + f
// (use_ratio), x0 = x, f0 = f. Sums: v.h (v optional), f.h, h.h, f.f; max |h|.
__global__ __launch_bounds__(kBlock) void k_min_begin(int n, int use_ratio, double ratio,
                                                      const float4* __restrict__ f, double4* __restrict__ h,
                                                      const double4* __restrict__ x, double4* __restrict__ x0,
                                                      double4* __restrict__ f0, const double4* __restrict__ v,
                                                      double* __restrict__ part)
{
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double mx = 0.0;
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const double4 fi = fdbl(f[i]);
        double4 hi = fi;
        if (use_ratio) {
            const double4 ho = h[i];
            hi = make_double4(ratio * ho.x + fi.x, ratio * ho.y + fi.y, ratio * ho.z + fi.z, 0.0);
        }
        h[i] = hi;
        x0[i] = x[i];
        f0[i] = fi;
        if (v) acc[0] += dot3(v[i], hi);
        acc[1] += dot3(fi, hi);
        acc[2] += dot3(hi, hi);
        acc[3] += dot3(fi, fi);
        mx = fmax(mx, amax3(hi));
    }
    block_sums_max_store<4>(acc, mx, part + (size_t)blockIdx.x * kPartW);
}

// h = (h + a u) g (u optional). Sums: v.h (v optional), f.h, h.h, f.f; max |h|.
// With d_dot set the coefficient comes from reduced sums already on the
// device (column 0 of earlier rows), so the two-loop recursion runs without a
// host round trip: a = rho (d_al[0] if given, else 0) - rho d_dot[0], i.e.
// alpha_i = -rho_i (s_i.h) in the first loop and -alpha_i - rho_i (y_i.h) in
// the second (src/min_l-bfgs.h:150-162).
__global__ __launch_bounds__(kBlock) void k_min_axpy(int n, double a, const double* __restrict__ d_dot,
                                                     const double* __restrict__ d_al, double rho,
                                                     const double4* __restrict__ u, double g,
                                                     double4* __restrict__ h, const double4* __restrict__ v,
                                                     const float4* __restrict__ f, double* __restrict__ part)
{
    if (d_dot) a = (d_al ? rho * d_al[0] : 0.0) - rho * d_dot[0];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double mx = 0.0;
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double4 hi = h[i];
        if (u) {
            const double4 ui = u[i];
            hi = make_double4(hi.x + a * ui.x, hi.y + a * ui.y, hi.z + a * ui.z, 0.0);
        }
        hi = make_double4(hi.x * g, hi.y * g, hi.z * g, 0.0);
        h[i] = hi;
        const double4 fi = fdbl(f[i]);
        if (v) acc[0] += dot3(v[i], hi);
        acc[1] += dot3(fi, hi);
        acc[2] += dot3(hi, hi);
        acc[3] += dot3(fi, fi);
        mx = fmax(mx, amax3(hi));
    }
    block_sums_max_store<4>(acc, mx, part + (size_t)blockIdx.x * kPartW);
}

// after an accepted step: s = t h, y = f0 - f (s optional). Sums: s.y, y.y,
// h.f0, h.f, f.f, f0.f0, f.f0; max |f|.
__global__ __launch_bounds__(kBlock) void k_min_pair(int n, double t, const double4* __restrict__ h,
                                                     const double4* __restrict__ f0, const float4* __restrict__ f,
                                                     double4* __restrict__ s, double4* __restrict__ y,
                                                     double* __restrict__ part)
{
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double mx = 0.0;
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const double4 fi = fdbl(f[i]), gi = f0[i], hi = h[i];
        if (s) {
            const double4 si = make_double4(t * hi.x, t * hi.y, t * hi.z, 0.0);
            const double4 yi = make_double4(gi.x - fi.x, gi.y - fi.y, gi.z - fi.z, 0.0);
            s[i] = si;
            y[i] = yi;
            acc[0] += dot3(si, yi);
            acc[1] += dot3(yi, yi);
        }
        acc[2] += dot3(hi, gi);
        acc[3] += dot3(hi, fi);
        acc[4] += dot3(fi, fi);
        acc[5] += dot3(gi, gi);
        acc[6] += dot3(fi, gi);
        mx = fmax(mx, amax3(fi));
    }
    block_sums_max_store<7>(acc, mx, part + (size_t)blockIdx.x * kPartW);
}

// MinMDHandler::F: x = x0 + t h, then the skin check of the list update
// (|x - x_build|^2 > skin^2 / 4 anywhere -> rebuild)
__global__ __launch_bounds__(kBlock) void k_min_move(int n, double t, const double4* __restrict__ x0,
                                                     const double4* __restrict__ h, double4* __restrict__ x,
                                                     const double4* __restrict__ xb, double skin_sq4,
                                                     int* __restrict__ flags, Rec16Frame R, uint4* __restrict__ rec)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double4 a = x0[i], d = h[i];
    const double4 xi = make_double4(a.x + t * d.x, a.y + t * d.y, a.z + t * d.z, a.w);
    x[i] = xi;
    // the 16-byte record of the new position (the force passes skip their record pass)
    rec[i] = make_uint4((unsigned)__double2ll_rn(xi.x * R.inv_q), (unsigned)__double2ll_rn(xi.y * R.inv_q),
                        (unsigned)__double2ll_rn(xi.z * R.inv_q), (unsigned)(int)xi.w);
    const double4 b = xb[i];
    if (rsq_exact(b.x, b.y, b.z, xi.x, xi.y, xi.z) > skin_sq4) atomicOr(&flags[1], 1);
}

// after a rebuild slot p holds the atom that was at slot perm[p]; with
// shift set the vector also takes the atom's wrap into the box (new minus
// old position), so x0 + t h stays next to the atom's listed images
__global__ __launch_bounds__(kBlock) void k_min_follow(int n, const int* __restrict__ perm,
                                                       const double4* __restrict__ x_new,
                                                       const double4* __restrict__ x_old, int shift,
                                                       const double4* __restrict__ src, double4* __restrict__ dst)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const int o = perm[p];
    double4 v = src[o];
    if (shift) {
        const double4 xn = x_new[p], xo = x_old[o];
        v.x += xn.x - xo.x;
        v.y += xn.y - xo.y;
        v.z += xn.z - xo.z;
    }
    dst[p] = v;
}

}  // namespace

void MDSystem::min_follow(int nvec, DevBuf<double4>** vecs, int nshift)
{
    const dim3 g((n_ + kBlock - 1) / kBlock);
    for (int k = 0; k < nvec; ++k) {
        // vals_alt_ holds the rebuild's permutation, x_alt_ the positions before it
        hipLaunchKernelGGL(k_min_follow, g, dim3(kBlock), 0, stream_, n_, vals_alt_.get(), x_.get(), x_alt_.get(),
                           k < nshift ? 1 : 0, vecs[k]->get(), mn_tmp_.get());
        std::swap(vecs[k]->p, mn_tmp_.p);
        std::swap(vecs[k]->n, mn_tmp_.n);
    }
}

MinReport MDSystem::minimize(const MinParams& P, std::vector<double>* tally)
{
    MAPP_REQUIRE(P.lbfgs == 0 || P.lbfgs == 1, ERR_ARG, "minimiser kind must be 0 (CG) or 1 (L-BFGS)");
    MAPP_REQUIRE(P.m >= 0 && P.m <= kMaxLbfgsM, ERR_ARG, "L-BFGS memory out of range");
    MAPP_REQUIRE(P.max_nsteps >= 0, ERR_ARG, "max_nsteps must be non-negative");
    MAPP_REQUIRE(P.max_dx > 0.0, ERR_ARG, "max_dx must be positive");
    LineSearchParams lp;
    lp.kind = P.ls;
    lp.c = P.ls_c;
    lp.rho = P.ls_rho;
    lp.min_alpha = P.ls_min_alpha;
    lp.max_iter = P.ls_max_iter;
    lp.brack = P.ls_brack;
    lp.tol = P.ls_tol;
    lp.zeps = P.ls_zeps;
    const char* ls_err = line_search_params_error(lp);
    MAPP_REQUIRE(ls_err == nullptr, ERR_ARG, ls_err ? ls_err : "");
    const bool lbfgs = P.lbfgs == 1;
    const int m = lbfgs ? P.m : 0;
    mn_x0_.reserve(n_);
    mn_h_.reserve(n_);
    mn_f0_.reserve(n_);
    mn_tmp_.reserve(n_);
    while ((int)mn_s_.size() < m) {
        mn_s_.push_back(std::make_unique<DevBuf<double4>>());
        mn_y_.push_back(std::make_unique<DevBuf<double4>>());
    }
    for (int i = 0; i < m; ++i) {
        mn_s_[i]->reserve(n_);
        mn_y_[i]->reserve(n_);
    }
    mn_part_.reserve((size_t)kMinRows * kPartW);
    mn_red_.reserve((size_t)(2 * kMaxLbfgsM + 4) * kPartW);
    const int G = std::min((n_ + kBlock - 1) / kBlock, kMinRows);
    const dim3 gmove((n_ + kBlock - 1) / kBlock);
    double red[kPartW];
    // reduce the partial sums into device row r of mn_red_; fetch copies it to red
    auto rows_dev = [&](int max_col, int r) {
        hipLaunchKernelGGL(k_min_rows, dim3(1), dim3(kBlock), 0, stream_, G, max_col, mn_part_.get(),
                           mn_red_.get() + (size_t)r * kPartW);
    };
    auto fetch = [&](int r) {
        MAPP_HIP(hipMemcpyAsync(red, mn_red_.get() + (size_t)r * kPartW, sizeof red, hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
    };
    auto rows = [&](int max_col) {
        rows_dev(max_col, 0);
        fetch(0);
    };
    // sums of the begin/axpy kernels
    struct Sums {
        double vh, fh, hh, ff, hmax;
    };
    auto sums4 = [&]() {
        rows(4);
        return Sums{red[0], red[1], red[2], red[3], red[4]};
    };
    auto begin = [&](int use_ratio, double ratio, const DevBuf<double4>* v) {
        prof_.begin(K_MD_MIN, stream_);
        hipLaunchKernelGGL(k_min_begin, dim3(G), dim3(kBlock), 0, stream_, n_, use_ratio, ratio, f_.get(),
                           mn_h_.get(), x_.get(), mn_x0_.get(), mn_f0_.get(), v ? v->get() : nullptr,
                           mn_part_.get());
        prof_.end(stream_);
        return sums4();
    };
    // h = (h + a u) g with a from device rows (d_dot >= 0) or the host value;
    // the sums go to device row r
    auto axpy_dev = [&](double a, int d_dot, int d_al, double rho_i, const DevBuf<double4>* u, double g,
                        const DevBuf<double4>* v, int r) {
        const double* red_d = mn_red_.get();
        prof_.begin(K_MD_MIN, stream_);
        hipLaunchKernelGGL(k_min_axpy, dim3(G), dim3(kBlock), 0, stream_, n_, a,
                           d_dot >= 0 ? red_d + (size_t)d_dot * kPartW : nullptr,
                           d_al >= 0 ? red_d + (size_t)d_al * kPartW : nullptr, rho_i, u ? u->get() : nullptr, g,
                           mn_h_.get(), v ? v->get() : nullptr, f_.get(), mn_part_.get());
        prof_.end(stream_);
        rows_dev(4, r);
    };
    auto begin_dev = [&](const DevBuf<double4>* v, int r) {
        prof_.begin(K_MD_MIN, stream_);
        hipLaunchKernelGGL(k_min_begin, dim3(G), dim3(kBlock), 0, stream_, n_, 0, 0.0, f_.get(), mn_h_.get(),
                           x_.get(), mn_x0_.get(), mn_f0_.get(), v ? v->get() : nullptr, mn_part_.get());
        prof_.end(stream_);
        rows_dev(4, r);
    };
    auto pair = [&](double t, DevBuf<double4>* s, DevBuf<double4>* y) {
        prof_.begin(K_MD_MIN, stream_);
        hipLaunchKernelGGL(k_min_pair, dim3(G), dim3(kBlock), 0, stream_, n_, t, mn_h_.get(), mn_f0_.get(), f_.get(),
                           s ? s->get() : nullptr, y ? y->get() : nullptr, mn_part_.get());
        prof_.end(stream_);
        rows(7);
    };
    // every trial point is evaluated with its forces (the EAM force passes
    // give the energy too): the accepted trial's forces are those the
    // reference's force_calc after the line search computes. f_fresh: f_ holds
    // the forces at the current positions, in the current slot order
    bool f_fresh = true;
    MinReport rep{};
    const int reb0 = nrebuild_;
    // MinMDHandler::F (src/min_cg.h:262-269) and F_reset (:273-278)
    // The force passes of a trial point are queued behind the move before the
    // skin check is read back (one host round trip per trial); when the check
    // asks for a rebuild their result is discarded and they run again on the
    // new list. On a stale list they read valid slots only.
    auto F = [&](double t, bool energy) {
        MAPP_HIP(hipMemsetAsync(flags_.get() + 1, 0, sizeof(int), stream_));
        prof_.begin(K_MD_MIN, stream_);
        hipLaunchKernelGGL(k_min_move, gmove, dim3(kBlock), 0, stream_, n_, t, mn_x0_.get(), mn_h_.get(), x_.get(),
                           x0_.get(), 0.25 * skin_ * skin_, flags_.get(), r16_, rec_.get());
        prof_.end(stream_);
        if (energy) force_passes(nullptr, true);
        int flag = 0;
        double e = 0.0;
        MAPP_HIP(hipMemcpyAsync(&flag, flags_.get() + 1, sizeof(int), hipMemcpyDeviceToHost, stream_));
        if (energy) MAPP_HIP(hipMemcpyAsync(&e, red_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
        MAPP_HIP(hipStreamSynchronize(stream_));
        // forces at the new positions come with every evaluated trial point
        f_fresh = energy;
        if (energy) {
            ++rep.evals;
            ++rep.force_calcs;
        }
        if (flag) {
            rebuild();
            DevBuf<double4>* vecs[3 + 2 * kMaxLbfgsM];
            int nv = 0;
            vecs[nv++] = &mn_x0_;
            vecs[nv++] = &mn_h_;
            vecs[nv++] = &mn_f0_;
            for (int i = 0; i < m; ++i) {
                vecs[nv++] = mn_s_[i].get();
                vecs[nv++] = mn_y_[i].get();
            }
            min_follow(nv, vecs, 1);
            if (energy) {
                force_passes(nullptr, true);  // the rebuild wrote the records
                MAPP_HIP(hipMemcpyAsync(&e, red_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
                MAPP_HIP(hipStreamSynchronize(stream_));
                ++rep.force_calcs;
            }
        }
        prof_.flush();
        return e;
    };
    auto force = [&]() {
        const MDReport r = force_calc();
        ++rep.force_calcs;
        f_fresh = true;
        return r.e[0];
    };
    auto f_trial = [&](double t) { return F(t, true); };
    auto f_reset = [&]() { F(0.0, false); };
    LineSearcher<decltype(f_trial), decltype(f_reset)> searcher(lp, f_trial, f_reset);
    // MinMDHandler::ls_prep (src/min_cg.h:231-257): max_dx bounds the largest
    // component of the step; then the chosen line search
    auto line_search = [&](double& nrgy, double& alpha, double f_h, const Sums& S, int init_flag) {
        double dfa = -f_h, max_a;
        if (S.hh == 0.0) {
            max_a = 0.0;
            dfa = 0.0;
        } else if (dfa >= 0.0) {
            max_a = 0.0;
            dfa = 1.0;
        } else {
            max_a = std::fabs(P.max_dx / S.hmax);
        }
        return searcher.min(nrgy, alpha, dfa, max_a, init_flag);
    };

    // NewDynamicMD::init (src/dynamic_md.h:145-163): a fresh list at the start
    rebuild();
    const double e0 = force();
    rep.e0 = e0;
    if (tally) {
        tally->push_back(0.0);
        tally->push_back(e0);
    }
    double e_curr = e0, e_prev = e0;
    int err = P.max_nsteps == 0 ? MIN_F_MAX_ITER : LS_S;
    int istep = 0;
    if (lbfgs) {
        std::vector<double> rho(m, 0.0);
        std::vector<int> dot_row(m, 0);
        std::vector<int> ord(m);  // ord[i]: buffer holding the i-th most recent pair
        for (int i = 0; i < m; ++i) ord[i] = i;
        int k = 0;
        double gamma = 1.0;
        for (; err == LS_S; ++istep) {
            // two-loop recursion h = H f over the k stored pairs, queued with
            // its coefficients computed on the device; dot_row[i] holds s_i.h
            int cur = 0;
            begin_dev(k > 0 ? mn_s_[ord[0]].get() : nullptr, cur);
            for (int i = 0; i < k; ++i) {
                dot_row[i] = cur;
                const bool last = i + 1 == k;
                axpy_dev(0.0, cur, -1, rho[i], mn_y_[ord[i]].get(), last ? gamma : 1.0,
                         last ? mn_y_[ord[k - 1]].get() : mn_s_[ord[i + 1]].get(), cur + 1);
                ++cur;
            }
            if (k == 0 && gamma != 1.0) {
                axpy_dev(0.0, -1, -1, 0.0, nullptr, gamma, nullptr, cur + 1);
                ++cur;
            }
            for (int i = k - 1; i > -1; --i) {
                axpy_dev(0.0, cur, dot_row[i], rho[i], mn_s_[ord[i]].get(), 1.0,
                         i > 0 ? mn_y_[ord[i - 1]].get() : nullptr, cur + 1);
                ++cur;
            }
            fetch(cur);
            Sums S{red[0], red[1], red[2], red[3], red[4]};
            e_prev = e_curr;
            double f_h = S.fh;
            if (f_h < 0.0) {
                S = begin(0, 0.0, nullptr);
                k = 0;
                f_h = S.ff;
            }
            double alpha_m = 0.0;
            err = line_search(e_curr, alpha_m, f_h, S, 0);
            if (err != LS_S) {
                --istep;
                continue;
            }
            if (tally) {
                tally->push_back((double)(istep + 1));
                tally->push_back(e_curr);
            }
            if (e_prev - e_curr < P.e_tol) err = MIN_S_TOLERANCE;
            if (istep + 1 == P.max_nsteps) err = MIN_F_MAX_ITER;
            if (err) continue;
            if (m) {
                if (k != m) ++k;
                // cyclic_shift of the first k pairs: the oldest buffer takes the new pair
                std::rotate(ord.begin(), ord.begin() + (k - 1), ord.begin() + k);
                for (int i = m - 1; i > 0; --i) rho[i] = rho[i - 1];
                pair(alpha_m, mn_s_[ord[0]].get(), mn_y_[ord[0]].get());
                const double inner0 = red[0], inner1 = red[1];
                gamma = inner0 / inner1;
                rho[0] = 1.0 / inner0;
            } else {
                pair(alpha_m, nullptr, nullptr);
                // alpha (h.f0 - h.f) / (f.f + f0.f0 - 2 f.f0)
                gamma = alpha_m * (red[2] - red[3]) / (red[4] + red[5] - 2.0 * red[6]);
            }
        }
    } else {
        Sums S = begin(0, 0.0, nullptr);
        double f0_f0 = S.ff, ratio = 0.0;
        bool first = true;
        for (; err == LS_S; ++istep) {
            if (f0_f0 == 0.0) {
                err = LS_F_GRAD0;
                continue;
            }
            if (!first) S = begin(1, ratio, nullptr);
            first = false;
            e_prev = e_curr;
            double f_h = S.fh;
            if (f_h < 0.0) {
                S = begin(0, 0.0, nullptr);
                f_h = f0_f0;
            }
            double alpha = 0.0;
            err = line_search(e_curr, alpha, f_h, S, 1);
            if (err != LS_S) {
                --istep;
                continue;
            }
            if (tally) {
                tally->push_back((double)(istep + 1));
                tally->push_back(e_curr);
            }
            if (e_prev - e_curr < P.e_tol) err = MIN_S_TOLERANCE;
            if (istep + 1 == P.max_nsteps) err = MIN_F_MAX_ITER;
            if (err) continue;
            pair(0.0, nullptr, nullptr);
            const double f_f = red[4], f_f0 = red[6];
            ratio = (f_f - f_f0) / f0_f0;
            f0_f0 = f_f;
        }
    }
    // a rebuild inside a failed line search reordered the slots after the
    // last force calculation: the forces at the same positions again
    if (!f_fresh) force();
    pair(0.0, nullptr, nullptr);
    rep.f_norm = std::sqrt(red[4]);
    rep.f_max = red[7];
    rep.steps = istep;
    rep.status = err;
    rep.e = e_curr;
    rep.rebuilds = nrebuild_ - reb0;
    prof_.flush();
    return rep;
}

}  // namespace mapp
