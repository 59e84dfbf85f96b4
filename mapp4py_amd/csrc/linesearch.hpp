// linesearch.hpp -- the reference's line searches over any energy along a
// line: LineSearchBackTrack::min (src/ls.h:711-769), LineSearch::bracket
// (src/ls.h:213-335), LineSearchGoldenSection::min (src/ls.h:341-520) and
// LineSearchBrent::min (src/ls.h:525-706), with the defaults of
// src/ls.cpp:95-101,262-268,456-461. Host code: f(alpha) sets the state to
// x0 + alpha h and returns its energy, reset() sets it back to x0.
#pragma once
#include <algorithm>
#include <cmath>

namespace mapp {

// status codes in the order of the reference's enum (src/ls.h:9-22)
enum { LS_S = 0, LS_F_DOWNHILL, LS_F_GRAD0, LS_MIN_ALPHA, MIN_S_TOLERANCE, MIN_F_MAX_ITER, B_S, B_F_MAX_ALPHA,
       B_F_DOWNHILL };

struct LineSearchParams {
    int kind = 0;  // 0 back-track, 1 golden section, 2 Brent
    double c = 0.4, rho = 0.5, min_alpha = 0.0;
    int max_iter = 5, brack = 1;
    double tol = 2.1073424255447017e-08, zeps = 2.220446049250313e-19;
};

// The parameter ranges under which every search ends: 0 < c < 1 and
// 0 < rho < 1 (the back-tracking step shrinks), min_alpha >= 0, max_iter,
// tol and zeps >= 0, all finite. Returns nullptr when p is valid, else the
// message of the first violated condition.
inline const char* line_search_params_error(const LineSearchParams& p)
{
    if (p.kind < 0 || p.kind > 2) return "line search must be 0 (back-track), 1 (golden) or 2 (Brent)";
    if (!(p.c > 0.0 && p.c < 1.0)) return "line search c must lie in (0, 1)";
    if (!(p.rho > 0.0 && p.rho < 1.0)) return "line search rho must lie in (0, 1)";
    if (!(p.min_alpha >= 0.0 && std::isfinite(p.min_alpha))) return "line search min_alpha must be finite and >= 0";
    if (p.max_iter < 0) return "line search max_iter must be non-negative";
    if (!(p.tol >= 0.0 && std::isfinite(p.tol)) || !(p.zeps >= 0.0 && std::isfinite(p.zeps)))
        return "line search tol and zeps must be finite and >= 0";
    return nullptr;
}

template <class Fn, class Reset>
class LineSearcher {
public:
    LineSearcher(const LineSearchParams& p, Fn f, Reset reset) : p_(p), f_(f), reset_(reset) {}
    // the back-tracking search's memory (LineSearch::prev_val; reset() sets -1)
    double prev_val = -1.0;

    // LineSearch*::min after MinMDHandler::ls_prep: dfa the slope along h
    // (-f.h), max_a the largest step; init_flag as the minimisers pass it
    int min(double& nrgy, double& alpha, double dfa, double max_a, int init_flag)
    {
        if (dfa == 0.0) return LS_F_GRAD0;
        if (max_a == 0.0) return LS_MIN_ALPHA;
        if (dfa >= 0.0) return LS_F_DOWNHILL;
        if (p_.kind == 1) return golden(nrgy, alpha, dfa, max_a);
        if (p_.kind == 2) return brent(nrgy, alpha, dfa, max_a);
        return backtrack(nrgy, alpha, dfa, max_a, init_flag);
    }

private:
    static constexpr double kEps = 2.220446049250313e-16;       // numeric_limits<double>::epsilon()
    static constexpr double kSqrtEps = 1.4901161193847656e-08;  // its square root
    static constexpr double kGolden = 1.618033988749895;        // (1 + sqrt 5) / 2

    // LineSearchBackTrack::min (src/ls.h:711-769)
    int backtrack(double& nrgy, double& alpha, double dfa, double max_a, int init_flag)
    {
        const double init_energy = nrgy;
        if (init_flag == 0) {
            max_a = std::min(max_a, 1.0);
        } else if (init_flag == 1) {
            if (prev_val > 0.0 && dfa < 0.0) max_a = std::min(max_a, -prev_val / dfa);
            prev_val = -dfa;
        } else if (init_flag == 2) {
            if (prev_val > 0.0 && dfa < 0.0) max_a = std::min(max_a, std::min(1.0, 2.02 * (prev_val + nrgy) / dfa));
            prev_val = -nrgy;
        }
        if (max_a <= p_.min_alpha) return (int)LS_MIN_ALPHA;
        alpha = max_a;
        while (true) {
            const double ideal = nrgy + alpha * p_.c * dfa;
            const double cur = f_(alpha);
            if (cur <= ideal) {
                if (init_flag == 1) prev_val *= alpha;
                nrgy = cur;
                return (int)LS_S;
            }
            alpha *= p_.rho;
            // a nan step (a 0/0 curvature pair once the forces stop changing)
            // never meets the reference's exit test; it ends the search here
            if (alpha <= p_.min_alpha || std::isnan(alpha)) {
                nrgy = init_energy;
                reset_();
                return (int)LS_MIN_ALPHA;
            }
        }
    }
    // LineSearch::bracket (src/ls.h:213-335): a < b < c with f(b) below f(a)
    // and f(c), or the shortest step that lowers the energy
    int bracket(double dfa, double max_a, double& a, double& b, double& c, double& fa, double& fb, double& fc)
    {
        const double q0 = 1.0e-14 / (max_a * std::fabs(dfa));
        if (0.01 < q0 && q0 < 1.0)
            b = q0 * max_a;
        else if (q0 >= 1.0)
            b = max_a;
        else
            b = 0.01 * max_a;
        fb = f_(b);
        if (fb >= fa) {
            c = b;
            fc = fb;
            int iter = 20;
            double r = (fa - fc) / (c * dfa);
            b = c * 0.5 / (1.0 + r);
            while (fb >= fa && iter && b > kEps) {
                fb = f_(b);
                if (fb < fa) continue;
                c = b;
                fc = fb;
                --iter;
                r = (fa - fc) / (c * dfa);
                b = c * 0.5 / (1.0 + r);
            }
            if (fb >= fa) {
                b = kSqrtEps < max_a ? kSqrtEps : max_a;
                fb = f_(b);
                if (fb >= fa) return (int)B_F_DOWNHILL;
            }
            return (int)B_S;
        }
        fc = fb;
        while (fb >= fc) {
            c = b + kGolden * (b - a);
            if (c >= max_a) {
                c = max_a;
                fc = f_(c);
                return (int)B_S;
            }
            fc = f_(c);
            if (fc > fb) continue;
            const double ulim = max_a;
            const double r = (b - a) * (fb - fc), q = (b - c) * (fb - fa);
            double u = b - ((b - c) * q - (b - a) * r) / (2.0 * (q - r));
            if (b < u && u < c) {
                const double fu = f_(u);
                if (fu < fc) {
                    a = b;
                    b = u;
                    fa = fb;
                    fb = fu;
                    return (int)B_S;
                } else if (fu > fb) {
                    c = u;
                    fc = fu;
                    return (int)B_S;
                }
                a = b;
                b = c;
                fa = fb;
                fb = fc;
            } else if (u > c) {
                u = u < ulim ? u : ulim;
                const double fu = f_(u);
                a = b;
                b = c;
                c = u;
                fa = fb;
                fb = fc;
                fc = fu;
            } else {
                a = b;
                b = c;
                fa = fb;
                fb = fc;
            }
        }
        return (int)B_S;
    }
    // LineSearchGoldenSection::min (src/ls.h:341-520)
    int golden(double& nrgy, double& alpha, double dfa, double max_a)
    {
        constexpr double gold = 0.61803399, cgold = 0.38196601;
        int calc = 0;
        double a = 0.0, b = 0.0, c = 0.0, fa = nrgy, fb = 0.0, fc = 0.0;
        double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0, f1 = 0.0, f2 = 0.0;
        if (p_.brack) {
            const int chk = bracket(dfa, max_a, a, b, c, fa, fb, fc);
            if (chk != B_S) {
                reset_();
                return chk;
            }
            x0 = a;
            x3 = c;
            if (c - b > b - a) {
                x1 = b;
                x2 = b + cgold * (c - b);
                f1 = fb;
                f2 = f_(x2);
                calc = 2;
            } else {
                x1 = b + cgold * (a - b);
                x2 = b;
                f1 = f_(x1);
                f2 = fb;
                calc = 1;
            }
            int iter = p_.max_iter;
            while (x3 - x0 > p_.tol * (x1 + x2) && x3 - x0 > kEps && iter) {
                if (f2 < f1) {
                    x0 = x1;
                    x1 = x2;
                    x2 = gold * x2 + cgold * x3;
                    f1 = f2;
                    f2 = f_(x2);
                    calc = 2;
                } else {
                    x3 = x2;
                    x2 = x1;
                    x1 = gold * x1 + cgold * x0;
                    f2 = f1;
                    f1 = f_(x1);
                    calc = 1;
                }
                --iter;
            }
        } else {
            double f0 = fa;
            x3 = max_a;
            double f3 = f_(x3);
            int iter = p_.max_iter;
            bool set_left = false, set_right = false;
            while (x3 - x0 > p_.tol * (x1 + x2) && x3 - x0 > kEps && iter) {
                const double delta = cgold * (x3 - x0);
                if (!set_left) {
                    x1 = x0 + delta;
                    f1 = f_(x1);
                    calc = 1;
                    set_left = true;
                }
                if (!set_right) {
                    x2 = x3 - delta;
                    f2 = f_(x2);
                    calc = 2;
                    set_right = true;
                }
                if ((f0 < f2 && f0 < f3) || (f1 < f2 && f1 < f3)) {
                    set_left = false;
                    x3 = x2;
                    x2 = x1;
                    f3 = f2;
                    f2 = f1;
                } else {
                    set_right = false;
                    x0 = x1;
                    x1 = x2;
                    f0 = f1;
                    f1 = f2;
                }
                --iter;
            }
            if (!set_left && f0 < f1) {
                if (x0 == 0.0) {
                    reset_();
                    alpha = 0.0;
                    return (int)LS_MIN_ALPHA;
                }
                x1 = x0;
                f1 = f_(x1);
                calc = 1;
            }
            if (!set_right && f3 < f2) {
                x2 = x3;
                f2 = f_(x2);
                calc = 2;
            }
        }
        if (f1 < f2) {
            if (calc == 2) f1 = f_(x1);
            nrgy = f1;
            alpha = x1;
        } else {
            if (calc == 1) f2 = f_(x2);
            nrgy = f2;
            alpha = x2;
        }
        prev_val = -dfa * alpha;
        return (int)LS_S;
    }
    // LineSearchBrent::min (src/ls.h:525-706)
    int brent(double& nrgy, double& alpha, double dfa, double max_a)
    {
        constexpr double cgold = 0.38196601;
        double a = 0.0, b = 0.0, fa = nrgy, fb = 0.0, x = 0.0, fx = 0.0;
        double d = 0.0, e = 0.0, u = 0.0, fu = 0.0;
        if (p_.brack) {
            const int chk = bracket(dfa, max_a, a, x, b, fa, fx, fb);
            if (chk != B_S) {
                reset_();
                return chk;
            }
        } else {
            b = max_a;
            x = cgold * (b - a);
            fx = f_(x);
        }
        double w = x, v = x, fw = fx, fv = fx;
        for (int iter = 0; iter < p_.max_iter; ++iter) {
            const double xm = 0.5 * (a + b);
            const double tol1 = p_.tol * std::fabs(x) + p_.zeps, tol2 = 2.0 * tol1;
            if (std::fabs(x - xm) <= (tol2 - 0.5 * (b - a))) {
                if (u != x) f_(x);
                nrgy = fx;
                alpha = x;
                prev_val = alpha;
                return (int)LS_S;
            }
            if (std::fabs(e) > tol1) {
                const double r = (x - w) * (fx - fv);
                double q = (x - v) * (fx - fw);
                double p = (x - v) * q - (x - w) * r;
                q = 2.0 * (q - r);
                if (q > 0.0) p = -p;
                q = std::fabs(q);
                const double etemp = e;
                e = d;
                if (std::fabs(p) >= std::fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
                    e = x >= xm ? a - x : b - x;
                    d = cgold * e;
                } else {
                    d = p / q;
                    u = x + d;
                    if (u - a < tol2 || b - u < tol2) d = xm - x >= 0.0 ? std::fabs(tol1) : -std::fabs(tol1);
                }
            } else {
                e = x >= xm ? a - x : b - x;
                d = cgold * e;
            }
            if (std::fabs(d) >= tol1)
                u = x + d;
            else
                u = d >= 0.0 ? x + std::fabs(tol1) : x - std::fabs(tol1);
            fu = f_(u);
            if (fu <= fx) {
                if (u >= x)
                    a = x;
                else
                    b = x;
                v = w;
                w = x;
                x = u;
                fv = fw;
                fw = fx;
                fx = fu;
            } else {
                if (u < x)
                    a = u;
                else
                    b = u;
                if (fu <= fw || w == x) {
                    v = w;
                    w = u;
                    fv = fw;
                    fw = fu;
                } else if (fu <= fv || v == x || v == w) {
                    v = u;
                    fv = fu;
                }
            }
        }
        if (fa < fx) {
            reset_();
            alpha = 0.0;
            return (int)LS_MIN_ALPHA;
        }
        if (u != x) f_(x);
        nrgy = fx;
        alpha = x;
        prev_val = alpha;
        return (int)LS_S;
    }

    LineSearchParams p_;
    Fn f_;
    Reset reset_;
};

}  // namespace mapp
