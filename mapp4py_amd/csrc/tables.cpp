// tables.cpp -- DYNAMO setfl / fs readers and the spline coefficient builder.
//
// Behaviour follows the reference importer (src/import_eam.h:52-64, 348-901;
// src/import_eam.cpp:62-115, 150-182): tokens are read across lines, '#'
// starts a comment, the per-element header line is skipped, the cutoff in
// the grid line is ignored and r_c is derived from the last non-zero knot.
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "common.hpp"

namespace mapp {

void spline_coefficients(int n, double* s, int stride)
{
    MAPP_REQUIRE(n >= 5, ERR_ARG, "spline needs at least 5 knots");
    auto y = [&](int k) { return s[(size_t)k * stride]; };
    auto c = [&](int k, int col) -> double& { return s[(size_t)k * stride + col]; };
    // slopes (per unit p): one-sided at both ends, 3-point next to them,
    // 5-point (exact for quartics) inside
    c(0, 1) = y(1) - y(0);
    c(1, 1) = 0.5 * (y(2) - y(0));
    c(n - 2, 1) = 0.5 * (y(n - 1) - y(n - 3));
    c(n - 1, 1) = y(n - 1) - y(n - 2);
    for (int k = 2; k < n - 2; ++k) c(k, 1) = ((y(k - 2) - y(k + 2)) + 8.0 * (y(k + 1) - y(k - 1))) / 12.0;
    // Hermite cubic on each interval from the end values and slopes
    for (int k = 0; k < n - 1; ++k) {
        const double dy = y(k + 1) - y(k);
        c(k, 2) = 3.0 * dy - 2.0 * c(k, 1) - c(k + 1, 1);
        c(k, 3) = c(k, 1) + c(k + 1, 1) - 2.0 * dy;
    }
    c(n - 1, 2) = 0.0;
    c(n - 1, 3) = 0.0;
    if (stride == 5) {
        for (int k = 0; k < n - 1; ++k) c(k, 4) = (c(k + 1, 2) - c(k, 2)) / 6.0 - 0.5 * c(k, 3);
        c(n - 1, 4) = 0.0;
    }
}

bool hermite_knots(const double* c, int nr, std::vector<double>& ys)
{
    ys.assign((size_t)2 * nr, 0.0);
    for (int m = 0; m < nr; ++m) {
        ys[2 * (size_t)m] = c[4 * (size_t)m];
        ys[2 * (size_t)m + 1] = c[4 * (size_t)m + 1];
    }
    for (int m = 0; m + 1 < nr; ++m) {
        const double y0 = ys[2 * (size_t)m], s0 = ys[2 * (size_t)m + 1];
        const double y1 = ys[2 * (size_t)m + 2], s1 = ys[2 * (size_t)m + 3];
        const double dy = y1 - y0;
        const double c2 = 3.0 * dy - 2.0 * s0 - s1, c3 = s0 + s1 - 2.0 * dy;
        const double scale = std::max({std::fabs(y0), std::fabs(y1), std::fabs(s0), std::fabs(s1), 1e-300});
        if (std::fabs(c[4 * (size_t)m + 2] - c2) > 1e-12 * scale || std::fabs(c[4 * (size_t)m + 3] - c3) > 1e-12 * scale)
            return false;
    }
    return true;
}

namespace {

EamTables alloc_tables(int ne, int nr, int nrho, double dr, double drho, bool dmd)
{
    EamTables t;
    t.ne = ne;
    t.nr = nr;
    t.nrho = nrho;
    t.dr = dr;
    t.drho = drho;
    t.dmd = dmd;
    t.rphi.assign((size_t)ne * ne * nr * 4, 0.0);
    t.rho.assign((size_t)ne * ne * nr * 4, 0.0);
    t.F.assign((size_t)ne * nrho * t.nf(), 0.0);
    t.rc.assign((size_t)ne * ne, 0.0);
    return t;
}

double* mrho(EamTables& t, int a, int b) { return t.rho.data() + ((size_t)a * t.ne + b) * t.nr * 4; }
double* mphi(EamTables& t, int a, int b) { return t.rphi.data() + ((size_t)a * t.ne + b) * t.nr * 4; }

// splines, r_c from the knots, and for the DMD force field the densities are
// turned into r*rho(r) before their final spline (src/ff_eam_dmd.cpp:80-94)
void finalize(EamTables& t)
{
    const int ne = t.ne;
    for (int a = 0; a < ne; ++a) {
        spline_coefficients(t.nrho, t.F.data() + (size_t)a * t.nrho * t.nf(), t.nf());
        for (int b = 0; b < ne; ++b) {
            spline_coefficients(t.nr, mrho(t, a, b), 4);
            spline_coefficients(t.nr, mphi(t, a, b), 4);
        }
    }
    for (int a = 0; a < ne; ++a)
        for (int b = 0; b <= a; ++b) {
            int k = t.nr - 1;
            while (k >= 1 && mphi(t, a, b)[(size_t)k * 4] == 0.0 && mrho(t, a, b)[(size_t)k * 4] == 0.0 &&
                   mrho(t, b, a)[(size_t)k * 4] == 0.0)
                --k;
            t.rc[(size_t)a * ne + b] = t.rc[(size_t)b * ne + a] = (double)k * t.dr;
        }
    if (t.dmd) {
        for (int a = 0; a < ne; ++a)
            for (int b = 0; b < ne; ++b) {
                double* s = mrho(t, a, b);
                for (int k = 0; k < t.nr; ++k) s[(size_t)k * 4] *= (double)k * t.dr;
                spline_coefficients(t.nr, s, 4);
            }
    }
}

struct LineReader {
    std::ifstream in;
    std::string path;
    int line_no = 0;
    explicit LineReader(const std::string& p) : in(p), path(p) {
        MAPP_REQUIRE(in.good(), ERR_FILE, "cannot open potential file " + p);
    }
    bool next(std::string& line) {
        if (!std::getline(in, line)) return false;
        ++line_no;
        return true;
    }
    void need(std::string& line) {
        MAPP_REQUIRE(next(line), ERR_FILE, path + " file ended immaturely");
    }
    static std::vector<std::string> split(const std::string& line) {
        std::string s = line.substr(0, line.find('#'));
        std::istringstream is(s);
        std::vector<std::string> out;
        std::string tok;
        while (is >> tok) out.push_back(tok);
        return out;
    }
    // n numbers spread over as many lines as needed
    void numbers(std::vector<double>& out, size_t n) {
        out.resize(n);
        size_t i = 0;
        std::string line;
        while (i < n) {
            need(line);
            auto tok = split(line);
            MAPP_REQUIRE(tok.size() + i <= n, ERR_FILE, path + " file has extra arguments at line " + std::to_string(line_no));
            for (auto& s : tok) out[i++] = std::atof(s.c_str());
        }
    }
};

}  // namespace

EamTables read_dynamo(const std::string& path, int fmt, const std::vector<std::string>& elems, bool dmd)
{
    MAPP_REQUIRE(fmt == 0 || fmt == 1, ERR_ARG, "unknown EAM file format");
    LineReader rd(path);
    std::string line;
    for (int k = 0; k < 3; ++k) rd.need(line);
    rd.need(line);
    auto tok = LineReader::split(line);
    MAPP_REQUIRE(tok.size() >= 2, ERR_FILE, "invalid element line in " + path + ": " + line);
    const int nfile = (int)tok.size() - 1;
    std::vector<std::string> file_names(tok.begin() + 1, tok.end());
    std::vector<int> map(nfile, -1);  // file element -> requested slot
    std::vector<std::string> names = elems.empty() ? file_names : elems;
    const int ne = (int)names.size();
    for (int f = 0; f < nfile; ++f)
        for (int e = 0; e < ne; ++e)
            if (names[e] == file_names[f] && map[f] < 0) {
                bool taken = false;
                for (int g = 0; g < f; ++g) taken |= map[g] == e;
                if (!taken) map[f] = e;
            }
    for (int e = 0; e < ne; ++e) {
        bool found = false;
        for (int f = 0; f < nfile; ++f) found |= map[f] == e;
        MAPP_REQUIRE(found, ERR_FILE, path + " file does not contain parameters for element " + names[e]);
    }
    rd.need(line);
    size_t nrho = 0, nr = 0;
    double drho = 0.0, dr = 0.0, rc_hdr = 0.0;
    {
        std::istringstream is(line);
        MAPP_REQUIRE(bool(is >> nrho >> drho >> nr >> dr >> rc_hdr), ERR_FILE, "invalid grid line in " + path + ": " + line);
    }
    MAPP_REQUIRE(nrho >= 5 && nr >= 5, ERR_FILE, "nr and nrho in " + path + " should be larger than 5");
    MAPP_REQUIRE(drho > 0.0 && dr > 0.0, ERR_FILE, "dr and drho in " + path + " should be larger than 0.0");
    EamTables t = alloc_tables(ne, (int)nr, (int)nrho, dr, drho, dmd);
    t.names = names;
    std::vector<double> buf;
    const size_t nblock = fmt == 1 ? nrho + nr * (size_t)nfile : nrho + nr;
    for (int f = 0; f < nfile; ++f) {
        rd.need(line);  // atomic number, mass, lattice constant, lattice
        rd.numbers(buf, nblock);
        const int a = map[f];
        if (a < 0) continue;
        for (size_t k = 0; k < nrho; ++k) t.F[((size_t)a * nrho + k) * t.nf()] = buf[k];
        for (int g = 0; g < nfile; ++g) {
            const int b = fmt == 1 ? map[g] : -1;
            if (fmt == 1 && b < 0) continue;
            const double* src = buf.data() + nrho + (fmt == 1 ? (size_t)g * nr : 0);
            if (fmt == 1) {
                for (size_t k = 0; k < nr; ++k) mrho(t, a, b)[k * 4] = src[k];
            } else {
                for (int bb = 0; bb < ne; ++bb)
                    for (size_t k = 0; k < nr; ++k) mrho(t, a, bb)[k * 4] = src[k];
                break;
            }
        }
    }
    rd.numbers(buf, (size_t)nfile * (nfile + 1) / 2 * nr);
    size_t pos = 0;
    for (int f = 0; f < nfile; ++f)
        for (int g = 0; g <= f; ++g) {
            const int a = map[f], b = map[g];
            if (a >= 0 && b >= 0)
                for (size_t k = 0; k < nr; ++k) mphi(t, a, b)[k * 4] = mphi(t, b, a)[k * 4] = buf[pos + k];
            pos += nr;
        }
    finalize(t);
    return t;
}

namespace {

// Value at offset p of interval k of the cubic through the knots with
// 5-point slopes (one-sided near the ends, linear on the last interval):
// ImportEAM::interpolate(arr, n, p, k), src/import_eam.cpp:9-52
double knot_cubic(const std::vector<double>& a, size_t n, double p, size_t k)
{
    const double y0 = a[k];
    double d1, d2, d3;
    if (k + 1 == n) return y0 + (a[n - 1] - a[n - 2]) * p;
    if (k == 0) {
        d1 = a[1] - a[0];
        d2 = (-a[0] + 2.0 * a[1] - a[2]) / 2.0;
        d3 = (a[0] - 2.0 * a[1] + a[2]) / 2.0;
    } else if (k == 1) {
        d1 = 0.5 * (a[2] - a[0]);
        d2 = (11.0 * a[0] - 28.0 * a[1] + 24.0 * a[2] - 8.0 * a[3] + a[4]) / 12.0;
        d3 = (-5.0 * a[0] + 16.0 * a[1] - 18.0 * a[2] + 8.0 * a[3] - a[4]) / 12.0;
    } else if (k + 2 == n) {
        d1 = 0.5 * (a[n - 1] - a[n - 3]);
        d2 = a[n - 3] - 3.0 * a[n - 2] + 2.0 * a[n - 1];
        d3 = 0.5 * (-a[n - 3] + 3.0 * a[n - 2] - 2.0 * a[n - 1]);
    } else {
        d1 = ((a[k - 2] - a[k + 2]) + 8.0 * (a[k + 1] - a[k - 1])) / 12.0;
        d2 = (-2.0 * a[k - 2] + 15.0 * a[k - 1] - 28.0 * a[k] + 21.0 * a[k + 1] - 6.0 * a[k + 2]) / 12.0;
        d3 = (a[k - 2] - 7.0 * a[k - 1] + 16.0 * a[k] - 17.0 * a[k + 1] + 7.0 * a[k + 2]) / 12.0;
    }
    return ((d3 * p + d2) * p + d1) * p + y0;
}

// interval and offset of x on a grid of n knots spaced d (the last interval
// takes everything beyond it, the offset is clamped to 1)
size_t knot_at(double x, double d, size_t n, double& p)
{
    double q = x / d;
    size_t l = (size_t)q;
    if (l > n - 2) l = n - 2;
    q -= (double)l;
    p = q < 1.0 ? q : 1.0;
    return l;
}

}  // namespace

EamTables read_funcfl(const std::vector<std::string>& paths, bool dmd)
{
    const int ne = (int)paths.size();
    MAPP_REQUIRE(ne >= 1 && ne <= 8, ERR_ARG, "1 to 8 funcfl files (one per element)");
    struct One {
        size_t nrho = 0, nr = 0;
        double drho = 0.0, dr = 0.0;
        std::vector<double> F, Z, rho;
    };
    std::vector<One> f(ne);
    for (int i = 0; i < ne; ++i) {
        LineReader rd(paths[i]);
        std::string line;
        rd.need(line);  // comment
        rd.need(line);  // atomic number, mass, lattice constant, lattice
        rd.need(line);
        double rc_hdr = 0.0;
        std::istringstream is(line);
        MAPP_REQUIRE(bool(is >> f[i].nrho >> f[i].drho >> f[i].nr >> f[i].dr >> rc_hdr), ERR_FILE,
                     "invalid line in " + paths[i] + ": " + line);
        MAPP_REQUIRE(f[i].nrho >= 5 && f[i].nr >= 5, ERR_FILE, "nr and nrho in " + paths[i] + " should be larger than 5");
        MAPP_REQUIRE(f[i].drho > 0.0 && f[i].dr > 0.0, ERR_FILE,
                     "dr and drho in " + paths[i] + " should be larger than 0.0");
        std::vector<double> v;
        rd.numbers(v, f[i].nrho + 2 * f[i].nr);
        f[i].F.assign(v.begin(), v.begin() + f[i].nrho);
        f[i].Z.assign(v.begin() + f[i].nrho, v.begin() + f[i].nrho + f[i].nr);
        f[i].rho.assign(v.begin() + f[i].nrho + f[i].nr, v.end());
    }
    // common grid: the largest spacing over the longest range (src/import_eam.h:214-231)
    double rmax = 0.0, rhomax = 0.0, dr = 0.0, drho = 0.0;
    for (const One& o : f) {
        rmax = std::max(rmax, (double)(o.nr - 1) * o.dr);
        rhomax = std::max(rhomax, (double)(o.nrho - 1) * o.drho);
        dr = std::max(dr, o.dr);
        drho = std::max(drho, o.drho);
    }
    const size_t nr = (size_t)(rmax / dr + 0.5), nrho = (size_t)(rhomax / drho + 0.5);
    for (int i = 0; i < ne; ++i)  // the reference would read past this file's Z table
        MAPP_REQUIRE(nr <= f[i].nr, ERR_FILE, "funcfl: " + paths[i] + " has fewer r knots than the common grid");
    std::vector<double> F((size_t)ne * nrho), rho((size_t)ne * ne * nr, 0.0), rphi((size_t)ne * ne * nr);
    double p;
    // r phi_ij = 27.2 * 0.529 Z_i Z_j (Hartree * Bohr to eV * A); Z is read on
    // interval k of the common grid with the offset of the file's own
    // interval, as src/import_eam.h:286-305 does
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j <= i; ++j)
            for (size_t k = 0; k < nr; ++k) {
                const double r = (double)k * dr;
                knot_at(r, f[i].dr, f[i].nr, p);
                const double zi = knot_cubic(f[i].Z, f[i].nr, p, k);
                knot_at(r, f[j].dr, f[j].nr, p);
                const double zj = knot_cubic(f[j].Z, f[j].nr, p, k);
                rphi[((size_t)i * ne + j) * nr + k] = rphi[((size_t)j * ne + i) * nr + k] = 27.2 * 0.529 * zi * zj;
            }
    for (int i = 0; i < ne; ++i)
        for (size_t k = 0; k < nr; ++k) {
            const size_t l = knot_at((double)k * dr, f[i].dr, f[i].nr, p);
            rho[(size_t)i * ne * nr + k] = knot_cubic(f[i].rho, f[i].nr, p, l);
        }
    for (int i = 0; i < ne; ++i)
        for (size_t k = 0; k < nrho; ++k) {
            const size_t l = knot_at((double)k * drho, f[i].drho, f[i].nrho, p);
            F[(size_t)i * nrho + k] = knot_cubic(f[i].F, f[i].nrho, p, l);
        }
    return tables_from_knots(ne, (int)nr, (int)nrho, dr, drho, F.data(), rho.data(), rphi.data(), true, dmd);
}

EamTables tables_from_knots(int ne, int nr, int nrho, double dr, double drho, const double* F,
                            const double* rho, const double* r_phi, bool setfl_alias, bool dmd)
{
    MAPP_REQUIRE(ne >= 1 && ne <= 8, ERR_ARG, "1 to 8 elements are supported");
    MAPP_REQUIRE(nr >= 5 && nrho >= 5 && dr > 0.0 && drho > 0.0, ERR_ARG, "invalid table grid");
    EamTables t = alloc_tables(ne, nr, nrho, dr, drho, dmd);
    for (int a = 0; a < ne; ++a) {
        t.names.push_back("E" + std::to_string(a));
        for (int k = 0; k < nrho; ++k) t.F[((size_t)a * nrho + k) * t.nf()] = F[(size_t)a * nrho + k];
        for (int b = 0; b < ne; ++b) {
            const int bb = setfl_alias ? 0 : b;
            for (int k = 0; k < nr; ++k) {
                mrho(t, a, b)[(size_t)k * 4] = rho[((size_t)a * ne + bb) * nr + k];
                mphi(t, a, b)[(size_t)k * 4] = r_phi[((size_t)a * ne + b) * nr + k];
            }
        }
    }
    finalize(t);
    return t;
}

}  // namespace mapp
