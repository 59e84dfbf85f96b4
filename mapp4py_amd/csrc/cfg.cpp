// cfg.cpp -- extended CFG files of DMD systems (host code).
#include "cfg.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>

#include "common.hpp"
#include "md.hpp"

namespace mapp {
namespace {

// the part of a line before '#', as whitespace-separated tokens
std::vector<std::string> tokens(const std::string& raw)
{
    std::vector<std::string> t;
    const size_t end = raw.find('#');
    const std::string s = raw.substr(0, end);
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        if (i >= s.size()) break;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) ++j;
        t.push_back(s.substr(i, j - i));
        i = j;
    }
    return t;
}

std::string join(const std::vector<std::string>& t)
{
    std::string s;
    for (size_t i = 0; i < t.size(); ++i) {
        if (i) s += ' ';
        s += t[i];
    }
    return s;
}

// square root of a symmetric positive-definite 3x3 matrix (cyclic Jacobi
// eigen-decomposition); false if an eigenvalue is not positive
bool sym_sqrt3(const double (&A)[3][3], double (&R)[3][3])
{
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::memcpy(a, A, sizeof a);
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double th = 0.5 * (a[q][q] - a[p][p]) / a[p][q];
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double ev[3];
    for (int k = 0; k < 3; ++k) {
        if (!(a[k][k] > 0.0)) return false;
        ev[k] = std::sqrt(a[k][k]);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) R[i][j] += v[i][k] * ev[k] * v[j][k];
        }
    return true;
}

}  // namespace

CfgDMD read_cfg_dmd(const char* path)
{
    std::ifstream in(path);
    MAPP_REQUIRE(in.good(), ERR_FILE, std::string("cannot open ") + path);
    const std::string file(path);
    // header (src/import_cfg.cpp:32-211)
    double basic_length = 1.0;
    double H0[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, eta[3][3] = {}, trns[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    int entry_count = 6, natms = -1;
    bool vel_xst = true, is_ext = false, header_done = false;
    std::string raw;
    std::vector<std::string> tk;
    while (!header_done && std::getline(in, raw)) {
        tk = tokens(raw);
        if (tk.empty()) continue;
        const std::string line = join(tk);
        const char* l = line.c_str();
        int i0, j0, n0;
        double v;
        if (std::sscanf(l, " Number of particles = %d ", &n0) == 1) {
            MAPP_REQUIRE(n0 >= 0, ERR_FILE, "Number of particles in " + file + " should be greater than 0");
            natms = n0;
            continue;
        }
        if (std::sscanf(l, " A = %lf ", &v) == 1) {
            MAPP_REQUIRE(v > 0.0, ERR_FILE, "A in " + file + " should be greater than 0.0");
            basic_length = v;
            continue;
        }
        if (line == ".NO_VELOCITY.") {
            if (!is_ext) {
                vel_xst = false;
                is_ext = true;
            }
            continue;
        }
        if (std::sscanf(l, " entry_count = %d ", &n0) == 1) {
            const int mincomp = 3 + (vel_xst ? 3 : 0);
            MAPP_REQUIRE(n0 >= mincomp, ERR_FILE, "entry_count in " + file + " is below the position entries");
            entry_count = n0;
            is_ext = true;
            continue;
        }
        char sbuf[8];
        if (std::sscanf(l, " R = %lf %7s", &v, sbuf) >= 1 && l[0] == 'R') continue;
        if (std::sscanf(l, " auxiliary [ %d ] =", &i0) == 1) {
            MAPP_REQUIRE(is_ext, ERR_FILE, "auxiliary[...] must come after entry_count in " + file);
            const int mincomp = 3 + (vel_xst ? 3 : 0);
            MAPP_REQUIRE(i0 + mincomp + 1 <= entry_count, ERR_FILE, "auxiliary index beyond entry_count in " + file);
            continue;
        }
        if (std::sscanf(l, " Transform ( %d , %d ) = %lf ", &i0, &j0, &v) == 3) {
            MAPP_REQUIRE(i0 >= 1 && i0 <= 3 && j0 >= 1 && j0 <= 3, ERR_FILE, "wrong Transform component in " + file);
            trns[i0 - 1][j0 - 1] = v;
            continue;
        }
        if (std::sscanf(l, " H0 ( %d , %d ) = %lf ", &i0, &j0, &v) == 3) {
            MAPP_REQUIRE(i0 >= 1 && i0 <= 3 && j0 >= 1 && j0 <= 3, ERR_FILE, "wrong H0 component in " + file);
            H0[i0 - 1][j0 - 1] = v;
            continue;
        }
        if (std::sscanf(l, " eta ( %d , %d ) = %lf ", &i0, &j0, &v) == 3) {
            MAPP_REQUIRE(i0 >= 1 && i0 <= 3 && j0 >= 1 && j0 <= 3, ERR_FILE, "wrong eta component in " + file);
            eta[i0 - 1][j0 - 1] = v;
            continue;
        }
        if ((tk.size() == 8 && !is_ext) || (tk.size() == 1 && is_ext)) {
            header_done = true;  // this line (the first mass) starts the body
            continue;
        }
        MAPP_REQUIRE(false, ERR_FILE, "invalid line in " + file + ": " + line);
    }
    MAPP_REQUIRE(natms >= 0, ERR_FILE, "no 'Number of particles' in " + file);
    MAPP_REQUIRE(header_done || natms == 0, ERR_FILE, "file " + file + " ended unexpectedly");
    // box: H0 Transform, then sqrt(1 + 2 eta), then lower-triangular form
    double Hx[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Hx[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) Hx[i][j] += H0[i][k] * trns[k][j];
        }
    bool has_eta = false;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) has_eta = has_eta || eta[i][j] != 0.0;
    if (has_eta) {
        double e[3][3], es[3][3], t[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) e[i][j] = 2.0 * eta[i][j] + (i == j ? 1.0 : 0.0);
        MAPP_REQUIRE(sym_sqrt3(e, es), ERR_FILE, "eta in " + file + " should be positive definite");
        std::memcpy(t, Hx, sizeof t);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                Hx[i][j] = 0.0;
                for (int k = 0; k < 3; ++k) Hx[i][j] += t[i][k] * es[k][j];
            }
    }
    // MSQ_2_MLT: the lower-triangular box with the same row Gram matrix
    double L[3][3] = {};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) {
            double g = Hx[i][0] * Hx[j][0] + Hx[i][1] * Hx[j][1] + Hx[i][2] * Hx[j][2];
            for (int k = 0; k < j; ++k) g -= L[j][k] * L[i][k];
            if (i == j) {
                MAPP_REQUIRE(g > 0.0, ERR_FILE, "degenerate box in " + file);
                L[i][i] = std::sqrt(g);
            } else {
                L[i][j] = g / L[j][j];
            }
        }
    CfgDMD out;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out.H[i * 3 + j] = L[i][j] * basic_length;
    out.natms = natms;
    if (natms == 0) return out;
    MAPP_REQUIRE(is_ext, ERR_FILE, "cannot read cfg standard file for dmd mode");
    // body (src/import_cfg.cpp:213-356): mass, symbol, then rows of
    // entry_count values; s wrapped into [0, 1)
    const int naux = entry_count - 3;
    std::vector<double> s, aux;
    s.reserve((size_t)3 * natms);
    aux.reserve((size_t)naux * natms);
    bool mass_flag = false, elem_init = false;
    double mass = 0.0;
    bool first = true;
    while (true) {
        if (!first) {
            if (!std::getline(in, raw)) break;
            tk = tokens(raw);
        }
        first = false;
        if (tk.empty()) continue;
        MAPP_REQUIRE(tk.size() == 1 || (int)tk.size() == entry_count, ERR_FILE,
                     "invalid line in " + file + ": " + join(tk));
        MAPP_REQUIRE(!(mass_flag && tk.size() != 1), ERR_FILE,
                     "expected chemical symbol of element after mass in " + file);
        if (tk.size() == 1) {
            if (!mass_flag) {
                mass = std::atof(tk[0].c_str());
                MAPP_REQUIRE(mass > 0.0, ERR_FILE, "mass in " + file + " should be greater than 0.0");
                mass_flag = true;
            } else {
                bool found = false;
                for (const std::string& nm : out.names) found = found || nm == tk[0];
                if (!found) {
                    out.names.push_back(tk[0]);
                    out.masses.push_back(mass);
                }
                mass_flag = false;
                elem_init = true;
            }
            continue;
        }
        MAPP_REQUIRE(elem_init, ERR_FILE, "a site line comes before any element in " + file);
        for (int d = 0; d < 3; ++d) {
            double v = std::atof(tk[d].c_str());
            while (v >= 1.0) v--;
            while (v < 0.0) v++;
            s.push_back(v);
        }
        for (int k = 3; k < entry_count; ++k) aux.push_back(std::atof(tk[k].c_str()));
    }
    const int nread = (int)(s.size() / 3);
    MAPP_REQUIRE(nread == natms, ERR_FILE,
                 "expected " + std::to_string(natms) + " sites in " + file + " but read " + std::to_string(nread));
    // ImportCFGDMD::operator(): alpha and c of every element, checks, and
    // the present elements compacted into components
    const int ne = (int)out.names.size();
    MAPP_REQUIRE(ne <= kMaxElems, ERR_CAPACITY, "too many elements in " + file);
    MAPP_REQUIRE(naux >= 2 * ne, ERR_FILE, "values for c and alpha of all elements should be provided for every atom");
    int c_dim = 0;
    for (int i = 0; i < natms; ++i) {
        double* al = aux.data() + (size_t)i * naux;
        const double* c = al + ne;
        int cnt = 0;
        double csum = 0.0;
        for (int j = 0; j < ne; ++j) {
            MAPP_REQUIRE(!(c[j] > 1.0 || (c[j] < 0.0 && c[j] != -1.0)), ERR_FILE,
                         "all values of c should be greater or equal to 0.0 and less than or equal to 1.0, "
                         "or equal to -1.0");
            MAPP_REQUIRE(!(c[j] != -1.0 && al[j] < 0.0), ERR_FILE,
                         "all values of alpha corresponding to present elements (i.e. c!=-1.0) should be greater "
                         "than 0.0");
            if (c[j] == -1.0 && al[j] < 0.0) al[j] = 0.0;
            if (c[j] != -1.0) {
                csum += c[j];
                ++cnt;
            }
        }
        MAPP_REQUIRE(!(csum > 1.0), ERR_FILE, "sum of present elements of one site should not exceed 1.0");
        c_dim = std::max(c_dim, cnt);
    }
    MAPP_REQUIRE(c_dim >= 1 && c_dim <= 4, ERR_CAPACITY, "components per site must be 1..4");
    out.nelems = ne;
    out.c_dim = c_dim;
    out.x.resize((size_t)3 * natms);
    out.elem.assign((size_t)natms * c_dim, 0);
    out.alpha.assign((size_t)natms * c_dim, 0.0);
    out.c.assign((size_t)natms * c_dim, -1.0);
    for (int i = 0; i < natms; ++i) {
        const double* al = aux.data() + (size_t)i * naux;
        const double* c = al + ne;
        int k = 0;
        for (int j = 0; j < ne; ++j)
            if (c[j] != -1.0) {
                out.elem[(size_t)i * c_dim + k] = (uint8_t)j;
                out.c[(size_t)i * c_dim + k] = c[j];
                out.alpha[(size_t)i * c_dim + k] = al[j];
                ++k;
            }
        // s2x with the lower-triangular box
        const double* si = s.data() + (size_t)3 * i;
        for (int d = 0; d < 3; ++d)
            out.x[(size_t)3 * i + d] = si[0] * out.H[0 + d] + si[1] * out.H[3 + d] + si[2] * out.H[6 + d];
    }
    return out;
}

void write_cfg_dmd(const char* path, int natms, int c_dim, const double* H, const double* x, const uint8_t* elem,
                   const double* alpha, const double* c, int nelems, const double* masses,
                   const std::vector<std::string>& names)
{
    MAPP_REQUIRE(nelems >= 1 && (int)names.size() == nelems, ERR_ARG, "one name and mass per element");
    FILE* fp = std::fopen(path, "w");
    MAPP_REQUIRE(fp != nullptr, ERR_FILE, std::string("cannot write ") + path);
    // write_header (src/export_cfg_dmd.cpp:25-58)
    std::fprintf(fp, "Number of particles = %d\n", natms);
    std::fprintf(fp, "A = %.16lf Angstrom (basic length-scale)\n", 1.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) std::fprintf(fp, "H0(%d,%d) = %.16lf A\n", i + 1, j + 1, H[i * 3 + j]);
    std::fprintf(fp, ".NO_VELOCITY.\n");
    std::fprintf(fp, "entry_count = %d\n", 3 + 2 * nelems);
    int icmp = 0;
    for (int i = 0; i < nelems; ++i) std::fprintf(fp, "auxiliary[%d] = %s_%d [reduced unit]\n", icmp++, "alpha", i);
    for (int i = 0; i < nelems; ++i) std::fprintf(fp, "auxiliary[%d] = %s_%d [reduced unit]\n", icmp++, "c", i);
    // write_body (src/export_cfg_dmd.cpp:127-170): fractional positions, one
    // alpha (absent 0) and one c (absent -1) per element; a new element block
    // whenever the element of largest c changes
    double B[9];  // inverse of the lower-triangular box
    B[0] = 1.0 / H[0];
    B[1] = B[2] = B[5] = 0.0;
    B[4] = 1.0 / H[4];
    B[8] = 1.0 / H[8];
    B[3] = -H[3] * B[0] * B[4];
    B[7] = -H[7] * B[4] * B[8];
    B[6] = -(H[6] * B[0] + H[7] * B[3]) * B[8];
    std::vector<double> ad(nelems), cdv(nelems);
    auto dump = [&](int i) {
        for (int j = 0; j < nelems; ++j) {
            ad[j] = 0.0;
            cdv[j] = -1.0;
        }
        for (int k = 0; k < c_dim; ++k) {
            const int e = elem[(size_t)i * c_dim + k];
            ad[e] = alpha[(size_t)i * c_dim + k];
            cdv[e] = c[(size_t)i * c_dim + k];
        }
    };
    auto max_pos = [&]() {
        int ie = 0;
        double m = cdv[0];
        for (int j = 1; j < nelems; ++j)
            if (m < cdv[j]) {
                m = cdv[j];
                ie = j;
            }
        return ie;
    };
    int first_elem = -1;
    if (natms) {
        dump(0);
        first_elem = max_pos();
    }
    for (int i = 0; i < nelems - 1; ++i) std::fprintf(fp, "%lf\n%s\n", masses[i], names[i].c_str());
    if (nelems - 1 != first_elem) std::fprintf(fp, "%lf\n%s\n", masses[nelems - 1], names[nelems - 1].c_str());
    int curr = -1;
    for (int i = 0; i < natms; ++i) {
        dump(i);
        const int e = max_pos();
        if (e != curr) {
            curr = e;
            std::fprintf(fp, "%lf\n", masses[curr]);
            std::fprintf(fp, "%s\n", names[curr].c_str());
        }
        const double* xi = x + (size_t)3 * i;
        // x2s with the lower-triangular inverse (s = x B, B lower triangular)
        const double s0 = xi[0] * B[0] + xi[1] * B[3] + xi[2] * B[6];
        const double s1 = xi[1] * B[4] + xi[2] * B[7];
        const double s2 = xi[2] * B[8];
        std::fprintf(fp, "%.16lf %.16lf %.16lf ", s0, s1, s2);
        for (int j = 0; j < nelems; ++j) std::fprintf(fp, "%.16lf ", ad[j]);
        for (int j = 0; j < nelems; ++j) std::fprintf(fp, "%.16lf ", cdv[j]);
        std::fprintf(fp, "\n");
    }
    const bool ok = std::fclose(fp) == 0;
    MAPP_REQUIRE(ok, ERR_FILE, std::string("cannot write ") + path);
}

}  // namespace mapp
