// cfg.cpp -- CFG files of MD and DMD systems (host code).
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

// what ImportCFG::read_header leaves (src/import_cfg.cpp:32-211): the box
// (lower triangular, times A), the particle count, the body layout and the
// tokens of the body's first line
struct Header {
    int natms = -1, entry_count = 6;
    bool vel_xst = true, is_ext = false, done = false;
    double R = 1.0;
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<std::string> first;
};

Header read_header(std::istream& in, const std::string& file)
{
    Header h;
    double basic_length = 1.0;
    double H0[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, eta[3][3] = {}, trns[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::string raw;
    std::vector<std::string> tk;
    while (!h.done && std::getline(in, raw)) {
        tk = tokens(raw);
        if (tk.empty()) continue;
        const std::string line = join(tk);
        const char* l = line.c_str();
        int i0, j0, n0;
        double v;
        if (std::sscanf(l, " Number of particles = %d ", &n0) == 1) {
            MAPP_REQUIRE(n0 >= 0, ERR_FILE, "Number of particles in " + file + " should be greater than 0");
            h.natms = n0;
            continue;
        }
        if (std::sscanf(l, " A = %lf ", &v) == 1) {
            MAPP_REQUIRE(v > 0.0, ERR_FILE, "A in " + file + " should be greater than 0.0");
            basic_length = v;
            continue;
        }
        if (line == ".NO_VELOCITY.") {
            if (!h.is_ext) {
                h.vel_xst = false;
                h.is_ext = true;
            }
            continue;
        }
        if (std::sscanf(l, " entry_count = %d ", &n0) == 1) {
            const int mincomp = 3 + (h.vel_xst ? 3 : 0);
            MAPP_REQUIRE(n0 >= mincomp, ERR_FILE, "entry_count in " + file + " is below the position entries");
            h.entry_count = n0;
            h.is_ext = true;
            continue;
        }
        if (l[0] == 'R' && std::sscanf(l, " R = %lf ", &v) == 1) {
            h.R = v;
            continue;
        }
        if (std::sscanf(l, " auxiliary [ %d ] =", &i0) == 1) {
            MAPP_REQUIRE(h.is_ext, ERR_FILE, "auxiliary[...] must come after entry_count in " + file);
            const int mincomp = 3 + (h.vel_xst ? 3 : 0);
            MAPP_REQUIRE(i0 + mincomp + 1 <= h.entry_count, ERR_FILE, "auxiliary index beyond entry_count in " + file);
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
        if ((tk.size() == 8 && !h.is_ext) || (tk.size() == 1 && h.is_ext)) {
            h.done = true;  // this line (the first mass, or the first atom) starts the body
            h.first = tk;
            continue;
        }
        MAPP_REQUIRE(false, ERR_FILE, "invalid line in " + file + ": " + line);
    }
    MAPP_REQUIRE(h.natms >= 0, ERR_FILE, "no 'Number of particles' in " + file);
    MAPP_REQUIRE(h.done || h.natms == 0, ERR_FILE, "file " + file + " ended unexpectedly");
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
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) h.H[i * 3 + j] = L[i][j] * basic_length;
    return h;
}

// the atom section: fractional positions wrapped into [0, 1), the element of
// every atom (Elements::add_type numbering: first appearance, the first mass
// given for a name kept), and the remaining columns
struct Body {
    std::vector<std::string> names;
    std::vector<double> masses;
    std::vector<uint8_t> elem;
    std::vector<double> s, aux;
    int naux = 0;
};

int add_type(Body& b, double mass, const std::string& name, const std::string& file)
{
    for (size_t i = 0; i < b.names.size(); ++i)
        if (b.names[i] == name) return (int)i;
    MAPP_REQUIRE(b.names.size() < 256, ERR_FILE, "cannot have more than 256 elements in " + file);
    b.names.push_back(name);
    b.masses.push_back(mass);
    return (int)b.names.size() - 1;
}

void push_s(Body& b, const std::vector<std::string>& tk, int at)
{
    for (int d = 0; d < 3; ++d) {
        double v = std::atof(tk[at + d].c_str());
        while (v >= 1.0) v--;
        while (v < 0.0) v++;
        b.s.push_back(v);
    }
}

// ImportCFG::read_body_ext (src/import_cfg.cpp:213-330): a mass line, a
// symbol line, then rows of entry_count values
Body read_body_ext(std::istream& in, const Header& h, const std::string& file)
{
    Body b;
    b.naux = h.entry_count - 3;
    std::string raw;
    std::vector<std::string> tk = h.first;
    bool mass_flag = false, elem_init = false, first = true;
    double mass = 0.0;
    int ielem = 0;
    while (true) {
        if (!first) {
            if (!std::getline(in, raw)) break;
            tk = tokens(raw);
        }
        first = false;
        if (tk.empty()) continue;
        MAPP_REQUIRE(tk.size() == 1 || (int)tk.size() == h.entry_count, ERR_FILE,
                     "invalid line in " + file + ": " + join(tk));
        MAPP_REQUIRE(!(mass_flag && tk.size() != 1), ERR_FILE,
                     "expected chemical symbol of element after mass in " + file);
        if (tk.size() == 1) {
            if (!mass_flag) {
                mass = std::atof(tk[0].c_str());
                MAPP_REQUIRE(mass > 0.0, ERR_FILE, "mass in " + file + " should be greater than 0.0");
                mass_flag = true;
            } else {
                ielem = add_type(b, mass, tk[0], file);
                mass_flag = false;
                elem_init = true;
            }
            continue;
        }
        MAPP_REQUIRE(elem_init, ERR_FILE, "a site line comes before any element in " + file);
        push_s(b, tk, 0);
        b.elem.push_back((uint8_t)ielem);
        for (int k = 3; k < h.entry_count; ++k) b.aux.push_back(std::atof(tk[k].c_str()));
    }
    return b;
}

// ImportCFG::read_body_std (src/import_cfg.cpp:334-480): every row is mass,
// symbol, three fractional coordinates and three reduced velocities
Body read_body_std(std::istream& in, const Header& h, const std::string& file)
{
    Body b;
    b.naux = 3;
    std::string raw;
    std::vector<std::string> tk = h.first;
    bool first = true;
    while (true) {
        if (!first) {
            if (!std::getline(in, raw)) break;
            tk = tokens(raw);
        }
        first = false;
        if (tk.empty()) continue;
        MAPP_REQUIRE(tk.size() == 8, ERR_FILE, "invalid line in " + file + ": " + join(tk));
        const double mass = std::atof(tk[0].c_str());
        MAPP_REQUIRE(mass > 0.0, ERR_FILE, "mass in " + file + " should be greater than 0.0");
        const int ielem = add_type(b, mass, tk[1], file);
        push_s(b, tk, 2);
        b.elem.push_back((uint8_t)ielem);
        for (int k = 5; k < 8; ++k) b.aux.push_back(std::atof(tk[k].c_str()));
    }
    return b;
}

void check_count(const Body& b, int natms, const std::string& file)
{
    const int nread = (int)(b.s.size() / 3);
    MAPP_REQUIRE(nread == natms, ERR_FILE,
                 "expected " + std::to_string(natms) + " atoms in " + file + " but read " + std::to_string(nread));
}

// inverse of a lower-triangular box (s = x B, B lower triangular)
void lower_inverse(const double* H, double* B)
{
    B[0] = 1.0 / H[0];
    B[1] = B[2] = B[5] = 0.0;
    B[4] = 1.0 / H[4];
    B[8] = 1.0 / H[8];
    B[3] = -H[3] * B[0] * B[4];
    B[7] = -H[7] * B[4] * B[8];
    B[6] = -(H[6] * B[0] + H[7] * B[3]) * B[8];
}

// x B of a lower-triangular B; wrapped into [0, 1) as Algebra::X2S does
void x2s(const double* x, const double* B, bool wrap, double* s)
{
    s[0] = x[0] * B[0] + x[1] * B[3] + x[2] * B[6];
    s[1] = x[1] * B[4] + x[2] * B[7];
    s[2] = x[2] * B[8];
    if (wrap)
        for (int d = 0; d < 3; ++d) {
            while (s[d] < 0.0) ++s[d];
            while (s[d] >= 1.0) --s[d];
        }
}

}  // namespace

CfgMD read_cfg_md(const char* path)
{
    std::ifstream in(path);
    MAPP_REQUIRE(in.good(), ERR_FILE, std::string("cannot open ") + path);
    const std::string file(path);
    const Header h = read_header(in, file);
    CfgMD out;
    std::memcpy(out.H, h.H, sizeof out.H);
    out.natms = h.natms;
    if (h.natms == 0) return out;
    const Body b = h.is_ext ? read_body_ext(in, h, file) : read_body_std(in, h, file);
    check_count(b, h.natms, file);
    out.nelems = (int)b.names.size();
    out.names = b.names;
    out.masses = b.masses;
    out.elem = b.elem;
    out.has_vel = h.vel_xst;
    out.x.resize((size_t)3 * h.natms);
    if (out.has_vel) out.v.resize((size_t)3 * h.natms);
    // s2x and x_d = s_d (R H), both with the lower-triangular box
    for (int i = 0; i < h.natms; ++i) {
        const double* si = b.s.data() + (size_t)3 * i;
        for (int d = 0; d < 3; ++d)
            out.x[(size_t)3 * i + d] = si[0] * h.H[0 + d] + si[1] * h.H[3 + d] + si[2] * h.H[6 + d];
        if (out.has_vel) {
            const double* vi = b.aux.data() + (size_t)b.naux * i;
            for (int d = 0; d < 3; ++d)
                out.v[(size_t)3 * i + d] =
                    h.R * (vi[0] * h.H[0 + d] + vi[1] * h.H[3 + d] + vi[2] * h.H[6 + d]);
        }
    }
    return out;
}

void write_cfg_md(const char* path, int natms, const double* H, const double* x, const uint8_t* elem,
                  const double* v, int naux, const double* aux, const std::vector<std::string>& aux_names,
                  int nelems, const double* masses, const std::vector<std::string>& names)
{
    MAPP_REQUIRE(nelems >= 1 && (int)names.size() == nelems, ERR_ARG, "one name and mass per element");
    MAPP_REQUIRE(naux >= 0 && (int)aux_names.size() == naux, ERR_ARG, "one name per auxiliary column");
    for (int i = 0; i < natms; ++i)
        MAPP_REQUIRE(elem[i] < nelems, ERR_ARG, "element index out of range");
    FILE* fp = std::fopen(path, "w");
    MAPP_REQUIRE(fp != nullptr, ERR_FILE, std::string("cannot write ") + path);
    // write_header (src/export_cfg_md.cpp:26-53)
    std::fprintf(fp, "Number of particles = %d\n", natms);
    std::fprintf(fp, "A = %.16lf Angstrom (basic length-scale)\n", 1.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) std::fprintf(fp, "H0(%d,%d) = %.16lf A\n", i + 1, j + 1, H[i * 3 + j]);
    if (v)
        std::fprintf(fp, "R = 1.0\n");
    else
        std::fprintf(fp, ".NO_VELOCITY.\n");
    std::fprintf(fp, "entry_count = %d\n", 3 + (v ? 3 : 0) + naux);
    for (int k = 0; k < naux; ++k) std::fprintf(fp, "auxiliary[%d] = %s [reduced unit]\n", k, aux_names[k].c_str());
    // write_body_sort (src/export_cfg_md.cpp:57-99): a mass and symbol line
    // whenever the element changes, then s, s_d = v B, the auxiliary columns
    double B[9];
    lower_inverse(H, B);
    int curr = -1;
    for (int i = 0; i < natms; ++i) {
        if (elem[i] != curr) {
            curr = elem[i];
            std::fprintf(fp, "%lf\n", masses[curr]);
            std::fprintf(fp, "%s\n", names[curr].c_str());
        }
        double s[3];
        x2s(x + (size_t)3 * i, B, true, s);
        std::fprintf(fp, "%.16lf %.16lf %.16lf ", s[0], s[1], s[2]);
        if (v) {
            x2s(v + (size_t)3 * i, B, false, s);
            std::fprintf(fp, "%.16lf %.16lf %.16lf ", s[0], s[1], s[2]);
        }
        for (int k = 0; k < naux; ++k) std::fprintf(fp, "%.16lf ", aux[(size_t)naux * i + k]);
        std::fprintf(fp, "\n");
    }
    const bool ok = std::fclose(fp) == 0;
    MAPP_REQUIRE(ok, ERR_FILE, std::string("cannot write ") + path);
}

CfgDMD read_cfg_dmd(const char* path)
{
    std::ifstream in(path);
    MAPP_REQUIRE(in.good(), ERR_FILE, std::string("cannot open ") + path);
    const std::string file(path);
    const Header h = read_header(in, file);
    CfgDMD out;
    std::memcpy(out.H, h.H, sizeof out.H);
    out.natms = h.natms;
    if (h.natms == 0) return out;
    MAPP_REQUIRE(h.is_ext, ERR_FILE, "cannot read cfg standard file for dmd mode");
    const Body b = read_body_ext(in, h, file);
    check_count(b, h.natms, file);
    const int natms = h.natms, naux = b.naux;
    std::vector<double> aux = b.aux;
    out.names = b.names;
    out.masses = b.masses;
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
        const double* si = b.s.data() + (size_t)3 * i;
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
    double B[9];
    lower_inverse(H, B);
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
        double si[3];
        x2s(x + (size_t)3 * i, B, true, si);
        std::fprintf(fp, "%.16lf %.16lf %.16lf ", si[0], si[1], si[2]);
        for (int j = 0; j < nelems; ++j) std::fprintf(fp, "%.16lf ", ad[j]);
        for (int j = 0; j < nelems; ++j) std::fprintf(fp, "%.16lf ", cdv[j]);
        std::fprintf(fp, "\n");
    }
    const bool ok = std::fclose(fp) == 0;
    MAPP_REQUIRE(ok, ERR_FILE, std::string("cannot write ") + path);
}

}  // namespace mapp
