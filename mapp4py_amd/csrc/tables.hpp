// tables.hpp -- tabulated EAM potentials: DYNAMO readers and spline tables.
#pragma once
#include <string>
#include <vector>

namespace mapp {

// Knot arrays turned into per-interval polynomial coefficients (value form:
// y(p) = c0 + c1 p + c2 p^2 + c3 p^3 [+ c4 p^4], p in [0,1] inside a knot
// interval). Layout follows the reference's type0 (*)[4] / (*)[5] tables.
struct EamTables {
    int ne = 0, nr = 0, nrho = 0;
    double dr = 0.0, drho = 0.0;
    bool dmd = false;            // quartic F and r*rho tables (DMD force field)
    std::vector<std::string> names;
    std::vector<double> rphi;    // [ne][ne][nr][4], r*phi, symmetric
    std::vector<double> rho;     // [ne][ne][nr][4]; rho[a][b]: block of element a, entry b
    std::vector<double> F;       // [ne][nrho][nf]
    std::vector<double> rc;      // [ne][ne]
    int nf() const { return dmd ? 5 : 4; }
    const double* rphi_at(int a, int b) const { return rphi.data() + ((size_t)a * ne + b) * nr * 4; }
    const double* rho_at(int a, int b) const { return rho.data() + ((size_t)a * ne + b) * nr * 4; }
    const double* F_at(int a) const { return F.data() + (size_t)a * nrho * nf(); }
    double max_rc() const {
        double m = 0.0;
        for (double v : rc) m = v > m ? v : m;
        return m;
    }
};

// Fills coefficient columns 1..stride-1 of an [n][stride] array whose column 0
// holds the knot values (stride 4: cubic Hermite with 5-point slopes; stride 5
// adds the quartic term used by the DMD embedding function).
void spline_coefficients(int n, double* s, int stride);

// The (value, slope per unit offset) pairs of the knots of one spline whose
// coefficient rows are c[4m .. 4m+3], into ys[2m], ys[2m+1], when every
// interval is the Hermite cubic of its two knots (c2 = 3 dy - 2 s0 - s1,
// c3 = s0 + s1 - 2 dy, as spline_coefficients makes them; checked to 1e-12
// of the knots' scale); false otherwise.
bool hermite_knots(const double* c, int nr, std::vector<double>& ys);

// fmt: 0 = setfl (eam/alloy), 1 = fs (eam/fs). elems selects and orders the
// file's elements (empty: all, in file order).
EamTables read_dynamo(const std::string& path, int fmt, const std::vector<std::string>& elems, bool dmd);

// ImportEAM::funcfl (src/import_eam.h:68-344): one DYNAMO funcfl file per
// element (F, Z, rho), resampled onto a common grid, r phi = 27.2*0.529 Z_i Z_j.
EamTables read_funcfl(const std::vector<std::string>& paths, bool dmd);

// Knots given directly: F[ne][nrho], rho[ne][ne][nr], r_phi[ne][ne][nr].
// setfl_alias makes every rho[a][b] equal rho[a][0] as the setfl format does.
EamTables tables_from_knots(int ne, int nr, int nrho, double dr, double drho, const double* F,
                            const double* rho, const double* r_phi, bool setfl_alias, bool dmd);

}  // namespace mapp
