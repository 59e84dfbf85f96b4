// cfg.hpp -- extended CFG (AtomEye) files of DMD systems: sites with a
// fractional position, one alpha and one c per element.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mapp {

// A DMD configuration as the importer leaves it: the present elements of
// each site compacted into c_dim components (absent: element 0, c = -1,
// alpha = 0), positions Cartesian, H lower triangular with rows = box vectors.
struct CfgDMD {
    int natms = 0, nelems = 0, c_dim = 0;
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<double> x;      // [natms][3]
    std::vector<uint8_t> elem;  // [natms][c_dim]
    std::vector<double> alpha;  // [natms][c_dim]
    std::vector<double> c;      // [natms][c_dim]
    std::vector<double> masses;
    std::vector<std::string> names;
};

// ImportCFGDMD::operator() (src/import_cfg.cpp:669-854) over
// ImportCFG::read_header / read_body_ext (src/import_cfg.cpp:32-330)
CfgDMD read_cfg_dmd(const char* path);

// ExportCFGDMD::write_header / write_body (src/export_cfg_dmd.cpp:25-170):
// sites in the given order, components expanded back to one alpha and one c
// per element
void write_cfg_dmd(const char* path, int natms, int c_dim, const double* H, const double* x, const uint8_t* elem,
                   const double* alpha, const double* c, int nelems, const double* masses,
                   const std::vector<std::string>& names);

}  // namespace mapp
